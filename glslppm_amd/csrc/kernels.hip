// Kernels of the SPPM pass loop for gfx950.  One pass (m_display, glslppm.cpp:214-536) is:
//   k_pass_begin    frame count, eye cadence, NumPhotons of the last pass (one block)
//   k_eye_begin, k_eye, k_grid  every max_path passes: AAOffset, eyeraytrace.fs + per-block
//                   max.fs/min.fs partials, the hit-point box and the grid uniforms of
//                   glslppm.cpp:292-311 (one block)
//   k_photon_stage  photontrace.fs emission/bounce, walk and deposit + hash.fs + scatter (atomic
//                   min of the photon index) + correction (atomic count), one thread per slot;
//                   sum.fs as per-wave counts; the clear of the next pass's hash table; then, in
//                   blocks of its own dispatched after the photon blocks, the gather of the pass
//                   before: progressive.fs (8-cell gather and the N/R/tau update) + re.fs draw
//   k_gather   the same gather for the last pass of a sequence (launched once at its end)
// Arithmetic follows the GLSL statement by statement in strict fp32 (see fmath.h).
//
// Each kernel is written once as a body over (job, block of that job).  The plain form takes the
// job by value and runs that job's blocks; the _batch form takes the job array in HBM and a block
// map, so that one launch advances every job of a batch.  Every body starts by reading its job's
// pass flags and returns when the job does not run this stage, so the launch sequence of a pass
// is fixed and a hipGraph of it stays valid from pass to pass.

#include "kernels.h"

namespace {

constexpr float kInf = __builtin_huge_valf();

// Streaming accesses: the per-slot photon records, read once and written once per pass (600 MB a
// launch at 4M photons), and the hash clear.  As nontemporal loads and stores their lines are the
// first to leave L2, so they evict fewer of the walk's tree and triangle records (measured: photon
// stage 1.370 -> 1.350 ms at 4M photons, 0.446 -> 0.436 ms at 1M).
typedef float v4f_stream __attribute__((ext_vector_type(4)));
typedef int v2i_stream __attribute__((ext_vector_type(2)));
PPM_DEV float4 ld_stream(const float4* p) {
    const v4f_stream v = __builtin_nontemporal_load(reinterpret_cast<const v4f_stream*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
PPM_DEV int2 ld_stream(const int2* p) {
    const v2i_stream v = __builtin_nontemporal_load(reinterpret_cast<const v2i_stream*>(p));
    return make_int2(v.x, v.y);
}
PPM_DEV void st_stream(float4* p, float4 x) {
    v4f_stream v;
    v.x = x.x;
    v.y = x.y;
    v.z = x.z;
    v.w = x.w;
    __builtin_nontemporal_store(v, reinterpret_cast<v4f_stream*>(p));
}
PPM_DEV void st_stream(int2* p, int2 x) {
    v2i_stream v;
    v.x = x.x;
    v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<v2i_stream*>(p));
}
PPM_DEV void st_stream(int* p, int x) { __builtin_nontemporal_store(x, p); }

PPM_DEV int cube_face(f3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return d.x >= 0.0f ? 0 : 1;
    if (ay >= az) return d.y >= 0.0f ? 2 : 3;
    return d.z >= 0.0f ? 4 : 5;
}

struct Hit {
    int tri;    // -1: no hit
    float t;
    float bx, by, bz;  // Barycentric = uvt.ywx
    f3 pos;            // Result.pos = V2 B.x + V0 B.y + V1 B.z (eyeraytrace.fs:273)
};

// The closest hit while a walk runs; the barycentrics of the final one are computed at the end.
struct Closest {
    int tri;
    float t;
};

// The Moller-Trumbore terms of a triangle record (vertices v0, v1, v2 in q0, q1, q2) for a ray
// (eyeraytrace.fs:118-147, photontrace.fs:174-205), in the shaders' operation order, the edges
// included (eyeraytrace.fs:125-126).
PPM_DEV void tri_terms(float4 q0, float4 q1, float4 q2, f3 org, f3 dir, float& det, float& u, float& v, float& t) {
    const f3 e0 = xyz(q1) - xyz(q0), e1 = xyz(q2) - xyz(q0);
    const f3 pv = cross3(dir, e1);
    det = dot3(e0, pv);
    const f3 tv = org - xyz(q0);
    const f3 qv = cross3(tv, e0);
    u = dot3(tv, pv) / det;
    v = dot3(dir, qv) / det;
    t = dot3(e1, qv) / det;
}

// Triangle test of a leaf.  Eye rays cull back faces on their first segment; photon rays skip
// light triangles.
template <bool PHOTON>
PPM_DEV void tri_test(float4 q0, float4 q1, float4 q2, f3 org, f3 dir, bool cull, int tri, Closest& h) {
    float det, u, v, t;
    tri_terms(q0, q1, q2, org, dir, det, u, v, t);
    if (!cull || det > 1e-10f) {
        const float w = (1.0f - u) - v;
        bool ok = (u >= 0.0f) && (v >= 0.0f) && (t >= 0.0f) && (w >= 0.0f) && (t < h.t);
        if (PHOTON) ok = ok && (__float_as_int(q0.w) != -1);
        if (ok) {
            h.tri = tri;
            h.t = t;
        }
    }
}

// The barycentrics tri_test computed when it took the closest hit: the same terms, recomputed.
PPM_DEV void hit_barycentrics(const DevScene& s, f3 org, f3 dir, Hit& h) {
    if (h.tri < 0) return;
    const float4* T = s.tri_isect + 4 * (size_t)h.tri;
    const float4 q0 = T[0], q1 = T[1], q2 = T[2];
    float det, u, v, t;
    tri_terms(q0, q1, q2, org, dir, det, u, v, t);
    h.bx = v;
    h.by = (1.0f - u) - v;
    h.bz = u;
    h.pos = (xyz(q2) * h.bx + xyz(q0) * h.by) + xyz(q1) * h.bz;
}

// Leaves a lane may hold before its triangle tests run.
constexpr int kLeafQueue = 2;

// The exact box of a leaf: read from its triangle record (q1.w, q2.w, q3) or, when every leaf's
// box is its triangle's vertex box, the smallest and largest vertex coordinate per axis (exact, no
// rounding; min and max of the same three values are the reference's box up to the sign of a
// zero, which no slab comparison sees).
PPM_DEV void leaf_box(bool stored, float4 q0, float4 q1, float4 q2, float4 q3, f3& lo, f3& hi) {
    if (stored) {
        lo = mk3(q1.w, q2.w, q3.x);
        hi = mk3(q3.y, q3.z, q3.w);
    } else {
        lo = mk3(fminf(fminf(q0.x, q1.x), q2.x), fminf(fminf(q0.y, q1.y), q2.y), fminf(fminf(q0.z, q1.z), q2.z));
        hi = mk3(fmaxf(fmaxf(q0.x, q1.x), q2.x), fmaxf(fmaxf(q0.y, q1.y), q2.y), fmaxf(fmaxf(q0.z, q1.z), q2.z));
    }
}

// float(binary16 in the low / high half of w) * r with one rounding: v_fma_mix_f32 widens the
// half exactly and adds -0, which leaves every product (signed zeros included) unchanged, so the
// result is bit-identical to a conversion followed by v_mul_f32, in one instruction.
PPM_DEV float mul_half_lo(unsigned int w, float r) {
    float d;
    asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[1,0,0]" : "=v"(d) : "v"(w), "v"(r));
    return d;
}
PPM_DEV float mul_half_hi(unsigned int w, float r) {
    float d;
    asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(w), "v"(r));
    return d;
}

// float(binary16 in the low / high half of w) * r + t with one rounding (a fused multiply-add).
PPM_DEV float fma_half_lo(unsigned int w, float r, float t) {
    float d;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(w), "v"(r), "v"(t));
    return d;
}
PPM_DEV float fma_half_hi(unsigned int w, float r, float t) {
    float d;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(w), "v"(r), "v"(t));
    return d;
}

// The reference's slab offset of one axis, t0 = -rcp * org (eyeraytrace.fs:88,
// photontrace.fs:144).
PPM_DEV float slab_t0(float rcp, float org) { return -rcp * org; }

// The axes of a face's node records: the major axis m = face / 2, then the other two in xyz order.
PPM_DEV f3 face_axes(int face, f3 v) {
    const int m = face >> 1;
    return mk3(m == 0 ? v.x : m == 1 ? v.y : v.z, m == 0 ? v.y : v.x, m == 2 ? v.y : v.z);
}

// The record tests serve a ray whose reciprocals lie in [2^-100, 2^100] and whose offsets in
// [-2^120, 2^120] in magnitude (unit directions always do unless parallel to an axis).
PPM_DEV bool sign_order_ok(f3 r, f3 t) {
    const float am = fabsf(r.x), aa = fabsf(r.y), ab = fabsf(r.z);
    return am >= 0x1p-100f && am <= 0x1p100f && aa >= 0x1p-100f && aa <= 0x1p100f && ab >= 0x1p-100f &&
           ab <= 0x1p100f && fabsf(t.x) <= 0x1p120f && fabsf(t.y) <= 0x1p120f && fabsf(t.z) <= 0x1p120f;
}

// The slack of a record test's fused slab values on one axis (wide_test).
PPM_DEV float slab_slack(float r, float t, float slab_eps) {
    return __builtin_fmaf(fabsf(r), slab_eps, __builtin_fmaf(fabsf(t), 0x1p-22f, 0x1p-140f));
}

// A ray's reciprocals in the axis order of its face's node records; sh_a / sh_b rotate a record
// word so that its low half is the near bound on that axis (16 when the direction is negative
// there).
struct WideRay {
    float rm, ra, rb;
    unsigned int sh_a, sh_b;
};

PPM_DEV WideRay wide_ray(int face, f3 rcp) {
    const f3 r = face_axes(face, rcp);
    WideRay w;
    w.rm = r.x;
    w.ra = r.y;
    w.rb = r.z;
    w.sh_a = signbit(w.ra) ? 16u : 0u;
    w.sh_b = signbit(w.rb) ? 16u : 0u;
    return w;
}

// The offsets of the record tests for the near and far bound of each axis: t0 -+ e for a ray
// the four-wide record tests serve (sign_order_ok, see wide_test), exactly t0 for any other.
struct SlabOffsets {
    float nm, fm, na, fa, nb, fb;
};

PPM_DEV SlabOffsets slab_offsets(const WideRay& w, int face, f3 org, float slab_eps, bool fused) {
    const f3 o = face_axes(face, org);
    const float tm = slab_t0(w.rm, o.x), ta = slab_t0(w.ra, o.y), tb = slab_t0(w.rb, o.z);
    const float em = fused ? slab_slack(w.rm, tm, slab_eps) : 0.0f;
    const float ea = fused ? slab_slack(w.ra, ta, slab_eps) : 0.0f;
    const float eb = fused ? slab_slack(w.rb, tb, slab_eps) : 0.0f;
    return {tm - em, tm + em, ta - ea, ta + ea, tb - eb, tb + eb};
}

PPM_DEV bool ray_fused(const WideRay& w, int face, f3 org) {
    const f3 o = face_axes(face, org);
    return sign_order_ok(mk3(w.rm, w.ra, w.rb), mk3(slab_t0(w.rm, o.x), slab_t0(w.ra, o.y), slab_t0(w.rb, o.z)));
}

// The reference's slab test of an exact fp32 box (eyeraytrace.fs:92-116, photontrace.fs:148-172:
// per axis rcp * bound + t0, min/max, then tmin <= tmax, tmin <= closest hit, tmax >= 0) with the
// slab terms in the face's axis order: the same products, offsets, sums and per-axis min/max; the
// axes enter tmin and tmax in another order, which NaN-rejecting min and max do not see (up to the
// sign of a zero, which no comparison sees).
PPM_DEV bool box_test_face(const WideRay& w, int face, f3 org, f3 a, f3 b, float ht) {
    const f3 o = face_axes(face, org), lo = face_axes(face, a), hi = face_axes(face, b);
    const float tm = slab_t0(w.rm, o.x), ta = slab_t0(w.ra, o.y), tb = slab_t0(w.rb, o.z);
    const float lm = w.rm * lo.x + tm, hm = w.rm * hi.x + tm;
    const float la = w.ra * lo.y + ta, ha = w.ra * hi.y + ta;
    const float lb = w.rb * lo.z + tb, hb = w.rb * hi.z + tb;
    const float tmin = max1(max1(min1(lm, hm), min1(la, ha)), min1(lb, hb));
    const float tmax = min1(min1(max1(lm, hm), max1(la, ha)), max1(lb, hb));
    return (tmin <= tmax) && (tmin <= ht) && (tmax >= 0.0f);
}

// Slab test of a pre-order node record's outward-rounded binary16 box (axes in the face's order,
// each as lo | hi << 16 or hi | lo << 16, rotated so that the low half is the near bound when the
// direction is negative on a minor axis), in box_test_face's arithmetic with the near and far
// offsets: per-axis min/max whatever the signs.  For a ray the record tests do not serve both
// offsets are t0 (up to the sign of a zero t0 + 0, which no comparison sees) and this is the
// reference's test on the rounded bounds; for one they serve, the near offset is <= t0 <= the far
// one, which only moves each value outward.
PPM_DEV bool node_test(uint4 r, const WideRay& w, const SlabOffsets& o, float ht) {
    const unsigned int a = __builtin_rotateright32(r.y, w.sh_a), b = __builtin_rotateright32(r.z, w.sh_b);
    const float lm = mul_half_lo(r.x, w.rm) + o.nm, hm = mul_half_hi(r.x, w.rm) + o.fm;
    const float la = mul_half_lo(a, w.ra) + o.na, ha = mul_half_hi(a, w.ra) + o.fa;
    const float lb = mul_half_lo(b, w.rb) + o.nb, hb = mul_half_hi(b, w.rb) + o.fb;
    const float tmin = max1(max1(min1(lm, hm), min1(la, ha)), min1(lb, hb));
    const float tmax = min1(min1(max1(lm, hm), max1(la, ha)), max1(lb, hb));
    return (tmin <= tmax) && (tmin <= ht) && (tmax >= 0.0f);
}

// Slab test of one member slot of a four-wide record, for rays that sign_order_ok accepts.  The
// near and far bounds are taken by the direction's signs instead of by min/max, and each slab
// value is one fused multiply-add, fma(b', rcp, t0 -+ e) on the binary16 bound b', with the
// offset moved outward by the slack e = |rcp| slab_eps + 2^-22 |t0| + 2^-140 (fp32 fmas, to
// nearest).  No slab value is NaN (a bound of magnitude over 65504 may pack as an infinity, whose
// product is infinite and exact; every other product and sum stays finite) and the two values of
// an axis are ordered by the sign of its reciprocal.  The fused value is never inward of the
// reference's rcp * b + t0 (two roundings) on the exact bound b: b' is b rounded outward, so
// rcp b' is on the outer side of rcp b; rounding rcp b errs by at most 2^-24 |rcp| B + 2^-150,
// where B is the largest bound magnitude capped at 2^17 and slab_eps = 2^-22 B (a bound beyond
// 2^17 packs as an infinity or as +-65504, which moves rcp b' outward by more than half of
// |rcp b|, far more than that product's rounding error); rounding t0 -+ e errs by at most
// 2^-24 (|t0| + e) + 2^-150; and e, rounded twice, is at least
// 3.99 (2^-24 |rcp| B + 2^-24 |t0|) + 2^-141, more than all of these together.  Rounding to
// nearest is monotone, so tmin is no larger and tmax no smaller than the reference's test on the
// exact box.  (The order in which the axes enter tmin and tmax does not matter: NaN-rejecting min
// and max of a set do not depend on it, up to the sign of a zero, which no comparison here sees.)
// Other rays, those parallel to an axis above all, keep the reference's arithmetic and NaN
// behaviour on the pre-order records: there, taking bounds by sign would leave a NaN slab out where
// min/max rejects the box, and such a ray would enter almost every box.
PPM_DEV bool wide_test(uint4 r, const WideRay& w, const SlabOffsets& o, float ht) {
    const unsigned int a = __builtin_rotateright32(r.y, w.sh_a), b = __builtin_rotateright32(r.z, w.sh_b);
    const float nm = fma_half_lo(r.x, w.rm, o.nm), fm = fma_half_hi(r.x, w.rm, o.fm);
    const float na = fma_half_lo(a, w.ra, o.na), fa = fma_half_hi(a, w.ra, o.fa);
    const float nb = fma_half_lo(b, w.rb, o.nb), fb = fma_half_hi(b, w.rb, o.fb);
    const float tmin = max1(max1(nm, na), nb);
    const float tmax = min1(min1(fm, fa), fb);
    return (tmin <= tmax) && (tmin <= ht) && (tmax >= 0.0f);
}

// The walk's stack: each lane's entries are walk_lds[k * blockDim.x + lane], k < Job::walk_stack.
extern __shared__ int walk_lds[];

// Closest hit over the threaded BVH of one cube face (eyeraytrace.fs:67-254,
// photontrace.fs:123-218).  The reference walks the face's pre-order (hit: next position, miss:
// the stored link) and tests the triangle of each leaf whose box is hit.
//
// This walk visits the same tree four-wide: a record holds up to four members (descendants of
// one node, in pre-order), all four boxes are tested at once, the first member hit is visited
// next and the others are pushed on the lane's stack in LDS, to be popped in pre-order.  A leaf
// member is queued when it is reached.  A lane whose stack would overflow walks on over the
// pre-order records from the first member hit (the pre-order from there covers everything still
// on its stack).
//
// The walk runs in two alternating phases so that a wave's triangle tests (three correctly
// rounded divisions each) run with many lanes at once instead of whenever any lane meets a leaf.
// In the box phase a lane that reaches a leaf queues it and walks on; the wave leaves the box
// phase when no lane can step or when more than a quarter of the lanes that were still walking
// when the phase began hold full queues.  The leaf phase then takes each lane's queued leaves in
// the order they were met: each gets the reference's slab test of its exact box with the current
// closest hit, then its triangle test.
//
// Box-phase boxes are binary16 boxes rounded outward.  A slab value rcp * b + t0 is monotone in
// b, and rounding outward keeps the sign of each bound (which decides the NaN axes of zero
// direction components), so a rounded box passes whenever the exact one does, with a tmin no
// larger (the four-wide records' fused slab values with their slack too, see wide_test).  Boxes are tested against a closest-hit distance no smaller than the reference's at
// the same point of the pre-order (members are tested before their turn, and leaves queued
// earlier may still be untested).  So the lane may enter boxes the reference skips, never skips
// one the reference enters, and reaches the reference's leaves in the reference's order.  A leaf
// inside a box the reference skips either fails its exact test or has an exact tmin >= that
// box's tmin (nested boxes, monotone slab test, NaN axes included) > the closest hit at that
// point; every other leaf is tested exactly as the reference tests it, with the closest hit the
// reference has at that point.  The triangle tests, their order and so every result, ties
// included, are the reference's.  On the pre-order records nodes i and i + 1 are fetched
// together; each face array ends with a padding node, so the second fetch stays in bounds.
constexpr int kDone = (int)0x80000000;  // no record or leaf left (never ~triangle: T < 2^31 - 1)

template <bool PHOTON>
PPM_DEV void walk(const DevScene& s, int face, f3 org, f3 dir, f3 rcp_in, bool cull, Hit& hit, int depth) {
    Closest h = {hit.tri, hit.t};
    const int n = s.num_nodes;
    const uint4* __restrict__ nd = s.nodes + (size_t)face * (size_t)(n + 1);
    const uint4* __restrict__ wr = s.wide + (size_t)face * (size_t)s.wide_stride * 4;
    // the lane's stack: entry k at walk_lds[k * blockDim.x + lane]; top points past the last one
    int* top = walk_lds + threadIdx.x;
    const int stride = blockDim.x;
    const WideRay wr4 = wide_ray(face, rcp_in);
    const bool fused = ray_fused(wr4, face, org);
    int cur = 0;  // >= 0: the record to visit; else ~triangle of a leaf to queue, or kDone
    int sp = 0;   // stack entries
    // >= 0: the lane walks the pre-order records from position i (from the start for a ray the
    // record tests do not cover)
    int i = fused ? -1 : 0;
    // queued leaves, newest first: q0, then q1 when two are queued
    int q0 = -1, q1 = -1;
    int nq = 0;
    for (;;) {
        const int live = __popcll(__ballot(cur != kDone));  // every queue is empty here
        // the slab offsets, recomputed for each box phase: held through the leaf phase, whose
        // two prefetched leaves are the walk's register peak, they would spill
        f3 ob = org;
        asm volatile("" : "+v"(ob.x), "+v"(ob.y), "+v"(ob.z));
        const SlabOffsets so = slab_offsets(wr4, face, ob, s.slab_eps, fused);
        for (;;) {
            const bool can = cur != kDone && nq < kLeafQueue;
            if (__ballot(can) == 0ull) break;
            if (4 * __popcll(__ballot(nq == kLeafQueue)) > live) break;
            if (can) {
                if (i >= 0) {
                    const uint4 r0 = nd[i], r1 = nd[i + 1];
                    bool hit = node_test(r0, wr4, so, h.t);
                    int link = (int)r0.w;
                    bool cont = hit || link < 0;
                    if (hit && link < 0) {
                        q1 = q0;
                        q0 = ~link;
                        nq++;
                    }
                    int next = cont ? i + 1 : link;
                    if (cont && i + 1 < n && nq < kLeafQueue) {
                        hit = node_test(r1, wr4, so, h.t);
                        link = (int)r1.w;
                        cont = hit || link < 0;
                        if (hit && link < 0) {
                            q1 = q0;
                            q0 = ~link;
                            nq++;
                        }
                        next = cont ? i + 2 : link;
                    }
                    i = next;
                    if (i >= n) cur = kDone;
                } else if (cur >= 0) {
                    const uint4* __restrict__ R = wr + 4 * (size_t)cur;
                    const uint4 a = R[0], b = R[1], c = R[2], d = R[3];
                    // the members' links are held in registers: left to itself the compiler
                    // fetches one again after the tests, a dependent load on the walk's chain
                    unsigned int wa = a.w, wb = b.w, wc = c.w, wd = d.w;
                    asm volatile("" : "+v"(wa), "+v"(wb), "+v"(wc), "+v"(wd));
                    // an empty slot's infinite near bound fails it (see the record packing)
                    const bool ha = wide_test(a, wr4, so, h.t), hb = wide_test(b, wr4, so, h.t);
                    const bool hc = wide_test(c, wr4, so, h.t), hd = wide_test(d, wr4, so, h.t);
                    const int nh = (int)ha + (int)hb + (int)hc + (int)hd;
                    if (sp + nh - 1 > depth) {
                        const uint4 ps = s.wide_pos[(size_t)face * (size_t)s.wide_stride + cur];
                        i = (int)(ha ? ps.x : hb ? ps.y : hc ? ps.z : ps.w);  // the stack is not used again
                    } else {
                        // members hit after the first are pushed last to first
                        int first = kDone;
                        if (hd) first = (int)wd;
                        if (hc) {
                            if (first != kDone) {
                                *top = first;
                                top += stride;
                                sp++;
                            }
                            first = (int)wc;
                        }
                        if (hb) {
                            if (first != kDone) {
                                *top = first;
                                top += stride;
                                sp++;
                            }
                            first = (int)wb;
                        }
                        if (ha) {
                            if (first != kDone) {
                                *top = first;
                                top += stride;
                                sp++;
                            }
                            first = (int)wa;
                        }
                        cur = first;
                        if (cur == kDone && sp > 0) {
                            sp--;
                            top -= stride;
                            cur = *top;
                        }
                        if (cur < 0 && cur != kDone) {
                            // a leaf next: queued in the same step (the queue had room)
                            q1 = q0;
                            q0 = ~cur;
                            nq++;
                            cur = kDone;
                            if (sp > 0) {
                                sp--;
                                top -= stride;
                                cur = *top;
                            }
                        }
                    }
                } else {
                    q1 = q0;
                    q0 = ~cur;
                    nq++;
                    cur = kDone;
                    if (sp > 0) {
                        sp--;
                        top -= stride;
                        cur = *top;
                    }
                }
            }
        }
        if (__ballot(nq > 0) == 0ull) break;
        // the leaves in the order they were met: with two queued, q1 came first
        const int qa = nq == 2 ? q1 : q0, qb = q0;
        // both queued leaves are fetched before the first test: the second leaf's tests need the
        // first test's result, its data does not
        float4 x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), x1 = x0, x2 = x0, x3 = x0;
        float4 y0 = x0, y1 = x0, y2 = x0, y3 = x0;
        const bool stored = s.leaf_box_stored != 0;
        if (nq > 0) {
            const float4* T = s.tri_isect + 4 * (size_t)qa;
            x0 = T[0];
            x1 = T[1];
            x2 = T[2];
            if (stored) x3 = T[3];
        }
        if (nq > 1) {
            const float4* T = s.tri_isect + 4 * (size_t)qb;
            y0 = T[0];
            y1 = T[1];
            y2 = T[2];
            if (stored) y3 = T[3];
        }
        f3 lo, hi;
        if (nq > 0) {
            leaf_box(stored, x0, x1, x2, x3, lo, hi);
            if (box_test_face(wr4, face, org, lo, hi, h.t)) tri_test<PHOTON>(x0, x1, x2, org, dir, cull, qa, h);
        }
        if (nq > 1) {
            leaf_box(stored, y0, y1, y2, y3, lo, hi);
            if (box_test_face(wr4, face, org, lo, hi, h.t)) tri_test<PHOTON>(y0, y1, y2, org, dir, cull, qb, h);
        }
        nq = 0;
    }
    hit.tri = h.tri;
    hit.t = h.t;
    hit_barycentrics(s, org, dir, hit);
}

template <bool PHOTON>
PPM_DEV Hit trace(const DevScene& s, f3 org, f3 dir, bool cull, int depth) {
    const f3 rcp = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    Hit h;
    h.tri = -1;
    h.t = PHOTON ? 1.0e30f : 1.0e20f;
    h.bx = h.by = h.bz = 0.0f;
    walk<PHOTON>(s, cube_face(dir), org, dir, rcp, cull, h, depth);
    return h;
}

struct Surface {
    f3 pos, nrm, gnrm, col;
    float eta, g;
    int brdf, mat;
};


// Shading fetch after the walk (eyeraytrace.fs:256-287).
PPM_DEV Surface shade(const DevScene& s, const Hit& h) {
    Surface r;
    const float4* t = s.tri_shade + (size_t)4 * h.tri;
    const float4 a = t[0], b = t[1], c = t[2], d = t[3];
    const f3 N0 = xyz(a), N1 = xyz(b), N2 = xyz(c);
    r.pos = h.pos;
    r.nrm = normalize3((N2 * h.bx + N0 * h.by) + N1 * h.bz);
    r.gnrm = xyz(d);
    r.mat = __float_as_int(d.w);
    const float4 m0 = s.mats[2 * r.mat], m1 = s.mats[2 * r.mat + 1];
    r.col = xyz(m0);
    r.eta = m0.w;
    r.brdf = (int)m1.x;
    r.g = m1.y;
    return r;
}

// onb (eyeraytrace.fs:293-322)
PPM_DEV f3 onb(f3 x, f3 n) {
    if (n.z < -0.9999999f) {
        const f3 u = mk3(x.z, x.x, 0.0f);
        return x.y * n - u;
    }
    const float a = 1.0f / (1.0f + n.z);
    const float b = -n.x * n.y * a;
    const f3 u = mk3(1.0f - n.x * n.x * a, b, -n.x);
    const f3 w = mk3(b, 1.0f - n.y * n.y * a, -n.y);
    return (x.x * u + x.y * n) + x.z * w;
}

// Fresnel (eyeraytrace.fs:350-367)
PPM_DEV float fresnel(f3 incom, f3 normal, float ni, float ne) {
    const float eta = ni / ne;
    const float c1 = dot3(incom, normal);
    float c2 = 1.0f - (eta * eta) * (1.0f - c1 * c1);
    if (c2 < 0.0f) return 1.0f;
    c2 = sqrtf(c2);
    const float rs = (ni * c1 - ne * c2) / (ni * c1 + ne * c2);
    const float rp = (ni * c2 - ne * c1) / (ni * c2 + ne * c1);
    return (rs * rs + rp * rp) * 0.5f;
}

// glossy_reflect (eyeraytrace.fs:370-392)
PPM_DEV f3 glossy_reflect(f3 d, f3 n, float g, float4& rndv) {
    const f3 r = normalize3(((1.0f - g) - g) * reflect3(d, n) + (g + g) * n);
    const float rnd1 = gpu_rnd(rndv);
    const float rnd2 = gpu_rnd(rndv);
    const float temp1 = 6.283184f * rnd1;
    const float temp2_ = ppm_powf(rnd2, g);
    const float temp2 = sqrtf(1.0f - temp2_ * temp2_);
    float st, ct;
    ppm_sincosf(temp1, st, ct);
    const f3 v = mk3(st * temp2, temp2_, ct * temp2);
    f3 result = normalize3(onb(v, r));
    if (dot3(result, n) < 0.0f) result = reflect3(result, n);
    return result;
}

// One axis of a GL_LINEAR, GL_REPEAT lookup: texel coordinate coord * size - 1/2, its floor
// wrapped into [0, size) and the next texel, and the fraction as the weight of the next texel.
// A coordinate too large for an exact floor (or NaN) reads texel 0 with weight 0.
PPM_DEV float tex_axis(float coord, int size, int& i0, int& i1) {
    const float a = coord * (float)size - 0.5f;
    if (!(fabsf(a) <= 1.0e18f)) {
        i0 = 0;
        i1 = size > 1 ? 1 : 0;
        return 0.0f;
    }
    const float fl = floorf(a);
    long long k = (long long)fl % (long long)size;
    if (k < 0) k += size;
    i0 = (int)k;
    i1 = i0 + 1 == size ? 0 : i0 + 1;
    return a - fl;
}

PPM_DEV f3 texel(const DevScene& s, int i, int j, int k) {
    const uchar4 c = s.volume[((size_t)k * s.tex_size + j) * s.tex_size + i];
    return mk3((float)c.x / 255.0f, (float)c.y / 255.0f, (float)c.z / 255.0f);
}

PPM_DEV f3 lerp3(f3 p, f3 q, float w) { return p * (1.0f - w) + q * w; }

// texture3D(VolumeTextureTextures, (u, v, r)).rgb (gpurt.cpp:158-198: RGB8, 512 x 512 x M,
// GL_LINEAR, GL_REPEAT): trilinear weights of the GL specification, interpolated along u, then v,
// then r, in fp32.
PPM_DEV f3 tex3d(const DevScene& s, float u, float v, float r) {
    int i0, i1, j0, j1, k0, k1;
    const float a = tex_axis(u, s.tex_size, i0, i1);
    const float b = tex_axis(v, s.tex_size, j0, j1);
    const float c = tex_axis(r, s.num_materials, k0, k1);
    const f3 y00 = lerp3(texel(s, i0, j0, k0), texel(s, i1, j0, k0), a);
    const f3 y10 = lerp3(texel(s, i0, j1, k0), texel(s, i1, j1, k0), a);
    const f3 y01 = lerp3(texel(s, i0, j0, k1), texel(s, i1, j0, k1), a);
    const f3 y11 = lerp3(texel(s, i0, j1, k1), texel(s, i1, j1, k1), a);
    return lerp3(lerp3(y00, y10, b), lerp3(y01, y11, b), c);
}

template <typename T>
PPM_DEV T wave_min(T v) {
    for (int o = 32; o > 0; o >>= 1) v = min1(v, __shfl_xor(v, o, 64));
    return v;
}
template <typename T>
PPM_DEV T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) v = max1(v, __shfl_xor(v, o, 64));
    return v;
}
PPM_DEV unsigned int wave_sum(unsigned int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// Pixel of this thread: tile of 2x2 waves, wave (0..3) = ww x wh pixels (ww = 2^wave_shift).
PPM_DEV void tile_pixel(const Tiles& tl, int tile, int wave, int& i, int& j) {
    const int sh = tl.wave_shift, ww = 1 << sh, wh = 64 >> sh;
    const int tx = tile % tl.tiles_x, ty = tile / tl.tiles_x;
    const int lane = threadIdx.x & 63;
    i = tx * 2 * ww + (wave & 1) * ww + (lane & (ww - 1));
    j = ty * 2 * wh + (wave >> 1) * wh + (lane >> sh);
}

}  // namespace

// ------------------------------------------------------------------------------------------

// Frame-0 state of a job: E, N, tau, R zero, every photon new, the hash and buckets clear.
__global__ void k_init_state(Job J) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < J.gp.count) {
        J.hs.e_n[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        J.hs.tau_r[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        J.hs.em[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        J.hs.em[J.hs.q + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (i < J.pp.count) {
        for (int k = 0; k < 2; k++) {
            const size_t o = (size_t)k * (size_t)J.pp.count + (size_t)i;
            J.ps.pos[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            J.ps.flux[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            J.ps.dir[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        J.ps.state[i] = make_int2(-1, 0);  // (-1, -1, 0, 1e30): every photon starts new
        J.ps.slot[i] = -1;
    }
    if (i < J.h.size)
        for (int k = 0; k < 3; k++)
            reinterpret_cast<uint2*>(J.h.slots)[(size_t)k * (size_t)J.h.size + (size_t)i] = make_uint2(0u, 0xffffffffu);
}

namespace {

// XORShift::m_frand (glslppm.cpp:77-91)
PPM_DEV float xs_frand(unsigned int* s) {
    const unsigned int t = s[0] ^ (s[0] << 11);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    return (float)s[3] * (1.0f / 4294967295.0f);
}

}  // namespace

// The seed grids (RandomizeTextures, glslppm.cpp:113-140): the eye grid's q texels, then the photon
// grid's p, four draws of the XORShift stream each, scaled by 2^22; g++ evaluates the TVector4
// constructor arguments last to first, so w takes the first draw and x the fourth.  Thread c
// draws texels [c chunk, (c + 1) chunk) of the two grids in sequence from starts[c], the state of
// the stream at that chunk.
__global__ void k_seed(float4* eye, int q, float4* photon, int p, const uint4* starts, int chunk) {
    const long long n = (long long)q + (long long)p;
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long k = c * chunk;
    if (k >= n) return;
    const uint4 st = starts[c];
    unsigned int s[4] = {st.x, st.y, st.z, st.w};
    const long long end = k + chunk < n ? k + chunk : n;
    for (; k < end; k++) {
        const float w = xs_frand(s) * 4194304.0f;
        const float z = xs_frand(s) * 4194304.0f;
        const float y = xs_frand(s) * 4194304.0f;
        const float x = xs_frand(s) * 4194304.0f;
        const float4 v = make_float4(x, y, z, w);
        if (k < q) eye[k] = v;
        else photon[k - q] = v;
    }
}

namespace {

// Per-pass uniforms of m_display (glslppm.cpp:214-287): eye pass when FrameCount % MaxPathLength
// is 0, with two AAOffset draws from the job's XORShift stream and NumEyeSamples = count + 1;
// FrameCount 0 is the first pass.  A job whose pass budget is spent stops advancing; the budget
// counts this pass.  The pass's hash table, photon half and the uniforms its gather reads (which
// runs one launch later) are fixed here.
PPM_DEV void pass_begin(const Job& J, unsigned int pass, int defer_all) {
    DevGlobals* g = J.g;
    const unsigned int left = g->passes_left;
    if (left == 0u) {
        g->active = 0;
        return;
    }
    if (left != 0xffffffffu) g->passes_left = left - 1u;
    const unsigned long long frame = g->frame;
    const unsigned long long L = (unsigned long long)J.ep.max_path;
    const int eye = (frame % L) == 0ull;
    g->active = 1;
    g->eye_now = eye;
    g->first_pass = frame == 0ull;
    if (eye) g->eye_samples = g->eye_samples + 1ull;
    const int half = (int)(frame & 1ull);
    g->table = (int)(frame % 3ull);
    g->half = half;
    GatherSnap& sn = g->snap[half];
    sn.first_pass = frame == 0ull;
    sn.eye_now = eye;
    sn.cb = (int)((g->eye_samples + 1ull) & 1ull);
    sn.table = g->table;
    sn.half = half;
    // the deposits of an eye pass wait for k_hash unless its eye stage has already set the grid up
    // (or, with defer_all, always: a batch that has not yet waited for the stages begun ahead)
    g->defer = eye && (defer_all || (unsigned long long)g->eye_stages_done < g->eye_samples);
    // for an eye stage begun ahead in this pass: the passes until the next eye pass, and whether
    // the budget lasts until then (left counts this pass)
    const unsigned long long to_next = L - frame % L;
    g->ahead_frames = (int)to_next;
    g->ahead_ok = left == 0xffffffffu || (unsigned long long)(left - 1u) >= to_next;
    g->begin_pass = pass;
    g->frame = frame + 1ull;
}

// The start of an eye stage (glslppm.cpp:238-287): the AAOffset draws of the job's XORShift
// stream, NumEyeSamples, and the buffers the stage writes.  kEyeNow begins the stage of the eye pass
// in progress unless it has already begun.  kEyeAhead begins the stage of the job's next eye pass
// (after an eye pass, beside the photon passes that follow it) once the stage before it is done,
// if the job's budget reaches that eye pass and, for a batch (pass != kAnyPass), only in the pass
// whose k_pass_begin last ran (a launch that runs late leaves it to a later one), not in an eye pass
// (the gather of the pass before it, which runs in that eye pass's photon-stage launch, still reads
// the buffer the stage would write; the plain loop starts its ahead stages after that launch) and
// with at least `lag` passes left before that eye pass: the batch's passes wait for the ahead
// stages of the pass `lag` passes back, which then covers this stage.  Stages begin in eye-pass order and draw the
// same numbers wherever they run, so the results do not depend on where a stage ran.  A job has one
// stage in progress at most: the two conditions exclude each other in any interleaving of their
// launches, and eye_go (which tags the stage's launches) is written only here and by k_grid.
PPM_DEV void eye_begin(const Job& J, int mode, unsigned int pass, int lag) {
    DevGlobals* g = J.g;
    if (!g->active) return;
    const unsigned long long begun = g->eye_stages;
    bool go;
    if (mode == kEyeAhead) {
        go = begun == g->eye_samples && (unsigned long long)g->eye_stages_done == begun && g->ahead_ok &&
             (pass == kAnyPass || (g->begin_pass == pass && !g->eye_now && g->ahead_frames >= lag));
    } else {
        go = g->eye_now && begun < g->eye_samples;
    }
    if (!go) return;
    unsigned int xs[4] = {g->xs[0], g->xs[1], g->xs[2], g->xs[3]};
    g->aa_x = (xs_frand(xs) - 0.5f) * 1.25f;
    g->aa_y = (xs_frand(xs) - 0.5f) * 1.25f;
    for (int k = 0; k < 4; k++) g->xs[k] = xs[k];
    const unsigned int e = g->eye_stages;
    g->eye_stage = (int)e;
    g->num_eye_samples = (int)(e + 1u);
    g->eye_stages = e + 1u;
    g->eye_go = mode;
}

// The buffer of the latest eye pass, which the passes after it read.
PPM_DEV int cur_buffer(const DevGlobals* g) { return (int)((g->eye_samples + 1ull) & 1ull); }

// The start of a pass (one block): its first wave folds the last pass's emitted counts into
// NumPhotons (glslppm.cpp:396-397, `NumPhotons += int(sum)` once per pass, in float), clears the
// counters, marks the pass before for the gather blocks of this pass's photon-stage launch (unless
// k_gather already ran it), with NumPhotons after it for its draw, and sets the pass's uniforms
// (pass_begin).  The pass's hash table was cleared by the photon stage of the pass before (the
// glClear of glslppm.cpp:421, 450; see next_table).
PPM_DEV void begin_pass(const Job& J, int blk, unsigned int pass, int defer_all) {
    DevGlobals* g = J.g;
    if (blk == 0 && threadIdx.x < 64) {
        unsigned int* sh = g->emit_shard + threadIdx.x * kEmitStride;
        const unsigned int v = wave_sum(*sh);
        *sh = 0u;
        if (threadIdx.x == 0) {
            g->num_photons_f = g->num_photons_f + (float)(int)v;
            g->photons_emitted += v;
            // `active` is still that of the pass before, begun by the last k_pass_begin
            const int lag = g->active && g->gathered != g->frame;
            g->lag_pending = lag;
            g->lag_idx = g->half;
            if (lag) g->snap[g->half].np = g->num_photons_f;
            pass_begin(J, pass, defer_all);
        }
    }
}

// Hash table k of three.  Pass f (0-based) deposits into table f % 3 (set by its k_pass_begin);
// the readback of the hash takes the same table after the last pass.
PPM_DEV unsigned int* table_at(const HashState& h, int k) { return h.slots + (size_t)2 * (size_t)h.size * (size_t)k; }
PPM_DEV unsigned int* hash_table(const HashState& h, const DevGlobals* g) { return table_at(h, g->table); }
// The table of the next pass, which the photon stage of the pass in progress clears.  The pass two
// before used it; its gather ran in the launch of the pass before, which has finished.
PPM_DEV unsigned int* next_table(const HashState& h, const DevGlobals* g) {
    return table_at(h, g->table == 2 ? 0 : g->table + 1);
}

// A view of the photon records with position, flux and direction in half k (rng, state and slot
// are single).
PPM_DEV PhotonState photon_half(const PhotonState& ps, int count, int k) {
    PhotonState v = ps;
    const size_t o = (size_t)count * (size_t)k;
    v.pos += o;
    v.flux += o;
    v.dir += o;
    return v;
}

// NumPhotons after the pass in progress: the folded count plus the pass's counters (every
// thread of the block calls it; one wave reads the counters).
PPM_DEV float num_photons_now(const DevGlobals* g) {
    __shared__ unsigned int pend;
    if (threadIdx.x < 64) {
        const unsigned int v = wave_sum(g->emit_shard[threadIdx.x * kEmitStride]);
        if (threadIdx.x == 0) pend = v;
    }
    __syncthreads();
    return g->num_photons_f + (float)(int)pend;
}

}  // namespace

namespace {

// eyeraytrace.fs:396-554 for one pixel, then the min/max of hit positions.
PPM_DEV void eye(const Job& J, int blk, int mode) {
    const DevGlobals* g = J.g;
    if (g->eye_go != mode) return;
    const DevScene& s = J.s;
    const DevCamera& cam = J.cam;
    const EyeParams& ep = J.ep;
    const HitState& hs = J.hs;
    const float aa_x = g->aa_x, aa_y = g->aa_y;
    // the tiles whose paths went deepest in the last eye pass first (k_grid's order): a path is a
    // chain of dependent walks and a wave lasts as long as its deepest path, so the longest waves
    // start first and the short ones fill in around them
    const int tile = g->order_tiles == J.tile_blocks ? hs.tile_order[blk] : blk;
    int px_i, px_j;
    tile_pixel(J.tl, tile, (int)(threadIdx.x >> 6), px_i, px_j);
    const int p = px_j * ep.width + px_i;
    f3 pos = mk3(0.0f, 0.0f, 0.0f);
    int segs = 0;
    const bool active = px_i < ep.width && px_j < ep.height;
    // this stage writes buffer b from the previous stage's buffer (generator states, emission)
    const size_t wb = (size_t)(g->eye_stage & 1) * (size_t)hs.q + (size_t)p;
    const size_t rb = (size_t)((g->eye_stage & 1) ^ 1) * (size_t)hs.q + (size_t)p;
    if (active) {
        float4 rndv = hs.rng[rb];
        const f3 C = mk3(cam.ox, cam.oy, cam.oz);
        const f3 U = mk3(cam.ux, cam.uy, cam.uz), V = mk3(cam.vx, cam.vy, cam.vz), W = mk3(cam.wx, cam.wy, cam.wz);
        const float ppx = ((float)px_i + 0.5f + aa_x) - cam.half_w;
        const float ppy = ((float)px_j + 0.5f + aa_y) - cam.half_h;
        const f3 rel = (ppx * U + ppy * V) + cam.dist * W;
        f3 dir = normalize3(rel);
        const float filter = (1.0f - fabsf(aa_x) / 1.5f) * (1.0f - fabsf(aa_y) / 1.5f);
        const f3 fp = C + dir * ep.focal_length;
        const float radius = sqrtf(gpu_rnd(rndv)) * ep.aperture;
        const float theta = 6.283184f * gpu_rnd(rndv);
        float st, ct;
        ppm_sincosf(theta, st, ct);
        const f3 lens = C + radius * (U * ct + V * st);
        dir = normalize3(fp - lens);
        f3 org = lens;
        f3 col = mk3(filter, filter, filter);
        f3 nrm = mk3(0.0f, 0.0f, 0.0f);
        f3 emi = mk3(0.0f, 0.0f, 0.0f);
        const float eps = 1e-5f;
        int last_tri = -1, last_mat = -1;
        float last_t = 1.0e20f;
        for (int j = 0; j < ep.max_path; j++) {
            segs = j + 1;
            const Hit h = trace<false>(s, org, dir, j == 0, J.walk_stack);
            last_tri = h.tri;
            last_t = h.t;
            if (h.tri < 0) {
                last_mat = -1;
                nrm = mk3(0.0f, 0.0f, 0.0f);
                if (s.light_area == 0.0f) emi = mk3(0.7f, 0.7f, 1.0f) * col;
                break;
            }
            const Surface i = shade(s, h);
            last_mat = i.mat;
            pos = i.pos;
            nrm = i.nrm;
            // Result.tex = T2 B.x + T0 B.y + T1 B.z; textured surfaces scale their colour
            // (eyeraytrace.fs:277-280, 443-446)
            f3 icol = i.col;
            if (s.tri_tex) {
                const float4 ta = s.tri_tex[2 * h.tri], tb = s.tri_tex[2 * h.tri + 1];
                const float tu = (tb.x * h.bx + ta.x * h.by) + ta.z * h.bz;
                const float tv = (tb.y * h.bx + ta.y * h.by) + ta.w * h.bz;
                if (fabsf(tu) < 1e10f) icol = icol * tex3d(s, tu, tv, s.material_rcp * ((float)i.mat + 0.5f));
            }
            if (i.brdf == -1) {
                nrm = mk3(0.0f, 0.0f, 0.0f);
                emi = col * icol;
                break;
            } else if (i.brdf == 0) {
                col = col * icol;
                break;
            } else if (i.brdf == 1 || i.brdf == 4) {
                org = pos + eps * i.gnrm;
                if (i.brdf == 4) nrm = glossy_reflect(nrm, nrm, i.g, rndv);
                dir = reflect3(dir, nrm);
                if (dot3(dir, i.gnrm) < 0.0f) dir = -dir;
                col = col * icol;
            } else if (i.brdf == 2 || i.brdf == 5) {
                if (i.brdf == 5) nrm = glossy_reflect(nrm, nrm, i.g, rndv);
                const float ln = dot3(nrm, dir);
                const float eta = i.eta;
                if (ln < 0.0f) {
                    const float re = fresnel(-dir, nrm, 1.0f, eta);
                    if (gpu_rnd(rndv) < re) {
                        org = pos + eps * i.gnrm;
                        dir = reflect3(dir, nrm);
                        if (dot3(dir, i.gnrm) < 0.0f) dir = -dir;
                    } else {
                        col = col * icol;
                        org = pos - eps * i.gnrm;
                        dir = refract3(dir, nrm, 1.0f / eta);
                        if (dot3(dir, i.gnrm) > 0.0f) dir = -dir;
                    }
                } else {
                    const float re = fresnel(-dir, -nrm, eta, 1.0f);
                    col = col * icol;
                    if (gpu_rnd(rndv) < re) {
                        org = pos - eps * i.gnrm;
                        dir = reflect3(dir, -nrm);
                        if (dot3(dir, -i.gnrm) < 0.0f) dir = -dir;
                    } else {
                        org = pos + eps * i.gnrm;
                        dir = refract3(dir, -nrm, eta);
                        if (dot3(dir, -i.gnrm) > 0.0f) dir = -dir;
                    }
                }
            } else if (i.brdf == 3 || i.brdf == 6) {
                if (i.brdf == 6) nrm = glossy_reflect(nrm, nrm, i.g, rndv);
                const float re = fresnel(-dir, nrm, 1.0f, i.eta);
                if (gpu_rnd(rndv) < re) {
                    org = pos + eps * i.gnrm;
                    dir = reflect3(dir, nrm);
                    if (dot3(dir, i.gnrm) < 0.0f) dir = -dir;
                } else {
                    col = col * icol;
                    break;
                }
            }
        }
        float4 en = hs.em[rb];
        const float a = 1.0f / (float)g->num_eye_samples;
        en.x = en.x * (1.0f - a) + emi.x * a;
        en.y = en.y * (1.0f - a) + emi.y * a;
        en.z = en.z * (1.0f - a) + emi.z * a;
        hs.pos[wb] = make_float4(pos.x, pos.y, pos.z, dir.x);
        hs.refl[wb] = make_float4(col.x, col.y, col.z, dir.y);
        hs.nrm[wb] = make_float4(nrm.x, nrm.y, nrm.z, dir.z);
        hs.rng[wb] = rndv;
        hs.tri[wb] = last_tri;
        hs.mat_t[wb] = make_float2(last_mat < 0 ? -1.0f : (float)last_mat, last_t);
        hs.em[wb] = en;
    }

    // max.fs / min.fs over every pixel's position (misses contribute their last position): one
    // partial per block, reduced by k_grid
    __shared__ float red[kWaves][6];
    __shared__ int deep[kWaves];
    const int d = wave_max(active ? segs : 0);
    float v[6] = {active ? pos.x : kInf, active ? pos.y : kInf, active ? pos.z : kInf,
                  active ? pos.x : -kInf, active ? pos.y : -kInf, active ? pos.z : -kInf};
    for (int k = 0; k < 3; k++) v[k] = wave_min(v[k]);
    for (int k = 3; k < 6; k++) v[k] = wave_max(v[k]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        for (int k = 0; k < 6; k++) red[wave][k] = v[k];
        deep[wave] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[6];
        for (int k = 0; k < 6; k++) r[k] = red[0][k];
        for (int w = 1; w < kWaves; w++) {
            for (int k = 0; k < 3; k++) r[k] = min1(r[k], red[w][k]);
            for (int k = 3; k < 6; k++) r[k] = max1(r[k], red[w][k]);
        }
        for (int k = 0; k < 6; k++) hs.partial[blk * 6 + k] = r[k];
        int dm = deep[0];
        for (int w = 1; w < kWaves; w++) dm = dm > deep[w] ? dm : deep[w];
        hs.tile_cost[tile] = dm;
    }
}

// The hit-point bounding box from the eye blocks' partials, then the grid scalars of
// glslppm.cpp:294-311 (one block of 1024 threads).
PPM_DEV void grid(const Job& J, int mode) {
    DevGlobals* g = J.g;
    if (g->eye_go != mode) return;
    GridScalars& gr = g->grid[g->eye_stage & 1];
    const EyeParams& ep = J.ep;
    const float* partial = J.hs.partial;
    const int n_partial = J.tile_blocks;
    __shared__ float red[16][6];
    float r[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
    for (int b = threadIdx.x; b < n_partial; b += 1024) {
        const float* pp = partial + b * 6;
        for (int k = 0; k < 3; k++) r[k] = min1(r[k], pp[k]);
        for (int k = 3; k < 6; k++) r[k] = max1(r[k], pp[k]);
    }
    for (int k = 0; k < 3; k++) r[k] = wave_min(r[k]);
    for (int k = 3; k < 6; k++) r[k] = wave_max(r[k]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 6; k++) red[wave][k] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        float bmin[3], bmax[3];
        for (int k = 0; k < 3; k++) {
            bmin[k] = red[0][k];
            bmax[k] = red[0][k + 3];
            for (int w = 1; w < 16; w++) {
                bmin[k] = min1(bmin[k], red[w][k]);
                bmax[k] = max1(bmax[k], red[w][k + 3]);
            }
        }
        float bbsize = 0.0f;
        for (int k = 0; k < 3; k++) bbsize += bmax[k] - bmin[k];
        float ir = (bbsize / 3.0f) / ep.footprint_width * ep.footprint;
        if (ep.fixed_radius > 0.0f) ir = ep.fixed_radius;
        for (int k = 0; k < 3; k++) {
            gr.bbmin[k] = bmin[k] - ir;
            gr.bbmax[k] = bmax[k] + ir;
        }
        const float gs = 0.5f / ir;
        gr.initial_radius = ir;
        gr.grid_scale = gs;
        gr.hash_scale = 4194304.0f / gs;
    }
    // the next eye pass's tile order: a counting sort of the tiles by the depth of their deepest
    // path, deepest first (any order of the tiles gives the same results)
    __shared__ unsigned int hist[33];
    const int n = J.tile_blocks;
    if (threadIdx.x < 33) hist[threadIdx.x] = 0u;
    __syncthreads();
    const HitState& hs = J.hs;
    for (int b = threadIdx.x; b < n; b += 1024) atomicAdd(&hist[min(max(hs.tile_cost[b], 0), 32)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int run = 0;
        for (int c = 32; c >= 0; c--) {
            const unsigned int k = hist[c];
            hist[c] = run;
            run += k;
        }
        g->order_tiles = n;
        g->eye_stages_done = (unsigned int)g->eye_stage + 1u;
        g->eye_go = 0;  // every thread read it before the first barrier
    }
    __syncthreads();
    for (int b = threadIdx.x; b < n; b += 1024) hs.tile_order[atomicAdd(&hist[min(max(hs.tile_cost[b], 0), 32)], 1u)] = b;
}

}  // namespace

namespace {

// GenerateLightSourceSample (photontrace.fs:345-440)
PPM_DEV void light_sample(const DevScene& s, float4& rndv, f3& flux, f3& dir, f3& org) {
    const float rnd = gpu_rnd(rndv);
    // lower bound over the texels at (mid - 0.5) / n, i.e. cdf[max(mid - 1, 0)]
    int low = 0, high = s.num_lights;
    while (low < high) {
        const int mid = low + ((high - low) / 2);
        const float c = s.light_cdf[mid > 0 ? mid - 1 : 0];
        if (c < rnd) low = mid + 1;
        else high = mid;
    }
    const int li = s.light_tri[high > 0 ? high - 1 : 0];
    const float rndv2 = gpu_rnd(rndv);
    const float t = sqrtf(1.0f - rndv2);
    const float sv = gpu_rnd(rndv);
    const float a = 1.0f - t;
    const float c = sv * t;
    const float b = t - c;
    const float4* tv = s.tri_isect + (size_t)4 * li;
    const float4* tr = s.tri_shade + (size_t)4 * li;
    const f3 V0 = xyz(tv[0]), V1 = xyz(tv[1]), V2 = xyz(tv[2]);
    const float4 q3 = tr[3];
    const f3 N0 = xyz(tr[0]), N1 = xyz(tr[1]), N2 = xyz(tr[2]);
    const int mat = __float_as_int(q3.w);
    const float4 m0 = s.mats[2 * mat];
    const float le = s.light_area * 3.141592f;
    flux = mk3(m0.x * le, m0.y * le, m0.z * le);
    // T = T2 c + T0 a + T1 b on textured lights (photontrace.fs:413-422)
    if (s.tri_tex) {
        const float4 ta = s.tri_tex[2 * li], tb = s.tri_tex[2 * li + 1];
        const float tu = (tb.x * c + ta.x * a) + ta.z * b;
        const float tv = (tb.y * c + ta.y * a) + ta.w * b;
        if (fabsf(tu) < 1e10f) flux = flux * tex3d(s, tu, tv, s.material_rcp * ((float)mat + 0.75f));
    }
    org = (V0 * a + V1 * b) + V2 * c;
    const f3 nrm = normalize3((N0 * a + N1 * b) + N2 * c);
    const f3 gnrm = xyz(q3);
    org = org + gnrm * 1.0e-5f;
    float rx = gpu_rnd(rndv);
    float ry = gpu_rnd(rndv);
    rx = 6.283184f * rx;
    ry = sqrtf(ry);
    float sx, cx;
    ppm_sincosf(rx, sx, cx);
    dir = onb(mk3(sx * ry, sqrtf(1.0f - ry * ry), cx * ry), nrm);
}

// GenerateIBLSample (photontrace.fs:283-341), used when the scene has no light triangle
PPM_DEV void ibl_sample(const DevScene& s, float4& rndv, f3& flux, f3& dir, f3& org) {
    const float area = 12.566368f * s.bsphere_r * s.bsphere_r;
    const float temp2 = 6.283184f * gpu_rnd(rndv);
    const float aaa = sqrtf(1.0f - gpu_rnd(rndv));
    float sin_a, cos_a, s2, c2;
    ppm_sincosf(aaa, sin_a, cos_a);
    ppm_sincosf(temp2, s2, c2);
    const float sin_t1 = (sin_a + sin_a) * cos_a;
    dir.x = sin_t1 * c2;
    dir.y = 1.0f - (sin_a + sin_a) * sin_a;
    dir.z = sin_t1 * s2;
    if (gpu_rnd(rndv) > 0.5f) {
        dir = mk3(1.0f, 1.0f, 1.0f);
        dir.y = dir.y + dir.y;
        dir.z = -dir.z;
        dir = normalize3(dir);
        flux = mk3(1.0f * area, 1.0f * area, 0.7f * area);
    } else {
        flux = mk3(0.7f * area, 0.7f * area, 1.0f * area);
    }
    const float radius = sqrtf(gpu_rnd(rndv));
    const float theta = 6.283184f * gpu_rnd(rndv);
    float st, ct;
    ppm_sincosf(theta, st, ct);
    const f3 temp = mk3(radius * ct, 0.0f, radius * st);
    org = (onb(temp * s.bsphere_r, dir) + dir * s.bsphere_r) + mk3(s.bsphere_x, s.bsphere_y, s.bsphere_z);
    dir = -dir;
}

}  // namespace

// photontrace.fs:464-720, hash.fs and the scatter/correction draws, as one kernel.
//
// k_photon_stage runs one thread per photon slot: emission or bounce with Russian roulette
// (photontrace.fs:466-686); a terminated photon writes its final state at once, a continuing one
// goes on through the walk and the deposit (photontrace.fs:690-710), hash.fs, scatter and
// correction.  Each wave adds its emitted flags (sum.fs) to one of the job's counters.  The slot
// table only sees commutative atomics, so the result equals the reference's one fragment per
// photon texel.
// (An earlier form compacted the continuing rays into a queue for a separate walk kernel; the
// queue's 56 bytes per photon written and read again, and two more launches, cost more than the
// idle lanes of the photons that end: 1.95 -> 2.02 G photon segments/s on the 205k-triangle
// spheres, 2.61 -> 2.81 G on the 2.26M-triangle terrain.)

namespace {

// Emission or bounce of slot p from its record in ps_in (the pass before's half).  Returns false
// when Russian roulette ends the photon, after writing the terminated state into ps (this pass's
// half; photontrace.fs:711-719).
PPM_DEV bool photon_gen(const DevScene& s, const PhotonParams& pp, const PhotonState& ps_in, const PhotonState& ps,
                        int p, f3& org_out, f3& dir_out, f3& flux_out, int& out_bits, unsigned int& emitted_out) {
    float4 rndv = ld_stream(ps.rng + p);
    const float4 ppos = ld_stream(ps_in.pos + p), pflux = ld_stream(ps_in.flux + p), pdir = ld_stream(ps_in.dir + p);
    const f3 prev_nrm = mk3(ppos.w, pflux.w, pdir.w);
    f3 flux = mk3(fabsf(pflux.x), fabsf(pflux.y), fabsf(pflux.z));
    const int2 st = ld_stream(ps.state + p);
    const int bounce = (st.y & 0xffff) + 1;
    f3 org = mk3(0.0f, 0.0f, 0.0f), dir = mk3(0.0f, 0.0f, 0.0f);
    bool cont;
    unsigned int emitted = 0;
    int out_bounce = bounce;
    if (st.x < 0 || bounce >= pp.max_path) {
        if (s.light_area != 0.0f) light_sample(s, rndv, flux, dir, org);
        else ibl_sample(s, rndv, flux, dir, org);
        emitted = 1;
        out_bounce = 0;
        cont = true;
    } else {
        const float4 q3 = s.tri_shade[(size_t)4 * st.x + 3];
        const f3 gnrm = xyz(q3);
        const int mat = __float_as_int(q3.w);
        const float4 m0 = s.mats[2 * mat], m1 = s.mats[2 * mat + 1];
        f3 col = xyz(m0);
        const int brdf = (int)m1.x;
        const float eta = m0.w, gl = m1.y;
        const f3 ipos = xyz(ppos);
        if (s.tri_tex) {
            // barycentrics of the stored hit from the three sub-triangle areas, then
            // i.tex = T2 B.x + T0 B.y + T1 B.z (photontrace.fs:519-544, 562-565)
            const float4* tvx = s.tri_isect + (size_t)4 * st.x;
            const f3 V0 = xyz(tvx[0]), V1 = xyz(tvx[1]), V2 = xyz(tvx[2]);
            const float bx = length3(cross3(V1 - V0, ipos - V0));
            const float by = length3(cross3(V2 - V1, ipos - V1));
            const float bz = length3(cross3(V0 - V2, ipos - V2));
            const float bw = (bx + by) + bz;
            const float wx = bx / bw, wy = by / bw, wz = bz / bw;
            const float4 ta = s.tri_tex[2 * st.x], tb = s.tri_tex[2 * st.x + 1];
            const float tu = (tb.x * wx + ta.x * wy) + ta.z * wz;
            const float tv = (tb.y * wx + ta.y * wy) + ta.w * wz;
            if (fabsf(tu) < 1e10f) col = col * tex3d(s, tu, tv, s.material_rcp * (((float)mat + 0.5f) + 0.25f));
        }
        f3 inrm = prev_nrm;
        const f3 pd = xyz(pdir);
        const float eps = 1e-6f;
        if (brdf == 0) {
            const float r0 = 6.283184f * gpu_rnd(rndv);
            const float r1 = sqrtf(gpu_rnd(rndv));
            float s0, c0;
            ppm_sincosf(r0, s0, c0);
            const f3 v = mk3(s0 * r1, sqrtf(1.0f - r1 * r1), c0 * r1);
            org = ipos + eps * gnrm;
            dir = onb(v, inrm);
            if (dot3(dir, gnrm) < 0.0f) dir = -dir;
        } else if (brdf == 1 || brdf == 4) {
            org = ipos + eps * gnrm;
            if (brdf == 4) inrm = glossy_reflect(inrm, inrm, gl, rndv);
            dir = reflect3(pd, inrm);
            if (dot3(dir, gnrm) < 0.0f) dir = -dir;
        } else if (brdf == 2 || brdf == 5) {
            if (brdf == 5) inrm = glossy_reflect(inrm, inrm, gl, rndv);
            const float ln = dot3(inrm, pd);
            if (ln < 0.0f) {
                const float re = fresnel(-pd, inrm, 1.0f, eta);
                if (gpu_rnd(rndv) < re) {
                    org = ipos + eps * gnrm;
                    dir = reflect3(pd, inrm);
                    col = mk3(1.0f, 1.0f, 1.0f);
                    if (dot3(dir, gnrm) < 0.0f) dir = -dir;
                } else {
                    org = ipos - eps * gnrm;
                    dir = refract3(pd, inrm, 1.0f / eta);
                    if (dot3(dir, gnrm) > 0.0f) dir = -dir;
                }
            } else {
                const float re = fresnel(-pd, -inrm, eta, 1.0f);
                if (gpu_rnd(rndv) < re) {
                    org = ipos - eps * gnrm;
                    dir = reflect3(pd, -inrm);
                    if (dot3(dir, -gnrm) < 0.0f) dir = -dir;
                } else {
                    org = ipos + eps * gnrm;
                    dir = refract3(pd, -inrm, eta);
                    if (dot3(dir, -gnrm) > 0.0f) dir = -dir;
                }
            }
        } else if (brdf == 3 || brdf == 6) {
            if (brdf == 6) inrm = glossy_reflect(inrm, inrm, gl, rndv);
            const float re = fresnel(-pd, inrm, 1.0f, eta);
            if (gpu_rnd(rndv) < re) {
                org = ipos + eps * gnrm;
                dir = reflect3(pd, inrm);
                col = mk3(1.0f, 1.0f, 1.0f);
                if (dot3(dir, gnrm) < 0.0f) dir = -dir;
            } else {
                const float r0 = 6.283184f * gpu_rnd(rndv);
                const float r1 = sqrtf(gpu_rnd(rndv));
                float s0, c0;
                ppm_sincosf(r0, s0, c0);
                const f3 v = mk3(s0 * r1, sqrtf(1.0f - r1 * r1), c0 * r1);
                org = ipos + eps * gnrm;
                dir = onb(v, inrm);
                if (dot3(dir, gnrm) < 0.0f) dir = -dir;
            }
        }
        // Russian roulette
        const float pr = max1(max1(col.x, col.y), col.z);
        if (pr < gpu_rnd(rndv)) {
            cont = false;
        } else {
            cont = true;
            col = col / pr;
            flux = flux * col;
        }
    }
    emitted_out = emitted;
    st_stream(ps.rng + p, rndv);
    if (!cont) {
        st_stream(ps.pos + p, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        st_stream(ps.flux + p, make_float4(-1.0f, -1.0f, -1.0f, 0.0f));
        st_stream(ps.dir + p, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        st_stream(ps.state + p, make_int2(-1, (int)(emitted << 16)));
        st_stream(ps.slot + p, -1);
        return false;
    }
    org_out = org;
    dir_out = dir;
    flux_out = flux;
    out_bits = out_bounce | (int)(emitted << 16);
    return true;
}

struct Walk {
    f3 org, dir, rcp;
    int face;
    Hit h;
};

// The slot of a photon that deposits at hpos (hash.fs), and the scatter and correction updates of
// that slot; -1 for a slot past the last row, which lands outside the hash viewport and is clipped.
PPM_DEV int deposit(const PhotonParams& pp, const HashState& h, const DevGlobals* g, int p, f3 hpos) {
    const GridScalars& gr = g->grid[cur_buffer(g)];
    const f3 cell = mk3(floorf((hpos.x - gr.bbmin[0]) * gr.grid_scale), floorf((hpos.y - gr.bbmin[1]) * gr.grid_scale),
                        floorf((hpos.z - gr.bbmin[2]) * gr.grid_scale));
    const float sf = cell_hash(cell, gr.grid_scale, gr.hash_scale, h.num_f);
    if (!(sf < h.num_f)) return -1;
    const int slot = (int)sf;
    unsigned int* t = hash_table(h, g);
    atomicAdd(t + 2 * slot, 1u);
    if (p < pp.rep_limit) atomicMin(t + 2 * slot + 1, (unsigned int)p);
    return slot;
}

PPM_DEV void photon_finish(const DevScene& s, const PhotonParams& pp, const PhotonState& ps, const HashState& h,
                           const DevGlobals* g, int p, const Walk& W, f3 flux, int out_bits) {
    const f3 dir = W.dir;
    f3 hpos = mk3(0.0f, 0.0f, 0.0f), hnrm = mk3(0.0f, 0.0f, 0.0f), hgn = mk3(0.0f, 0.0f, 0.0f);
    int hbrdf = 0;
    if (W.h.tri >= 0) {
        const Surface i = shade(s, W.h);
        hpos = i.pos;
        hnrm = i.nrm;
        hgn = i.gnrm;
        hbrdf = i.brdf;
    }
    // flux = -flux on a miss or a non-diffuse hit, as a sign-bit flip selected without a
    // branch (a divergent short-circuit here was compiled into a lost negation)
    const bool no_deposit = (W.h.tri < 0) | ((hbrdf != 0) & (hbrdf != 6) & (hbrdf != 3));
    const unsigned int sign_bit = no_deposit ? 0x80000000u : 0u;
    flux = mk3(__uint_as_float(__float_as_uint(flux.x) ^ sign_bit),
               __uint_as_float(__float_as_uint(flux.y) ^ sign_bit),
               __uint_as_float(__float_as_uint(flux.z) ^ sign_bit));
    // a miss has no geometric normal: 0/0 is NaN and the NaN-rejecting min gives sqrt(5)
    const float gf = min1(fabsf(dot3(dir, hnrm) / dot3(dir, hgn)), sqrtf(5.0f));
    flux = flux * gf;
    st_stream(ps.pos + p, make_float4(hpos.x, hpos.y, hpos.z, hnrm.x));
    st_stream(ps.flux + p, make_float4(flux.x, flux.y, flux.z, hnrm.y));
    st_stream(ps.dir + p, make_float4(dir.x, dir.y, dir.z, hnrm.z));
    st_stream(ps.state + p, make_int2(W.h.tri, out_bits));
    // hash.fs: deposits with flux.x >= 0 (including -0) go to the grid.  In an eye pass the grid
    // is still being set up while the photons fly, so the deposit waits for k_hash (kSlotPending).
    int slot = -1;
    if (!(flux.x < 0.0f)) slot = g->defer ? kSlotPending : deposit(pp, h, g, p, hpos);
    st_stream(ps.slot + p, slot);
}

// The deposits an eye pass held back, once k_grid has set the grid up.
PPM_DEV void hash_deferred(const Job& J, int blk) {
    const DevGlobals* g = J.g;
    if (!g->active || !g->defer) return;
    const int p = blk * kBlock + threadIdx.x;
    if (p >= J.pp.count || J.ps.slot[p] != kSlotPending) return;
    const float4* pos = J.ps.pos + (size_t)J.pp.count * (size_t)g->half;
    J.ps.slot[p] = deposit(J.pp, J.h, g, p, xyz(pos[p]));
}

// The photon stage in one launch: each thread takes one photon slot through emission or bounce
// (photon_gen), then a continuing photon through its walk and deposit (photon_finish).  No ray
// queue: a photon ended by Russian roulette (about a sixth of them per pass) leaves its lane idle
// during the walk, which costs little, since the walk's time follows the lanes' node fetches.
PPM_DEV void photon_stage(const Job& J, int blk) {
    const DevGlobals* g = J.g;
    if (!g->active) return;
    const DevScene& s = J.s;
    const PhotonParams& pp = J.pp;
    // the next pass's hash table, cleared beside this pass's walks
    {
        uint2* nt = reinterpret_cast<uint2*>(next_table(J.h, g));
        const int stride = J.trace_blocks * (int)blockDim.x;
        for (int i = blk * (int)blockDim.x + (int)threadIdx.x; i < J.h.size; i += stride)
            st_stream(reinterpret_cast<int2*>(nt + i), make_int2(0, -1));
    }
    const int p = blk * (int)blockDim.x + (int)threadIdx.x;
    unsigned int emitted = 0;
    bool live = false;
    f3 org, dir, flux;
    int out_bits = 0;
    const PhotonState ps_out = photon_half(J.ps, pp.count, g->half);
    if (p < pp.count) live = photon_gen(s, pp, photon_half(J.ps, pp.count, g->half ^ 1), ps_out, p, org, dir, flux,
                                        out_bits, emitted);
    // sum.fs: each wave's emitted flags go to its block's counter (folded in by the next pass's
    // k_pass_begin; the gather's draw adds them itself)
    const unsigned int es = wave_sum(emitted);
    if ((threadIdx.x & 63) == 0 && es != 0u) atomicAdd(J.g->emit_shard + (blk % kEmitShards) * kEmitStride, es);
    if (!live) return;
    Walk W;
    W.org = org;
    W.dir = dir;
    W.rcp = mk3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    W.face = cube_face(dir);
    W.h.tri = -1;
    W.h.t = 1.0e30f;
    W.h.bx = W.h.by = W.h.bz = 0.0f;
    walk<true>(s, W.face, W.org, W.dir, W.rcp, false, W.h, J.walk_stack);
    photon_finish(s, pp, ps_out, J.h, g, p, W, flux, out_bits);
}

// re.fs: tone map of one channel
PPM_DEV float srgb(float c) {
    if (c < 0.0031308f) return 12.92f * c;
    const float a = 0.055f;
    return (1.0f + a) * ppm_powf(c, 0.416666666667f) - a;
}

PPM_DEV void radiance(float4 tr, float4 en, float np, float4& lin, float4& mapped) {
    const float den = tr.w * tr.w * 3.141592f * np;
    lin = make_float4(tr.x / den + en.x, tr.y / den + en.y, tr.z / den + en.z, 1.0f + 0.0f);
    const float ex = 60000.0f;
    mapped.x = srgb(1.0f - ppm_expf(-lin.x * ex));
    mapped.y = srgb(1.0f - ppm_expf(-lin.y * ex));
    mapped.z = srgb(1.0f - ppm_expf(-lin.z * ex));
    mapped.w = 1.0f - ppm_expf(-lin.w * ex);
}

PPM_DEV unsigned char unorm8(float c) {
    c = c > 1.0f ? 1.0f : c;
    c = c > 0.0f ? c : 0.0f;
    return (unsigned char)(int)(c * 255.0f + 0.5f);
}

// progressive.fs:98-142 for one hit point, then the re.fs draw of that pixel: the gather of pass
// f, for wave `wave` (0..3) of pixel tile `tile`.  kGatherLagged: pass f is the pass before the one
// in progress (the gather blocks of the photon-stage launch), and the gather reads the uniforms
// pass f's k_pass_begin saved (GatherSnap); kGatherFlush: pass f is the last one begun (k_gather at
// the end of a pass sequence; every thread of the block calls it, since the draw's NumPhotons is a
// block-wide sum).  The gather reads pass f's hash table and photon half, which the photon stage
// running beside it does not write, and its hit points, which no eye stage writes before the
// gather is done (an eye stage writes the buffer the next eye pass reads).
PPM_DEV void gather(const Job& J, int tile, int wave, int mode) {
    const DevGlobals* g = J.g;
    GatherSnap sn;
    float np = 0.0f;
    const GatherParams& gp = J.gp;
    if (mode == kGatherLagged) {
        if (!g->lag_pending) return;
        sn = g->snap[g->lag_idx];
        np = sn.np;
    } else {
        if (!g->active) return;
        sn = g->snap[g->half];
        if (gp.draw) np = num_photons_now(g);
        // the pass is gathered: the next k_pass_begin leaves it out of the lagged gather
        if (tile == 0 && threadIdx.x == 0) J.g->gathered = g->frame;
    }
    const Tiles& tl = J.tl;
    const HitState& hs = J.hs;
    const PhotonState ps = photon_half(J.ps, J.pp.count, sn.half);
    const HashState& h = J.h;
    int px_i, px_j;
    tile_pixel(tl, tile, wave, px_i, px_j);
    if (px_i >= tl.width || px_j >= tl.height) return;
    const int p = px_j * tl.width + px_i;
    // the hit points, emission and grid of the latest eye pass at pass f
    const int cb = sn.cb;
    const GridScalars& gr = g->grid[cb];
    const size_t hp = (size_t)cb * (size_t)hs.q + (size_t)p;
    float4 tr = sn.first_pass ? make_float4(0.0f, 0.0f, 0.0f, gr.initial_radius) : hs.tau_r[p];
    float4 en = hs.e_n[p];
    if (sn.eye_now) {
        // the emission changes with the eye stage only: between eye passes e_n holds it already
        const float4 em = hs.em[hp];
        en.x = em.x;
        en.y = em.y;
        en.z = em.z;
    }
    if (hs.tri[hp] >= 0) {
        const f3 qpos = xyz(hs.pos[hp]);
        const f3 qnrm = xyz(hs.nrm[hp]);
        const f3 refl = xyz(hs.refl[hp]);
        const float R = tr.w;
        const float N = en.w;
        const f3 bbmin = mk3(gr.bbmin[0], gr.bbmin[1], gr.bbmin[2]);
        const float gs = gr.grid_scale, hsc = gr.hash_scale;
        const f3 d = qpos - bbmin;
        const f3 ii = mk3(floorf(fabsf(d.x) * gs - 0.5f), floorf(fabsf(d.y) * gs - 0.5f),
                          floorf(fabsf(d.z) * gs - 0.5f));
        f3 flux = mk3(0.0f, 0.0f, 0.0f);
        float m = 0.0f;
        const int clamp_slot = (h.height - 1) * h.width;
        const unsigned int* table = table_at(h, sn.table);
        // Each cell's box is (cell / GridScale + BBoxMin, (cell + 1) / GridScale + BBoxMin)
        // (progressive.fs:79-80); along an axis the two cells' boxes share their middle bound, so the
        // block of eight has three bounds per axis, computed once with the same operations.
        const f3 c1 = mk3(ii.x + 1.0f, ii.y + 1.0f, ii.z + 1.0f);
        const f3 c2 = mk3(c1.x + 1.0f, c1.y + 1.0f, c1.z + 1.0f);
        const f3 k0 = mk3(ii.x + 0.0f, ii.y + 0.0f, ii.z + 0.0f) / gs + bbmin;
        const f3 k1 = c1 / gs + bbmin;
        const f3 k2 = c2 / gs + bbmin;
        // A cell whose box lies at least R from the hit point holds no photon the test below can
        // accept, so it is neither hashed nor read, and the sums stay bit-identical.  The test
        // accepts a photon p strictly inside the box with length(p - q) < R.  Per axis, p > rmin
        // and p < rmax give |fl(p - q)| >= g = max(fl(rmin - q), fl(q - rmax), 0), since rounding
        // is monotone; squares, sums and sqrt of non-negative numbers are monotone too, so the
        // shader's length(p - q) = sqrt((dx dx + dy dy) + dz dz) is at least the same expression
        // of the gaps.  That expression >= R therefore rules out every photon of the cell.  (A NaN
        // anywhere makes the comparison false: nothing is skipped.)
        for (int iz = 0; iz <= 1; iz++)
            for (int iy = 0; iy <= 1; iy++)
                for (int ix = 0; ix <= 1; ix++) {
                    const f3 cell = mk3(ii.x + (float)ix, ii.y + (float)iy, ii.z + (float)iz);
                    const f3 rmin = mk3(ix ? k1.x : k0.x, iy ? k1.y : k0.y, iz ? k1.z : k0.z);
                    const f3 rmax = mk3(ix ? k2.x : k1.x, iy ? k2.y : k1.y, iz ? k2.z : k1.z);
                    const f3 gap = mk3(fmaxf(fmaxf(rmin.x - qpos.x, qpos.x - rmax.x), 0.0f),
                                       fmaxf(fmaxf(rmin.y - qpos.y, qpos.y - rmax.y), 0.0f),
                                       fmaxf(fmaxf(rmin.z - qpos.z, qpos.z - rmax.z), 0.0f));
                    if (length3(gap) >= R) continue;
                    const float sf = cell_hash(cell, gs, hsc, h.num_f);
                    // texel fetch clamps a slot past the last row to the start of that row; a
                    // NaN slot (zero-extent grid) clamps to texel (0, 0)
                    const int slot = sf < h.num_f ? (int)sf : (sf >= h.num_f ? clamp_slot : 0);
                    const uint2 rec = reinterpret_cast<const uint2*>(table)[slot];
                    const unsigned int cnt = rec.x;
                    if (cnt == 0u) continue;  // adds exactly +0 to flux and count
                    unsigned int rep = rec.y;
                    if (rep == 0xffffffffu) rep = 0u;  // cleared slot reads photon texel (0, 0)
                    const f3 ppos = xyz(ps.pos[rep]);
                    if (ppos.x > rmin.x && ppos.y > rmin.y && ppos.z > rmin.z && ppos.x < rmax.x &&
                        ppos.y < rmax.y && ppos.z < rmax.z) {
                        const f3 diff = ppos - qpos;
                        const f3 pdir = xyz(ps.dir[rep]);
                        if (length3(diff) < R && dot3(qnrm, -pdir) > 0.0f) {
                            const float corr = (float)cnt;
                            flux = flux + xyz(ps.flux[rep]) * corr;
                            m = m + corr;
                        }
                    }
                }
        flux = flux * (refl * 0.31830995240629591621063460818591f);
        const float gg = min1((N + gp.alpha * m) / (N + m), 1.0f);
        tr.w = R * sqrtf(gg);
        en.w = N + m * gp.alpha;
        tr.x = (tr.x + flux.x) * gg;
        tr.y = (tr.y + flux.y) * gg;
        tr.z = (tr.z + flux.z) * gg;
    }
    hs.tau_r[p] = tr;
    hs.e_n[p] = en;
    if (gp.draw) {
        float4 lin, mapped;
        radiance(tr, en, np, lin, mapped);
        hs.rgba8[p] = make_uchar4(unorm8(mapped.x), unorm8(mapped.y), unorm8(mapped.z), unorm8(mapped.w));
    }
}

// The photon-stage launch: trace blocks [0, trace_blocks) run the photon stage of the pass in
// progress; the blocks after them run the gather of the pass before, one wave per quarter of a
// pixel tile.  Being dispatched last, the gather waves fill the GPU while the last walks finish.
PPM_DEV void photon_launch(const Job& J, int blk) {
    if (blk < J.trace_blocks) {
        photon_stage(J, blk);
        return;
    }
    const int w = (blk - J.trace_blocks) * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (w < 4 * J.tile_blocks) gather(J, w >> 2, w & 3, kGatherLagged);
}

// k_gather: the gather of the last pass begun, in 256-thread tiles.
PPM_DEV void gather_flush(const Job& J, int blk) { gather(J, blk, (int)(threadIdx.x >> 6), kGatherFlush); }

}  // namespace

// ------------------------------------------------------------------------------------------
// Plain (one job) and batched entry points of every body.

#define PPM_JOB_KERNEL(name, body, threads)                                                       \
    __global__ __launch_bounds__(threads) void name(Job J) { body(J, (int)blockIdx.x); }          \
    __global__ __launch_bounds__(threads) void name##_batch(const Job* __restrict__ jobs,          \
                                                            const BlockRef* __restrict__ map) {    \
        const BlockRef m = map[blockIdx.x];                                                         \
        body(jobs[m.job], m.block);                                                                 \
    }
// A floor on waves per SIMD caps a kernel's VGPRs: the eye kernel's bounce loop at 6 waves (80,
// with a small spill; it would take 105 at 4), the photon walk at 7 (72; 8 waves would spill it),
// which shortens their dependent-load stalls (measured: eye pass 1.04 -> 0.99 ms on a 1024^2 image).
#define PPM_JOB_KERNEL_WAVES(name, body, threads, waves)                                          \
    __global__ __launch_bounds__(threads) __attribute__((amdgpu_waves_per_eu(waves))) void name(Job J) { \
        body(J, (int)blockIdx.x);                                                                   \
    }                                                                                               \
    __global__ __launch_bounds__(threads) __attribute__((amdgpu_waves_per_eu(waves))) void name##_batch( \
        const Job* __restrict__ jobs, const BlockRef* __restrict__ map) {                           \
        const BlockRef m = map[blockIdx.x];                                                         \
        body(jobs[m.job], m.block);                                                                 \
    }
__global__ __launch_bounds__(kBlock) void k_pass_begin(Job J, unsigned int pass) {
    begin_pass(J, (int)blockIdx.x, pass, 0);
}
__global__ __launch_bounds__(kBlock) void k_pass_begin_batch(const Job* __restrict__ jobs,
                                                             const BlockRef* __restrict__ map, unsigned int pass,
                                                             int defer_all) {
    const BlockRef m = map[blockIdx.x];
    begin_pass(jobs[m.job], m.block, pass, defer_all);
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_eye(Job J, int mode) {
    eye(J, (int)blockIdx.x, mode);
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_eye_batch(
    const Job* __restrict__ jobs, const BlockRef* __restrict__ map, int mode) {
    const BlockRef m = map[blockIdx.x];
    eye(jobs[m.job], m.block, mode);
}

__global__ void k_eye_begin(Job J, int mode, unsigned int pass, int lag) {
    if (threadIdx.x == 0) eye_begin(J, mode, pass, lag);
}
__global__ void k_eye_begin_batch(const Job* __restrict__ jobs, int n_jobs, int mode, unsigned int pass, int lag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_jobs) eye_begin(jobs[j], mode, pass, lag);
}

__global__ __launch_bounds__(1024) void k_grid(Job J, int mode) { grid(J, mode); }
__global__ __launch_bounds__(1024) void k_grid_batch(const Job* __restrict__ jobs, int mode) {
    grid(jobs[blockIdx.x], mode);
}
PPM_JOB_KERNEL_WAVES(k_photon_stage, photon_launch, kTraceBlockMax, 7)
PPM_JOB_KERNEL(k_hash, hash_deferred, kBlock)
PPM_JOB_KERNEL(k_gather, gather_flush, kBlock)


// (frame, passes_left) of every job of a batch, for the host's progress snapshot.
__global__ void k_batch_status(const Job* __restrict__ jobs, int n_jobs, unsigned long long* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_jobs) return;
    const DevGlobals* g = jobs[j].g;
    out[2 * j] = g->frame;
    out[2 * j + 1] = g->passes_left;
}

__global__ void k_resolve(int count, HitState hs, const DevGlobals* g, float4* linear, float4* mapped) {
    const float np = num_photons_now(g);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    float4 lin, mp;
    radiance(hs.tau_r[p], hs.e_n[p], np, lin, mp);
    linear[p] = lin;
    mapped[p] = mp;
    hs.rgba8[p] = make_uchar4(unorm8(mp.x), unorm8(mp.y), unorm8(mp.z), unorm8(mp.w));
}
