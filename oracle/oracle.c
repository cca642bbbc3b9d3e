/*
 * oracle.c -- CPU restatement of the glslppm SPPM pass loop.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker for the HIP implementation: only tests, the smoke check and the CPU
 * baseline of bench.py load it.  Nothing in the product links, imports or calls it.
 *
 * Each function restates one GLSL pass or host step of the reference statement by statement in
 * strict fp32 (built with -ffp-contract=off, no fast-math): division and sqrt are the IEEE
 * operations, min/max reject NaN, sin/cos/exp/log/pow follow the published Cephes single
 * precision routines (S. L. Moshier) so that a GPU build of the same steps gives the same bits.
 * Where the reference relies on GL fixed function, the oracle emulates it literally rather than
 * using the closed forms the HIP kernels use:
 *   - BVH walk over the canonical link table (the six hit/miss link tiles, gpurt.cpp:408-473),
 *     cube-face selection by the largest direction component;
 *   - scatter.vs points drawn in photon order against an fp32 depth buffer cleared to 1.0 with
 *     GL_LESS and clipping at |z| > 1 (glslppm.cpp:418-442);
 *   - correction.vs additive blend (glslppm.cpp:445-470);
 *   - texel addressing of the hash texture in progressive.fs with GL_CLAMP + GL_NEAREST.
 *
 * Parity pins (tests/): xorshift128 against an independent integer restatement; GPURnd against
 * exact integer Schrage arithmetic; the transcendentals against float64 references within a few
 * ulp; the BVH walk against brute-force closest hit; hash/scatter/correction against brute
 * force; the progressive update against its closed-form recurrence; re.fs against closed form.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ppm.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } v4;
typedef struct { float x, y, z; } v3;

static v3 V3(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v4 V4(float x, float y, float z, float w) { v4 r = {x, y, z, w}; return r; }
static v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 vscale(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vcross(v3 a, v3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static v3 vnormalize(v3 a) { return vscale(a, 1.0f / sqrtf(vdot(a, a))); }
static v3 vreflect(v3 i, v3 n) { return vsub(i, vscale(n, 2.0f * vdot(n, i))); }
static v3 vrefract(v3 i, v3 n, float eta) {
    float d = vdot(n, i);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return V3(0.0f, 0.0f, 0.0f);
    return vsub(vscale(i, eta), vscale(n, eta * d + sqrtf(k)));
}
static float gl_min(float a, float b) { return fminf(a, b); }
static float gl_max(float a, float b) { return fmaxf(a, b); }
static float gl_fract(float x) { return x - floorf(x); }
static float gl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static float gl_mod(float x, float y) { return x - y * floorf(x / y); }

static float bits_f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ------------------------------------------------------------------------------------------
 * Transcendentals after Cephes sinf.c, cosf.c, expf.c, logf.c.
 * ------------------------------------------------------------------------------------------ */

static float ldexp_steps(float p, int n) {
    if (n > 127) { p = p * 1.7014118346046923e38f; n -= 127; }
    if (n < -126) { p = p * 1.1754943508222875e-38f; n += 126; }
    if (n < -126) { p = p * 1.1754943508222875e-38f; n += 126; }
    return p * bits_f((uint32_t)(n + 127) << 23);
}

static float sin_poly(float x) {
    float z = x * x;
    float y = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x;
    return y + x;
}
static float cos_poly(float x) {
    float z = x * x;
    float y = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z;
    y = y - 0.5f * z;
    return y + 1.0f;
}
/* x >= 0: octant j and the reduced argument */
static float reduce_pio4(float x, int* jj) {
    int j = (int)(1.27323954473516f * x);
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    *jj = j;
    return ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
}
float oracle_sinf(float x) {
    int sign = 1, j;
    if (x < 0.0f) { x = -x; sign = -1; }
    float r = reduce_pio4(x, &j);
    if (j > 3) { sign = -sign; j -= 4; }
    float y = (j == 1 || j == 2) ? cos_poly(r) : sin_poly(r);
    return sign < 0 ? -y : y;
}
float oracle_cosf(float x) {
    int sign = 1, j;
    if (x < 0.0f) x = -x;
    float r = reduce_pio4(x, &j);
    if (j > 3) { j -= 4; sign = -sign; }
    if (j > 1) sign = -sign;
    float y = (j == 1 || j == 2) ? sin_poly(r) : cos_poly(r);
    return sign < 0 ? -y : y;
}
float oracle_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283905206835f) return bits_f(0x7f800000u);
    if (x < -103.278929903431851103f) return 0.0f;
    float z = floorf(1.44269504088896341f * x + 0.5f);
    x = x - z * 0.693359375f;
    x = x - z * -2.12194440e-4f;
    int n = (int)z;
    float zz = x * x;
    float p = ((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
               1.6666665459e-1f) * x + 5.0000001201e-1f;
    p = p * zz + x + 1.0f;
    return ldexp_steps(p, n);
}
float oracle_logf(float x) {
    if (x != x) return x;
    if (x <= 0.0f) return x == 0.0f ? bits_f(0xff800000u) : bits_f(0x7fc00000u);
    if (x == bits_f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.1754943508222875e-38f) { x = x * 33554432.0f; e = -25; }
    uint32_t b = f_bits(x);
    e += (int)((b >> 23) & 0xffu) - 126;
    x = bits_f((b & 0x807fffffu) | 0x3f000000u);
    if (x < 0.707106781186547524f) { e -= 1; x = x + x - 1.0f; }
    else x = x - 1.0f;
    float z = x * x;
    float y = ((((((((7.0376836292e-2f * x - 1.1514610310e-1f) * x + 1.1676998740e-1f) * x - 1.2420140846e-1f) * x +
                  1.4249322787e-1f) * x - 1.6668057665e-1f) * x + 2.0000714765e-1f) * x - 2.4999993993e-1f) * x +
               3.3333331174e-1f) * x * z;
    float fe = (float)e;
    y += -2.12194440e-4f * fe;
    y += -0.5f * z;
    z = x + y;
    z += 0.693359375f * fe;
    return z;
}
float oracle_powf(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : bits_f(0x7f800000u);
    if (x < 0.0f) return bits_f(0x7fc00000u);
    return oracle_expf(y * oracle_logf(x));
}

/* ------------------------------------------------------------------------------------------
 * GPURnd (eyeraytrace.fs:326-347) and hash() (hash.fs:32-48), vec4 arithmetic lane by lane.
 * ------------------------------------------------------------------------------------------ */

static const float Q_[4] = {1225.0f, 1585.0f, 2457.0f, 2098.0f};
static const float R_[4] = {1112.0f, 367.0f, 92.0f, 265.0f};
static const float A_[4] = {3423.0f, 2646.0f, 1707.0f, 1999.0f};
static const float M_[4] = {4194287.0f, 4194277.0f, 4194191.0f, 4194167.0f};

static void schrage(const float n[4], float out[4]) {
    for (int k = 0; k < 4; k++) {
        float q_ = 1.0f / Q_[k];
        float beta = floorf(n[k] * q_);
        float p = A_[k] * (n[k] - beta * Q_[k]) - beta * R_[k];
        float m_2 = 0.5f * M_[k];
        beta = gl_sign(-p) * m_2 + m_2;
        out[k] = p + beta;
    }
}
static float signed_sum(const float n[4]) {
    float v[4];
    for (int k = 0; k < 4; k++) v[k] = n[k] * (1.0f / M_[k]);
    /* dot(v, vec4(1, -1, 1, -1)) */
    return ((v[0] * 1.0f + v[1] * -1.0f) + v[2] * 1.0f) + v[3] * -1.0f;
}
float oracle_gpurnd(float* n) {
    float o[4];
    schrage(n, o);
    memcpy(n, o, sizeof(o));
    return gl_fract(signed_sum(n));
}
static float gpurnd4(v4* s) {
    float n[4] = {s->x, s->y, s->z, s->w};
    float r = oracle_gpurnd(n);
    *s = V4(n[0], n[1], n[2], n[3]);
    return r;
}
float oracle_hash(const float idx[3], float grid_scale, float hash_scale1, float hash_num) {
    float n[4] = {idx[0] * hash_scale1, idx[1] * hash_scale1, idx[2] * hash_scale1, (grid_scale * 0.5f) * hash_scale1};
    float o[4];
    schrage(n, o);
    return floorf(gl_fract(signed_sum(o)) * hash_num);
}

/* ------------------------------------------------------------------------------------------
 * XORShift (glslppm.cpp:77-91) and RandomizeTextures (glslppm.cpp:113-140).
 * ------------------------------------------------------------------------------------------ */

static float xorshift_frand(uint32_t* s) {
    uint32_t t = s[0] ^ (s[0] << 11);
    s[0] = s[1]; s[1] = s[2]; s[2] = s[3];
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    return (float)s[3] * (1.0f / 4294967295.0f);
}
/* the texel constructor's arguments are evaluated last to first by g++: w, z, y, x */
void oracle_seed_grid(uint32_t* state, float* out, size_t texels) {
    for (size_t i = 0; i < texels; i++) {
        float w = (float)((double)xorshift_frand(state) * 4194304.0);
        float z = (float)((double)xorshift_frand(state) * 4194304.0);
        float y = (float)((double)xorshift_frand(state) * 4194304.0);
        float x = (float)((double)xorshift_frand(state) * 4194304.0);
        out[4 * i + 0] = x; out[4 * i + 1] = y; out[4 * i + 2] = z; out[4 * i + 3] = w;
    }
}
float oracle_xorshift_frand(uint32_t* state) { return xorshift_frand(state); }

/* ------------------------------------------------------------------------------------------
 * Scene access in the reference's terms.
 * ------------------------------------------------------------------------------------------ */

typedef struct {
    const ppm_scene* s;
} Scene;

static v3 vertex(const ppm_scene* s, int tri, int k) {
    const float* p = s->positions + (size_t)tri * 9 + k * 3;
    return V3(p[0], p[1], p[2]);
}
/* textureTriangles stores (n.x, n.y, sgn(n.z) (mat + 1)) with sgn(v) = v > 0 ? 1 : -1
 * (gpurt.cpp:103-113, 151-158); the shaders rebuild n.z = sign(.) sqrt(|1 - x^2 - y^2|). */
static v3 normal_rebuilt(const ppm_scene* s, int tri, int k) {
    const float* n = s->normals + (size_t)tri * 9 + k * 3;
    float packed = (n[2] > 0.0f ? 1.0f : -1.0f) * (float)(s->material[tri] + 1);
    float z = gl_sign(packed) * sqrtf(fabsf(1.0f - n[0] * n[0] - n[1] * n[1]));
    return V3(n[0], n[1], z);
}

/* Texture coordinate k (0: u, 1: v) of a triangle vertex: texcoords or, without any, 1e30
 * (mesh.cpp:664-668). */
static float texcoord(const ppm_scene* s, int tri, int vert, int k) {
    return s->texcoords ? s->texcoords[(size_t)tri * 6 + vert * 2 + k] : 1.0e30f;
}

/* texture3D of the RGB8 volume (gpurt.cpp:158-198), GL_LINEAR and GL_REPEAT on all three axes
 * (GL specification, linear filtering: i0 = wrap(floor(u - 1/2)), weight frac(u - 1/2)); the
 * restatement interpolates along u, then v, then r, in fp32.  Coordinates too large for an
 * exact floor, or NaN, read texel 0 with weight 0.  Pinned against a float64 evaluation of the
 * same formula and against texel centres in tests/test_oracle_texture.py. */
static float tex_axis(float coord, int size, int* i0, int* i1) {
    float a = coord * (float)size - 0.5f;
    if (!(fabsf(a) <= 1.0e18f)) {
        *i0 = 0;
        *i1 = size > 1 ? 1 : 0;
        return 0.0f;
    }
    float fl = floorf(a);
    long long k = (long long)fl % (long long)size;
    if (k < 0) k += size;
    *i0 = (int)k;
    *i1 = *i0 + 1 == size ? 0 : *i0 + 1;
    return a - fl;
}

static v3 texel(const ppm_scene* s, int i, int j, int k) {
    const size_t S = (size_t)s->texture_size;
    const uint8_t* c = s->textures + 3 * (((size_t)k * S + (size_t)j) * S + (size_t)i);
    return V3((float)c[0] / 255.0f, (float)c[1] / 255.0f, (float)c[2] / 255.0f);
}

static v3 lerp3(v3 p, v3 q, float w) {
    float r = 1.0f - w;
    return V3(p.x * r + q.x * w, p.y * r + q.y * w, p.z * r + q.z * w);
}

v3 oracle_texture3d(const ppm_scene* s, float u, float v, float r) {
    int i0, i1, j0, j1, k0, k1;
    float a = tex_axis(u, s->texture_size, &i0, &i1);
    float b = tex_axis(v, s->texture_size, &j0, &j1);
    float c = tex_axis(r, s->num_materials, &k0, &k1);
    v3 y00 = lerp3(texel(s, i0, j0, k0), texel(s, i1, j0, k0), a);
    v3 y10 = lerp3(texel(s, i0, j1, k0), texel(s, i1, j1, k0), a);
    v3 y01 = lerp3(texel(s, i0, j0, k1), texel(s, i1, j0, k1), a);
    v3 y11 = lerp3(texel(s, i0, j1, k1), texel(s, i1, j1, k1), a);
    return lerp3(lerp3(y00, y10, b), lerp3(y01, y11, b), c);
}

/* unit entry point for the texture pins */
void oracle_texture_sample(const ppm_scene* s, float u, float v, float r, float* out) {
    v3 c = oracle_texture3d(s, u, v, r);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

static int textured(const ppm_scene* s) { return s->texcoords && s->textures && s->texture_size > 0; }

/* MaterialNumRcp (glslppm.cpp:256, 376) */
static float material_rcp(const ppm_scene* s) { return 1.0f / (float)s->num_materials; }

typedef struct {
    float t;
    v3 pos, nrm, gnrm, col;
    float tex_u, tex_v; /* Result.tex */
    int brdf;
    float eta, g;
    int matid;
    /* LastIntersection: triangle (-1 on a miss), material, t */
    int isect_tri;
    float isect_mat;
    float isect_t;
} Isect;

static int cube_face(v3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return d.x >= 0.0f ? 0 : 1;
    if (ay >= az) return d.y >= 0.0f ? 2 : 3;
    return d.z >= 0.0f ? 4 : 5;
}

/* raytrace() of eyeraytrace.fs:67-290 (photon = 0) and photontrace.fs:123-255 (photon = 1) */
static Isect raytrace(const ppm_scene* s, v3 org, v3 dir, int cull, int photon) {
    Isect R;
    memset(&R, 0, sizeof(R));
    R.t = 1.0e30f;
    R.nrm = V3(0, 0, 0);
    R.pos = V3(0, 0, 0);
    R.col = V3(0, 0, 0);
    R.gnrm = V3(0, 0, 0);
    R.brdf = 0;
    int isect_tri = -1;
    float isect_mat = -1.0f;
    float isect_w = photon ? 1.0e30f : 1.0e20f;
    v3 bary = V3(0, 0, 0);
    const int face = cube_face(dir);
    const int32_t* links = s->node_links + (size_t)face * s->num_nodes * 2;
    v3 rcp = V3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    v3 t0 = vmul(vneg(rcp), org);
    int node = 0;
    for (;;) {
        const float* b = s->node_bounds + (size_t)node * 6;
        v3 lo = vadd(vmul(rcp, V3(b[0], b[1], b[2])), t0);
        v3 hi = vadd(vmul(rcp, V3(b[3], b[4], b[5])), t0);
        v3 a = V3(gl_min(lo.x, hi.x), gl_min(lo.y, hi.y), gl_min(lo.z, hi.z));
        v3 c = V3(gl_max(lo.x, hi.x), gl_max(lo.y, hi.y), gl_max(lo.z, hi.z));
        float tmin = gl_max(gl_max(a.x, a.y), a.z);
        float tmax = gl_min(gl_min(c.x, c.y), c.z);
        int bbhit = (tmin <= tmax) && (tmin <= isect_w) && (tmax >= 0.0f);
        int tri = s->node_triangle[node];
        if (tri >= 0 && bbhit) {
            v3 V0 = vertex(s, tri, 0), V1 = vertex(s, tri, 1), V2 = vertex(s, tri, 2);
            int mat = s->material[tri];
            int brdf = s->materials[mat].brdf;
            v3 e0 = vsub(V1, V0), e1 = vsub(V2, V0);
            v3 pv = vcross(dir, e1);
            float det = vdot(e0, pv);
            if ((cull && (det > 1e-10f)) || !cull) {
                v3 tv = vsub(org, V0);
                v3 qv = vcross(tv, e0);
                float ux = vdot(tv, pv), uy = vdot(dir, qv), uz = vdot(e1, qv);
                ux = ux / det; uy = uy / det; uz = uz / det;
                float uw = 1.0f - ux - uy;
                int ok = ux >= 0.0f && uy >= 0.0f && uz >= 0.0f && uw >= 0.0f && uz < isect_w;
                if (photon) ok = ok && brdf != -1;
                if (ok) {
                    bary = V3(uy, uw, ux);
                    isect_tri = tri;
                    isect_mat = (float)mat;
                    isect_w = uz;
                }
            }
        }
        int next = bbhit ? links[2 * node] : links[2 * node + 1];
        if (next < 0) break;
        node = next;
    }
    if (isect_tri >= 0) {
        int tri = isect_tri;
        v3 V0 = vertex(s, tri, 0), V1 = vertex(s, tri, 1), V2 = vertex(s, tri, 2);
        v3 N0 = normal_rebuilt(s, tri, 0), N1 = normal_rebuilt(s, tri, 1), N2 = normal_rebuilt(s, tri, 2);
        R.t = isect_w;
        R.pos = vadd(vadd(vscale(V2, bary.x), vscale(V0, bary.y)), vscale(V1, bary.z));
        R.nrm = vnormalize(vadd(vadd(vscale(N2, bary.x), vscale(N0, bary.y)), vscale(N1, bary.z)));
        R.gnrm = vnormalize(vcross(vsub(V1, V0), vsub(V2, V0)));
        /* Result.tex = T2 * Barycentric.x + T0 * Barycentric.y + T1 * Barycentric.z */
        R.tex_u = (texcoord(s, tri, 2, 0) * bary.x + texcoord(s, tri, 0, 0) * bary.y) + texcoord(s, tri, 1, 0) * bary.z;
        R.tex_v = (texcoord(s, tri, 2, 1) * bary.x + texcoord(s, tri, 0, 1) * bary.y) + texcoord(s, tri, 1, 1) * bary.z;
        const ppm_material* m = &s->materials[(int)isect_mat];
        R.col = V3(m->color[0], m->color[1], m->color[2]);
        R.brdf = m->brdf;
        R.g = m->g;
        R.eta = m->eta;
        R.matid = (int)isect_mat;
    }
    R.isect_tri = isect_tri;
    R.isect_mat = isect_mat;
    R.isect_t = isect_w;
    return R;
}

/* unit entry point for the traversal pin: returns the triangle, writes t and barycentrics */
int oracle_trace(const ppm_scene* s, const float org[3], const float dir[3], int cull, int photon, float* t_out,
                 float* bary_out) {
    Isect r = raytrace(s, V3(org[0], org[1], org[2]), V3(dir[0], dir[1], dir[2]), cull, photon);
    if (t_out) *t_out = r.isect_t;
    if (bary_out && r.isect_tri >= 0) {
        /* recompute is not needed: expose position instead */
        bary_out[0] = r.pos.x; bary_out[1] = r.pos.y; bary_out[2] = r.pos.z;
    }
    return r.isect_tri;
}

static v3 onb(v3 x, v3 n) {
    if (n.z < -0.9999999f) {
        v3 u = V3(x.z, x.x, 0.0f);
        return vsub(vscale(n, x.y), u);
    }
    float a = 1.0f / (1.0f + n.z);
    float b = -n.x * n.y * a;
    v3 u = V3(1.0f - n.x * n.x * a, b, -n.x);
    v3 w = V3(b, 1.0f - n.y * n.y * a, -n.y);
    return vadd(vadd(vscale(u, x.x), vscale(n, x.y)), vscale(w, x.z));
}

static float fresnel(v3 incom, v3 normal, float ni, float ne) {
    float eta = ni / ne;
    float c1 = vdot(incom, normal);
    float c2 = 1.0f - (eta * eta) * (1.0f - c1 * c1);
    if (c2 < 0.0f) return 1.0f;
    c2 = sqrtf(c2);
    float rs = (ni * c1 - ne * c2) / (ni * c1 + ne * c2);
    float rp = (ni * c2 - ne * c1) / (ni * c2 + ne * c1);
    return (rs * rs + rp * rp) * 0.5f;
}

static v3 glossy_reflect(v3 d, v3 n, float g, v4* rndv) {
    v3 r = vnormalize(vadd(vscale(vreflect(d, n), 1.0f - g - g), vscale(n, g + g)));
    float rnd1 = gpurnd4(rndv);
    float rnd2 = gpurnd4(rndv);
    float temp1 = 6.283184f * rnd1;
    float temp2_ = oracle_powf(rnd2, g);
    float temp2 = sqrtf(1.0f - temp2_ * temp2_);
    v3 v = V3(oracle_sinf(temp1) * temp2, temp2_, oracle_cosf(temp1) * temp2);
    v3 result = vnormalize(onb(v, r));
    if (vdot(result, n) < 0.0f) result = vreflect(result, n);
    return result;
}

/* ------------------------------------------------------------------------------------------
 * Pass state: the reference's textures.
 * ------------------------------------------------------------------------------------------ */

typedef struct {
    const ppm_scene* s;
    ppm_params prm;
    int Q, P, H, Qw, Qh, Pw, Ph, Hw, Hh;
    uint32_t xs[4];
    uint64_t frame, eye_samples, photons_emitted;
    float num_photons;
    /* hit points */
    v4 *qpos, *qrefl, *qnrm, *qrng, *qflux_r, *qemi_n;
    int* qtri;
    float *qmat, *qt;
    /* photons */
    v4 *ppos, *pflux, *pdir, *prng;
    int* ptri;
    int* pbounce;
    int* pemitted;
    int* pslot;  /* PhotonIndexTexture.zw as a linear slot, -1 outside the viewport */
    /* hash */
    int* hrep;   /* photon index in the colour of the hash texture, -1 = cleared */
    float* hdepth;
    float* hcount;
    /* grid */
    float bbmin[3], bbmax[3], ir, grid_scale, hash_scale;
} Oracle;

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

void oracle_destroy(void* h) {
    Oracle* o = (Oracle*)h;
    if (!o) return;
    free(o->qpos); free(o->qrefl); free(o->qnrm); free(o->qrng); free(o->qflux_r); free(o->qemi_n);
    free(o->qtri); free(o->qmat); free(o->qt);
    free(o->ppos); free(o->pflux); free(o->pdir); free(o->prng); free(o->ptri); free(o->pbounce);
    free(o->pemitted); free(o->pslot);
    free(o->hrep); free(o->hdepth); free(o->hcount);
    free(o);
}

int oracle_create(const ppm_scene* s, const ppm_params* p, void** out) {
    if (!s || !p || !out) return PPM_E_INVAL;
    Oracle* o = (Oracle*)calloc(1, sizeof(Oracle));
    if (!o) return PPM_E_OOM;
    o->s = s;
    o->prm = *p;
    o->Qw = p->image_width; o->Qh = p->image_height;
    o->Pw = p->photon_width; o->Ph = p->photon_height;
    o->Hw = p->hash_width; o->Hh = p->hash_height;
    o->Q = o->Qw * o->Qh; o->P = o->Pw * o->Ph; o->H = o->Hw * o->Hh;
    memcpy(o->xs, p->xorshift, sizeof(o->xs));
    size_t Q = o->Q, P = o->P, H = o->H;
    o->qpos = xcalloc(Q, sizeof(v4)); o->qrefl = xcalloc(Q, sizeof(v4)); o->qnrm = xcalloc(Q, sizeof(v4));
    o->qrng = xcalloc(Q, sizeof(v4)); o->qflux_r = xcalloc(Q, sizeof(v4)); o->qemi_n = xcalloc(Q, sizeof(v4));
    o->qtri = xcalloc(Q, sizeof(int)); o->qmat = xcalloc(Q, sizeof(float)); o->qt = xcalloc(Q, sizeof(float));
    o->ppos = xcalloc(P, sizeof(v4)); o->pflux = xcalloc(P, sizeof(v4)); o->pdir = xcalloc(P, sizeof(v4));
    o->prng = xcalloc(P, sizeof(v4)); o->ptri = xcalloc(P, sizeof(int)); o->pbounce = xcalloc(P, sizeof(int));
    o->pemitted = xcalloc(P, sizeof(int)); o->pslot = xcalloc(P, sizeof(int));
    o->hrep = xcalloc(H, sizeof(int)); o->hdepth = xcalloc(H, sizeof(float)); o->hcount = xcalloc(H, sizeof(float));
    if (!o->qpos || !o->qrefl || !o->qnrm || !o->qrng || !o->qflux_r || !o->qemi_n || !o->qtri || !o->qmat ||
        !o->qt || !o->ppos || !o->pflux || !o->pdir || !o->prng || !o->ptri || !o->pbounce || !o->pemitted ||
        !o->pslot || !o->hrep || !o->hdepth || !o->hcount) {
        oracle_destroy(o);
        return PPM_E_OOM;
    }
    oracle_seed_grid(o->xs, (float*)o->qrng, Q);
    oracle_seed_grid(o->xs, (float*)o->prng, P);
    for (size_t i = 0; i < H; i++) { o->hrep[i] = -1; o->hcount[i] = 0.0f; }
    for (size_t i = 0; i < P; i++) o->pslot[i] = -1;
    *out = o;
    return PPM_OK;
}

/* ---------------- eyeraytrace.fs main() ---------------- */
/* eyeraytrace.fs:396-420: the camera ray of pixel (i, j) with the pass's AA offset, through the
 * thin lens (two draws even at aperture 0), and the tent filter weight. */
static void camera_ray(const ppm_scene* s, const ppm_params* prm, int i, int j, float aax, float aay, v4* rndv,
                       v3* org, v3* dir, float* filter) {
    const ppm_camera* c = &s->camera;
    v3 C = V3(c->origin[0], c->origin[1], c->origin[2]);
    v3 U = V3(c->u[0], c->u[1], c->u[2]), V = V3(c->v[0], c->v[1], c->v[2]), W = V3(c->w[0], c->w[1], c->w[2]);
    float fcx = (float)i + 0.5f, fcy = (float)j + 0.5f;
    float ppx = (fcx + aax) - c->half_width;
    float ppy = (fcy + aay) - c->half_height;
    v3 rel = vadd(vadd(vscale(U, ppx), vscale(V, ppy)), vscale(W, c->distance));
    v3 rdir = vnormalize(rel);
    *filter = (1.0f - fabsf(aax) / 1.5f) * (1.0f - fabsf(aay) / 1.5f);
    v3 fp = vadd(C, vscale(rdir, prm->focal_length));
    float radius = sqrtf(gpurnd4(rndv)) * prm->aperture;
    float theta = 6.283184f * gpurnd4(rndv);
    v3 lens = vadd(C, vscale(vadd(vscale(U, oracle_cosf(theta)), vscale(V, oracle_sinf(theta))), radius));
    *dir = vnormalize(vsub(fp, lens));
    *org = lens;
}

/* One eye path (eyeraytrace.fs:396-547): the camera ray, then up to max_path_length segments
 * through the specular BRDFs; the outputs are what main() writes to the hit-point textures
 * (eyeraytrace.fs:548-553) before the emission average. */
typedef struct {
    v3 pos, col, nrm, dir, emi;
    int tri;
    float mat, t;
} EyePath;

static void eye_path(const ppm_scene* s, const ppm_params* prm, int i, int j, float aax, float aay, v4* rng,
                     EyePath* out) {
    v4 rndv = *rng;
    v3 rorg, rdir;
    float filter;
    camera_ray(s, prm, i, j, aax, aay, &rndv, &rorg, &rdir, &filter);
    v3 col = vscale(V3(1.0f, 1.0f, 1.0f), filter);
    v3 nrm = V3(0, 0, 0), pos = V3(0, 0, 0), emi = V3(0, 0, 0);
    const float eps = 1e-5f;
    int last_tri = -1;
    float last_mat = -1.0f, last_t = 1.0e20f;
    for (int k = 0; k < prm->max_path_length; k++) {
        Isect I = raytrace(s, rorg, rdir, k == 0, 0);
        last_tri = I.isect_tri; last_mat = I.isect_mat; last_t = I.isect_t;
        if (I.t == 1.0e30f) {
            nrm = V3(0, 0, 0);
            if (s->light_area == 0.0f) emi = vmul(V3(0.7f, 0.7f, 1.0f), col);
            break;
        }
        pos = I.pos;
        nrm = I.nrm;
        /* eyeraytrace.fs:443-446 */
        if (textured(s) && fabsf(I.tex_u) < 1e+10f)
            I.col = vmul(I.col, oracle_texture3d(s, I.tex_u, I.tex_v, material_rcp(s) * ((float)I.matid + 0.5f)));
        if (I.brdf == -1) {
            nrm = V3(0, 0, 0);
            emi = vmul(col, I.col);
            break;
        } else if (I.brdf == 0) {
            col = vmul(col, I.col);
            break;
        } else if (I.brdf == 1 || I.brdf == 4) {
            rorg = vadd(pos, vscale(I.gnrm, eps));
            if (I.brdf == 4) nrm = glossy_reflect(nrm, nrm, I.g, &rndv);
            rdir = vreflect(rdir, nrm);
            if (vdot(rdir, I.gnrm) < 0.0f) rdir = vneg(rdir);
            col = vmul(col, I.col);
        } else if (I.brdf == 2 || I.brdf == 5) {
            if (I.brdf == 5) nrm = glossy_reflect(nrm, nrm, I.g, &rndv);
            float ln = vdot(nrm, rdir);
            float eta = I.eta;
            if (ln < 0.0f) {
                float Re = fresnel(vneg(rdir), nrm, 1.0f, eta);
                if (gpurnd4(&rndv) < Re) {
                    rorg = vadd(pos, vscale(I.gnrm, eps));
                    rdir = vreflect(rdir, nrm);
                    if (vdot(rdir, I.gnrm) < 0.0f) rdir = vneg(rdir);
                } else {
                    col = vmul(col, I.col);
                    rorg = vsub(pos, vscale(I.gnrm, eps));
                    rdir = vrefract(rdir, nrm, 1.0f / eta);
                    if (vdot(rdir, I.gnrm) > 0.0f) rdir = vneg(rdir);
                }
            } else {
                float Re = fresnel(vneg(rdir), vneg(nrm), eta, 1.0f);
                col = vmul(col, I.col);
                if (gpurnd4(&rndv) < Re) {
                    rorg = vsub(pos, vscale(I.gnrm, eps));
                    rdir = vreflect(rdir, vneg(nrm));
                    if (vdot(rdir, vneg(I.gnrm)) < 0.0f) rdir = vneg(rdir);
                } else {
                    rorg = vadd(pos, vscale(I.gnrm, eps));
                    rdir = vrefract(rdir, vneg(nrm), eta);
                    if (vdot(rdir, vneg(I.gnrm)) > 0.0f) rdir = vneg(rdir);
                }
            }
        } else if (I.brdf == 3 || I.brdf == 6) {
            if (I.brdf == 6) nrm = glossy_reflect(nrm, nrm, I.g, &rndv);
            float Re = fresnel(vneg(rdir), nrm, 1.0f, I.eta);
            if (gpurnd4(&rndv) < Re) {
                rorg = vadd(pos, vscale(I.gnrm, eps));
                rdir = vreflect(rdir, nrm);
                if (vdot(rdir, I.gnrm) < 0.0f) rdir = vneg(rdir);
            } else {
                col = vmul(col, I.col);
                break;
            }
        }
    }
    *rng = rndv;
    out->pos = pos;
    out->col = col;
    out->nrm = nrm;
    out->dir = rdir;
    out->emi = emi;
    out->tri = last_tri;
    out->mat = last_tri < 0 ? -1.0f : last_mat;
    out->t = last_t;
}

static void eye_fragment(Oracle* o, int i, int j, float aax, float aay, int num_eye_samples) {
    const int px = j * o->Qw + i;
    v4 rndv = o->qrng[px];
    EyePath e;
    eye_path(o->s, &o->prm, i, j, aax, aay, &rndv, &e);
    v4 en = o->qemi_n[px];
    float a = 1.0f / (float)num_eye_samples;
    en.x = en.x * (1.0f - a) + e.emi.x * a;
    en.y = en.y * (1.0f - a) + e.emi.y * a;
    en.z = en.z * (1.0f - a) + e.emi.z * a;
    o->qpos[px] = V4(e.pos.x, e.pos.y, e.pos.z, e.dir.x);
    o->qrefl[px] = V4(e.col.x, e.col.y, e.col.z, e.dir.y);
    o->qnrm[px] = V4(e.nrm.x, e.nrm.y, e.nrm.z, e.dir.z);
    o->qrng[px] = rndv;
    o->qtri[px] = e.tri;
    o->qmat[px] = e.mat;
    o->qt[px] = e.t;
    o->qemi_n[px] = en;
}

/* ---------------- photontrace.fs ---------------- */
static int light_sample(const ppm_scene* s, v4* rndv, v3* flux, v3* dir, v3* org) {
    float Rnd = gpurnd4(rndv);
    int low = 0;
    float stride = 1.0f / (float)s->num_lights;
    int high = (int)(1.0f / stride + 0.5f);
    while (low < high) {
        int mid = low + ((high - low) / 2);
        /* texel at ((mid - 0.5) * stride, 0) with GL_CLAMP + GL_NEAREST */
        float u = ((float)mid - 0.5f) * stride;
        if (u < 0.0f) u = 0.0f;
        int texel = (int)floorf(u * (float)s->num_lights);
        if (texel > s->num_lights - 1) texel = s->num_lights - 1;
        if (s->light_cdf[texel] < Rnd) low = mid + 1;
        else high = mid;
    }
    float u = ((float)high - 0.5f) * stride;
    if (u < 0.0f) u = 0.0f;
    int texel = (int)floorf(u * (float)s->num_lights);
    if (texel > s->num_lights - 1) texel = s->num_lights - 1;
    int tri = s->light_triangle[texel];
    float rndv2 = gpurnd4(rndv);
    float t = sqrtf(1.0f - rndv2);
    float sv = gpurnd4(rndv);
    float a = 1.0f - t;
    float c = sv * t;
    float b = t - c;
    v3 V0 = vertex(s, tri, 0), V1 = vertex(s, tri, 1), V2 = vertex(s, tri, 2);
    v3 N0 = normal_rebuilt(s, tri, 0), N1 = normal_rebuilt(s, tri, 1), N2 = normal_rebuilt(s, tri, 2);
    const ppm_material* m = &s->materials[s->material[tri]];
    *flux = V3(m->color[0], m->color[1], m->color[2]);
    float le = s->light_area * 3.141592f;
    *flux = vmul(*flux, V3(le, le, le));
    /* photontrace.fs:413-422: T = T2 * c + T0 * a + T1 * b */
    if (textured(s)) {
        float tu = (texcoord(s, tri, 2, 0) * c + texcoord(s, tri, 0, 0) * a) + texcoord(s, tri, 1, 0) * b;
        float tv = (texcoord(s, tri, 2, 1) * c + texcoord(s, tri, 0, 1) * a) + texcoord(s, tri, 1, 1) * b;
        if (fabsf(tu) < 1e+10f)
            *flux = vmul(*flux, oracle_texture3d(s, tu, tv, material_rcp(s) * ((float)s->material[tri] + 0.75f)));
    }
    v3 o = vadd(vadd(vscale(V0, a), vscale(V1, b)), vscale(V2, c));
    v3 nrm = vnormalize(vadd(vadd(vscale(N0, a), vscale(N1, b)), vscale(N2, c)));
    v3 gnrm = vnormalize(vcross(vsub(V1, V0), vsub(V2, V0)));
    o = vadd(o, vscale(gnrm, 1.0e-5f));
    float rx = gpurnd4(rndv);
    float ry = gpurnd4(rndv);
    rx = 6.283184f * rx;
    ry = sqrtf(ry);
    *dir = onb(V3(oracle_sinf(rx) * ry, sqrtf(1.0f - ry * ry), oracle_cosf(rx) * ry), nrm);
    *org = o;
    return tri;
}

static void ibl_sample(const ppm_scene* s, v4* rndv, v3* flux, v3* dir, v3* org) {
    const float* bs = s->bsphere;
    float area = 12.566368f * bs[3] * bs[3];
    float temp2 = 6.283184f * gpurnd4(rndv);
    float aaa = sqrtf(1.0f - gpurnd4(rndv));
    float sinA = oracle_sinf(aaa);
    float sinT1 = (sinA + sinA) * oracle_cosf(aaa);
    v3 d;
    d.x = sinT1 * oracle_cosf(temp2);
    d.y = 1.0f - (sinA + sinA) * sinA;
    d.z = sinT1 * oracle_sinf(temp2);
    if (gpurnd4(rndv) > 0.5f) {
        d = V3(1.0f, 1.0f, 1.0f);
        d.y += d.y;
        d.z = -d.z;
        d = vnormalize(d);
        *flux = vscale(V3(1.0f, 1.0f, 0.7f), area);
    } else {
        *flux = vscale(V3(0.7f, 0.7f, 1.0f), area);
    }
    float radius = sqrtf(gpurnd4(rndv));
    float theta = 6.283184f * gpurnd4(rndv);
    v3 temp = V3(radius * oracle_cosf(theta), 0.0f, radius * oracle_sinf(theta));
    *org = vadd(vadd(onb(vscale(temp, bs[3]), d), vscale(d, bs[3])), V3(bs[0], bs[1], bs[2]));
    *dir = vneg(d);
}

/* photontrace.fs:487-686: the bounce of a stored photon off triangle tri at ipos (shading normal
 * pnrm, arriving along pd): BRDF sampling, then Russian roulette on max(col).  Returns 0 when the
 * photon ends; otherwise sets the new ray and scales the flux. */
static int photon_bounce(const ppm_scene* s, int tri, v3 ipos, v3 pnrm, v3 pd, v4* rndv, v3* flux, v3* rorg,
                         v3* rdir) {
    v3 V0 = vertex(s, tri, 0), V1 = vertex(s, tri, 1), V2 = vertex(s, tri, 2);
    v3 gnrm = vnormalize(vcross(vsub(V1, V0), vsub(V2, V0)));
    const ppm_material* m = &s->materials[s->material[tri]];
    v3 icol = V3(m->color[0], m->color[1], m->color[2]);
    int brdf = m->brdf;
    float eta = m->eta, g = m->g;
    v3 inrm = pnrm;
    const float eps = 1e-6f;
    /* photontrace.fs:519-544, 562-565: barycentrics from sub-triangle areas, then i.tex */
    if (textured(s)) {
        float bx = vlength(vcross(vsub(V1, V0), vsub(ipos, V0)));
        float by = vlength(vcross(vsub(V2, V1), vsub(ipos, V1)));
        float bz = vlength(vcross(vsub(V0, V2), vsub(ipos, V2)));
        float bw = (bx + by) + bz;
        float wx = bx / bw, wy = by / bw, wz = bz / bw;
        float tu = (texcoord(s, tri, 2, 0) * wx + texcoord(s, tri, 0, 0) * wy) + texcoord(s, tri, 1, 0) * wz;
        float tv = (texcoord(s, tri, 2, 1) * wx + texcoord(s, tri, 0, 1) * wy) + texcoord(s, tri, 1, 1) * wz;
        if (fabsf(tu) < 1e+10f)
            icol = vmul(icol, oracle_texture3d(s, tu, tv,
                                               material_rcp(s) * (((float)s->material[tri] + 0.5f) + 0.25f)));
    }
    if (brdf == 0) {
        float r0 = 6.283184f * gpurnd4(rndv);
        float r1 = sqrtf(gpurnd4(rndv));
        v3 v = V3(oracle_sinf(r0) * r1, sqrtf(1.0f - r1 * r1), oracle_cosf(r0) * r1);
        *rorg = vadd(ipos, vscale(gnrm, eps));
        *rdir = onb(v, inrm);
        if (vdot(*rdir, gnrm) < 0.0f) *rdir = vneg(*rdir);
    } else if (brdf == 1 || brdf == 4) {
        *rorg = vadd(ipos, vscale(gnrm, eps));
        if (brdf == 4) inrm = glossy_reflect(inrm, inrm, g, rndv);
        *rdir = vreflect(pd, inrm);
        if (vdot(*rdir, gnrm) < 0.0f) *rdir = vneg(*rdir);
    } else if (brdf == 2 || brdf == 5) {
        if (brdf == 5) inrm = glossy_reflect(inrm, inrm, g, rndv);
        float ln = vdot(inrm, pd);
        if (ln < 0.0f) {
            float Re = fresnel(vneg(pd), inrm, 1.0f, eta);
            if (gpurnd4(rndv) < Re) {
                *rorg = vadd(ipos, vscale(gnrm, eps));
                *rdir = vreflect(pd, inrm);
                icol = V3(1.0f, 1.0f, 1.0f);
                if (vdot(*rdir, gnrm) < 0.0f) *rdir = vneg(*rdir);
            } else {
                *rorg = vsub(ipos, vscale(gnrm, eps));
                *rdir = vrefract(pd, inrm, 1.0f / eta);
                if (vdot(*rdir, gnrm) > 0.0f) *rdir = vneg(*rdir);
            }
        } else {
            float Re = fresnel(vneg(pd), vneg(inrm), eta, 1.0f);
            if (gpurnd4(rndv) < Re) {
                *rorg = vsub(ipos, vscale(gnrm, eps));
                *rdir = vreflect(pd, vneg(inrm));
                if (vdot(*rdir, vneg(gnrm)) < 0.0f) *rdir = vneg(*rdir);
            } else {
                *rorg = vadd(ipos, vscale(gnrm, eps));
                *rdir = vrefract(pd, vneg(inrm), eta);
                if (vdot(*rdir, vneg(gnrm)) > 0.0f) *rdir = vneg(*rdir);
            }
        }
    } else if (brdf == 3 || brdf == 6) {
        if (brdf == 6) inrm = glossy_reflect(inrm, inrm, g, rndv);
        float Re = fresnel(vneg(pd), inrm, 1.0f, eta);
        if (gpurnd4(rndv) < Re) {
            *rorg = vadd(ipos, vscale(gnrm, eps));
            *rdir = vreflect(pd, inrm);
            icol = V3(1.0f, 1.0f, 1.0f);
            if (vdot(*rdir, gnrm) < 0.0f) *rdir = vneg(*rdir);
        } else {
            float r0 = 6.283184f * gpurnd4(rndv);
            float r1 = sqrtf(gpurnd4(rndv));
            v3 v = V3(oracle_sinf(r0) * r1, sqrtf(1.0f - r1 * r1), oracle_cosf(r0) * r1);
            *rorg = vadd(ipos, vscale(gnrm, eps));
            *rdir = onb(v, inrm);
            if (vdot(*rdir, gnrm) < 0.0f) *rdir = vneg(*rdir);
        }
    }
    float p = gl_max(gl_max(icol.x, icol.y), icol.z);
    if (p < gpurnd4(rndv)) {
        return 0;
    } else {
        icol = V3(icol.x / p, icol.y / p, icol.z / p);
        *flux = vmul(*flux, icol);
    }
    return 1;
}

/* Exported for the oracle's own tests: the camera ray, the light and IBL emission samples, the
 * glossy perturbation, one eye path, the Fresnel term and the photon bounce above, on caller-held
 * generator states (rnd: 4 floats, read and advanced). */
void oracle_camera_ray(const ppm_scene* s, const ppm_params* p, int i, int j, float aax, float aay, float* rnd,
                       float* org, float* dir, float* filter) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    v3 o, d;
    camera_ray(s, p, i, j, aax, aay, &r, &o, &d, filter);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}
int oracle_light_sample(const ppm_scene* s, float* rnd, float* flux, float* dir, float* org) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    v3 f, d, o;
    int tri = light_sample(s, &r, &f, &d, &o);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    flux[0] = f.x; flux[1] = f.y; flux[2] = f.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    return tri;
}
void oracle_ibl_sample(const ppm_scene* s, float* rnd, float* flux, float* dir, float* org) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    v3 f, d, o;
    ibl_sample(s, &r, &f, &d, &o);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    flux[0] = f.x; flux[1] = f.y; flux[2] = f.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
}
void oracle_glossy_reflect(const float* d, const float* n, float g, float* rnd, float* out) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    v3 res = glossy_reflect(V3(d[0], d[1], d[2]), V3(n[0], n[1], n[2]), g, &r);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    out[0] = res.x; out[1] = res.y; out[2] = res.z;
}
/* out: pos xyz, col xyz, nrm xyz, dir xyz, emission xyz, material, t (17 floats); returns the
 * triangle of the last segment, -1 for a miss. */
int oracle_eye_path(const ppm_scene* s, const ppm_params* p, int i, int j, float aax, float aay, float* rnd,
                    float* out) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    EyePath e;
    eye_path(s, p, i, j, aax, aay, &r, &e);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    const v3 v[5] = {e.pos, e.col, e.nrm, e.dir, e.emi};
    for (int k = 0; k < 5; k++) {
        out[3 * k] = v[k].x; out[3 * k + 1] = v[k].y; out[3 * k + 2] = v[k].z;
    }
    out[15] = e.mat;
    out[16] = e.t;
    return e.tri;
}
float oracle_fresnel(const float* incom, const float* normal, float ni, float ne) {
    return fresnel(V3(incom[0], incom[1], incom[2]), V3(normal[0], normal[1], normal[2]), ni, ne);
}
int oracle_photon_bounce(const ppm_scene* s, int tri, const float* ipos, const float* pnrm, const float* pd,
                         float* rnd, float* flux, float* org, float* dir) {
    v4 r = V4(rnd[0], rnd[1], rnd[2], rnd[3]);
    v3 f = V3(flux[0], flux[1], flux[2]), o = V3(0, 0, 0), d = V3(0, 0, 0);
    int cont = photon_bounce(s, tri, V3(ipos[0], ipos[1], ipos[2]), V3(pnrm[0], pnrm[1], pnrm[2]),
                             V3(pd[0], pd[1], pd[2]), &r, &f, &o, &d);
    rnd[0] = r.x; rnd[1] = r.y; rnd[2] = r.z; rnd[3] = r.w;
    flux[0] = f.x; flux[1] = f.y; flux[2] = f.z;
    org[0] = o.x; org[1] = o.y; org[2] = o.z;
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    return cont;
}

static int photon_fragment(Oracle* o, int px) {
    const ppm_scene* s = o->s;
    const int L = o->prm.max_path_length;
    v4 rndv = o->prng[px];
    v4 PP = o->ppos[px], PF = o->pflux[px], PD = o->pdir[px];
    v3 pnrm = V3(PP.w, PF.w, PD.w);
    v3 flux = V3(fabsf(PF.x), fabsf(PF.y), fabsf(PF.z));
    v3 rorg = V3(0, 0, 0), rdir = V3(0, 0, 0);
    int cont, emitted;
    int isect_tri = o->ptri[px];
    float isect_z = (float)o->pbounce[px] + 1.0f;
    if (isect_tri < 0 || isect_z >= (float)L) {
        if (s->light_area != 0.0f) light_sample(s, &rndv, &flux, &rdir, &rorg);
        else ibl_sample(s, &rndv, &flux, &rdir, &rorg);
        emitted = 1;
        isect_z = 0.0f;
        cont = 1;
    } else {
        cont = photon_bounce(s, isect_tri, V3(PP.x, PP.y, PP.z), pnrm, V3(PD.x, PD.y, PD.z), &rndv, &flux, &rorg,
                             &rdir);
        emitted = 0;
    }
    if (cont) {
        Isect I = raytrace(s, rorg, rdir, 0, 1);
        if ((I.t == 1.0e30f) || ((I.brdf != 0) && (I.brdf != 6) && (I.brdf != 3))) flux = vneg(flux);
        /* on a miss the shader's geometric normal is unset; it is taken as zero, 0/0 = NaN, and
         * the NaN-rejecting min returns sqrt(5) */
        float gf = gl_min(fabsf(vdot(rdir, I.nrm) / vdot(rdir, I.gnrm)), sqrtf(5.0f));
        flux = vscale(flux, gf);
        o->ppos[px] = V4(I.pos.x, I.pos.y, I.pos.z, I.nrm.x);
        o->pflux[px] = V4(flux.x, flux.y, flux.z, I.nrm.y);
        o->pdir[px] = V4(rdir.x, rdir.y, rdir.z, I.nrm.z);
        o->ptri[px] = I.isect_tri;
        o->pbounce[px] = (int)isect_z;
    } else {
        o->ppos[px] = V4(0, 0, 0, 0);
        o->pflux[px] = V4(-1.0f, -1.0f, -1.0f, 0.0f);
        o->pdir[px] = V4(0, 0, 0, 0);
        o->ptri[px] = -1;
        o->pbounce[px] = 0;
    }
    o->prng[px] = rndv;
    o->pemitted[px] = emitted;
    return emitted;
}

/* ---------------- hash.fs, scatter.vs, correction.vs ---------------- */
static void hash_fragment(Oracle* o, int px, float* pix_index) {
    /* PhotonIndex = (own texcoord, convert1Dto2D(hash)) or zw = -1 */
    v4 PP = o->ppos[px];
    float cell[3] = {floorf((PP.x - o->bbmin[0]) * o->grid_scale), floorf((PP.y - o->bbmin[1]) * o->grid_scale),
                     floorf((PP.z - o->bbmin[2]) * o->grid_scale)};
    float h = oracle_hash(cell, o->grid_scale, o->hash_scale, (float)o->H);
    float zx = gl_mod(h, (float)o->Hw) + 0.5f;
    float zy = floorf(h * (1.0f / (float)o->Hw)) + 0.5f;
    if (o->pflux[px].x < 0.0f) { zx = -1.0f; zy = -1.0f; }
    pix_index[0] = zx;
    pix_index[1] = zy;
}

static void build_hash(Oracle* o) {
    const int P = o->P;
    float* idx = (float*)malloc((size_t)P * 2 * sizeof(float));
    if (!idx) return;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; p++) hash_fragment(o, p, idx + 2 * p);
    /* glClear(colour | depth) */
    for (int h = 0; h < o->H; h++) { o->hrep[h] = -1; o->hdepth[h] = 1.0f; o->hcount[h] = 0.0f; }
    const float bz = 1.0f / (float)o->Hw, bw = 1.0f / (float)o->Hh;
    const float ps1 = 1.0f / (float)o->Pw;
    for (int p = 0; p < P; p++) {
        int i = p % o->Pw, j = p / o->Pw;
        float s = ((float)i + 0.5f) / (float)o->Pw, t = ((float)j + 0.5f) / (float)o->Ph;
        float zx = idx[2 * p], zy = idx[2 * p + 1];
        /* gl_Position.xy = zw * 2/size - 1, viewport transform back to pixels */
        float nx = zx * (bz + bz) - 1.0f, ny = zy * (bw + bw) - 1.0f;
        int slot = -1;
        if (nx >= -1.0f && nx <= 1.0f && ny >= -1.0f && ny <= 1.0f) {
            float wx = (nx + 1.0f) * 0.5f * (float)o->Hw, wy = (ny + 1.0f) * 0.5f * (float)o->Hh;
            int x = (int)floorf(wx), y = (int)floorf(wy);
            if (x >= 0 && x < o->Hw && y >= 0 && y < o->Hh) slot = y * o->Hw + x;
        }
        o->pslot[p] = slot;
        if (slot < 0) continue;
        /* scatter: depth z = s/Pw + t, clipped outside [-1, 1], GL_LESS against 1.0 */
        float z = (s * ps1) + t;
        if (z <= 1.0f) {
            float depth = (z + 1.0f) * 0.5f;
            if (depth < o->hdepth[slot]) { o->hdepth[slot] = depth; o->hrep[slot] = p; }
        }
        /* correction: z = 0.5, blend ONE, ONE of 1.0 */
        o->hcount[slot] = o->hcount[slot] + 1.0f;
    }
    free(idx);
}

/* ---------------- progressive.fs ---------------- */
static void accumulate(Oracle* o, v3 qpos, v3 qnrm, float qr, v3 cell, v3* flux, float* count) {
    float hidx[3] = {cell.x, cell.y, cell.z};
    float h = oracle_hash(hidx, o->grid_scale, o->hash_scale, (float)o->H);
    float u = (gl_mod(h, (float)o->Hw) + 0.5f) * (1.0f / (float)o->Hw);
    float v = (floorf(h * (1.0f / (float)o->Hw)) + 0.5f) * (1.0f / (float)o->Hh);
    /* GL_CLAMP + GL_NEAREST; the clamp rejects NaN like GLSL min/max, so a NaN coordinate (a
     * zero-extent hit-point box gives grid_scale = inf and NaN cells) reads texel (0, 0) */
    u = gl_min(gl_max(u, 0.0f), 1.0f);
    v = gl_min(gl_max(v, 0.0f), 1.0f);
    int x = (int)floorf(u * (float)o->Hw), y = (int)floorf(v * (float)o->Hh);
    if (x > o->Hw - 1) x = o->Hw - 1;
    if (y > o->Hh - 1) y = o->Hh - 1;
    int slot = y * o->Hw + x;
    /* the hash texture's colour holds the representative's texcoord; cleared texels read (0, 0) */
    int rep = o->hrep[slot] < 0 ? 0 : o->hrep[slot];
    v4 PF = o->pflux[rep], PP = o->ppos[rep], PD = o->pdir[rep];
    v3 ppos = V3(PP.x, PP.y, PP.z);
    v3 bb = V3(o->bbmin[0], o->bbmin[1], o->bbmin[2]);
    v3 rmin = vadd(V3(cell.x / o->grid_scale, cell.y / o->grid_scale, cell.z / o->grid_scale), bb);
    v3 c1 = vadd(cell, V3(1.0f, 1.0f, 1.0f));
    v3 rmax = vadd(V3(c1.x / o->grid_scale, c1.y / o->grid_scale, c1.z / o->grid_scale), bb);
    if (ppos.x > rmin.x && ppos.y > rmin.y && ppos.z > rmin.z && ppos.x < rmax.x && ppos.y < rmax.y &&
        ppos.z < rmax.z) {
        v3 diff = vsub(ppos, qpos);
        if ((vlength(diff) < qr) && (vdot(qnrm, vneg(V3(PD.x, PD.y, PD.z))) > 0.0f)) {
            float corr = o->hcount[slot];
            *flux = vadd(*flux, vscale(V3(PF.x, PF.y, PF.z), corr));
            *count = *count + corr;
        }
    }
}

/* progressive.fs:129-137: BRDF scaling of the gathered flux, then the N/R/tau update with
 * g = min((N + alpha M) / (N + M), 1), NaN-rejecting so that N + M = 0 gives g = 1. */
void oracle_update(float* tau, float* radius, float* count, const float* phi, float m, const float* refl,
                   float alpha) {
    v3 flux = vmul(V3(phi[0], phi[1], phi[2]), vscale(V3(refl[0], refl[1], refl[2]),
                                                      0.31830995240629591621063460818591f));
    float qn = *count, qr = *radius;
    float g = gl_min((qn + alpha * m) / (qn + m), 1.0f);
    qr = qr * sqrtf(g);
    qn = qn + m * alpha;
    v3 t = vscale(vadd(V3(tau[0], tau[1], tau[2]), flux), g);
    tau[0] = t.x; tau[1] = t.y; tau[2] = t.z;
    *radius = qr;
    *count = qn;
}

static void progressive_fragment(Oracle* o, int px, float alpha) {
    v4 QP = o->qpos[px], QN = o->qnrm[px], QFR = o->qflux_r[px], QEN = o->qemi_n[px], QR = o->qrefl[px];
    v3 qflux = V3(QFR.x, QFR.y, QFR.z);
    float qn = QEN.w, qr = QFR.w;
    if (o->qtri[px] >= 0) {
        v3 qpos = V3(QP.x, QP.y, QP.z);
        v3 d = vsub(qpos, V3(o->bbmin[0], o->bbmin[1], o->bbmin[2]));
        v3 ii = V3(floorf(fabsf(d.x) * o->grid_scale - 0.5f), floorf(fabsf(d.y) * o->grid_scale - 0.5f),
                   floorf(fabsf(d.z) * o->grid_scale - 0.5f));
        v3 flux = V3(0, 0, 0);
        float m = 0.0f;
        for (int iz = 0; iz <= 1; iz++)
            for (int iy = 0; iy <= 1; iy++)
                for (int ix = 0; ix <= 1; ix++)
                    accumulate(o, qpos, V3(QN.x, QN.y, QN.z), qr,
                               vadd(ii, V3((float)ix, (float)iy, (float)iz)), &flux, &m);
        float tau[3] = {qflux.x, qflux.y, qflux.z}, phi[3] = {flux.x, flux.y, flux.z}, refl[3] = {QR.x, QR.y, QR.z};
        oracle_update(tau, &qr, &qn, phi, m, refl, alpha);
        qflux = V3(tau[0], tau[1], tau[2]);
    }
    o->qflux_r[px] = V4(qflux.x, qflux.y, qflux.z, qr);
    o->qemi_n[px] = V4(QEN.x, QEN.y, QEN.z, qn);
}

/* ---------------- re.fs ---------------- */
static float srgb(float c) {
    if (c < 0.0031308f) return 12.92f * c;
    const float a = 0.055f;
    return (1.0f + a) * oracle_powf(c, 0.416666666667f) - a;
}
void oracle_radiance(const float* flux_r, const float* emi, float total_photons, float* lin, float* mapped) {
    float den = flux_r[3] * flux_r[3] * 3.141592f * total_photons;
    float c[4] = {flux_r[0] / den + emi[0], flux_r[1] / den + emi[1], flux_r[2] / den + emi[2], 1.0f + 0.0f};
    for (int k = 0; k < 4; k++) lin[k] = c[k];
    for (int k = 0; k < 4; k++) mapped[k] = 1.0f - oracle_expf(-c[k] * 60000.0f);
    for (int k = 0; k < 3; k++) mapped[k] = srgb(mapped[k]);
}

/* ---------------- m_display() ---------------- */
int oracle_pass(void* h, uint32_t npasses) {
    Oracle* o = (Oracle*)h;
    if (!o) return PPM_E_INVAL;
    const int L = o->prm.max_path_length;
    for (uint32_t it = 0; it < npasses; it++) {
        if (o->frame == 0) {
            for (int i = 0; i < o->Q; i++) o->qemi_n[i] = V4(0, 0, 0, 0);
        }
        if (o->frame % (uint64_t)L == 0) {
            float aax = (xorshift_frand(o->xs) - 0.5f) * 1.25f;
            float aay = (xorshift_frand(o->xs) - 0.5f) * 1.25f;
            int nes = (int)o->eye_samples + 1;
            o->eye_samples++;
#pragma omp parallel for schedule(dynamic, 64)
            for (int p = 0; p < o->Q; p++) eye_fragment(o, p % o->Qw, p / o->Qw, aax, aay, nes);
            /* max.fs / min.fs over QueryPositionTexture */
            float mx[3] = {-INFINITY, -INFINITY, -INFINITY}, mn[3] = {INFINITY, INFINITY, INFINITY};
            for (int p = 0; p < o->Q; p++) {
                const float* v = &o->qpos[p].x;
                for (int k = 0; k < 3; k++) { mx[k] = gl_max(mx[k], v[k]); mn[k] = gl_min(mn[k], v[k]); }
            }
            float bbsize = 0.0f;
            for (int k = 0; k < 3; k++) bbsize += mx[k] - mn[k];
            o->ir = (bbsize / 3.0f) / (float)o->Qw * o->prm.initial_footprint;
            if (o->prm.fixed_radius > 0.0f) o->ir = o->prm.fixed_radius;
            for (int k = 0; k < 3; k++) { o->bbmin[k] = mn[k] - o->ir; o->bbmax[k] = mx[k] + o->ir; }
            o->grid_scale = 0.5f / o->ir;
            o->hash_scale = 4194304.0f / o->grid_scale;
        }
        if (o->frame == 0) {
            for (int i = 0; i < o->Q; i++) o->qflux_r[i] = V4(0, 0, 0, o->ir);
            for (int p = 0; p < o->P; p++) { o->ptri[p] = -1; o->pbounce[p] = 0; }
            o->num_photons = 0.0f;
        }
        unsigned long long emitted = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : emitted)
        for (int p = 0; p < o->P; p++) emitted += (unsigned long long)photon_fragment(o, p);
        o->num_photons += (float)(int)emitted;
        o->photons_emitted += emitted;
        build_hash(o);
        const float alpha = o->prm.alpha;
#pragma omp parallel for schedule(static)
        for (int p = 0; p < o->Q; p++) progressive_fragment(o, p, alpha);
        o->frame++;
    }
    return PPM_OK;
}

size_t oracle_buffer_bytes(void* h, int which) {
    Oracle* o = (Oracle*)h;
    if (!o) return 0;
    size_t Q = o->Q, P = o->P, H = o->H;
    switch (which) {
        case PPM_BUF_RADIANCE: case PPM_BUF_TONEMAPPED: case PPM_BUF_HIT_POSITION: case PPM_BUF_HIT_REFLECTANCE:
        case PPM_BUF_HIT_NORMAL: case PPM_BUF_EYE_RNG: case PPM_BUF_FLUX_RADIUS: case PPM_BUF_EMISSION_COUNT:
            return Q * 16;
        case PPM_BUF_RGBA8: return Q * 4;
        case PPM_BUF_HIT_TRIANGLE: return Q * 8;
        case PPM_BUF_HIT_T: return Q * 4;
        case PPM_BUF_PHOTON_POSITION: case PPM_BUF_PHOTON_FLUX: case PPM_BUF_PHOTON_DIRECTION:
        case PPM_BUF_PHOTON_RNG: return P * 16;
        case PPM_BUF_PHOTON_STATE: return P * 12;
        case PPM_BUF_PHOTON_SLOT: return P * 4;
        case PPM_BUF_SLOT_COUNT: case PPM_BUF_SLOT_REP: return H * 4;
        case PPM_BUF_COUNTERS: return sizeof(ppm_counters);
        case PPM_BUF_GRID: return sizeof(ppm_grid);
        default: return 0;
    }
}

static unsigned char to_unorm8(float c) {
    c = c > 1.0f ? 1.0f : c;
    c = c > 0.0f ? c : 0.0f;
    return (unsigned char)(int)(c * 255.0f + 0.5f);
}

int oracle_readback(void* h, int which, void* dst, size_t bytes) {
    Oracle* o = (Oracle*)h;
    if (!o || !dst) return PPM_E_INVAL;
    if (bytes != oracle_buffer_bytes(h, which) || bytes == 0) return PPM_E_INVAL;
    size_t Q = o->Q, P = o->P, H = o->H;
    switch (which) {
        case PPM_BUF_RADIANCE: case PPM_BUF_TONEMAPPED: case PPM_BUF_RGBA8: {
            for (size_t i = 0; i < Q; i++) {
                float lin[4], mp[4];
                oracle_radiance(&o->qflux_r[i].x, &o->qemi_n[i].x, o->num_photons, lin, mp);
                if (which == PPM_BUF_RADIANCE) memcpy((float*)dst + 4 * i, lin, 16);
                else if (which == PPM_BUF_TONEMAPPED) memcpy((float*)dst + 4 * i, mp, 16);
                else for (int k = 0; k < 4; k++) ((unsigned char*)dst)[4 * i + k] = to_unorm8(mp[k]);
            }
            return PPM_OK;
        }
        case PPM_BUF_HIT_POSITION: memcpy(dst, o->qpos, Q * 16); return PPM_OK;
        case PPM_BUF_HIT_REFLECTANCE: memcpy(dst, o->qrefl, Q * 16); return PPM_OK;
        case PPM_BUF_HIT_NORMAL: memcpy(dst, o->qnrm, Q * 16); return PPM_OK;
        case PPM_BUF_EYE_RNG: memcpy(dst, o->qrng, Q * 16); return PPM_OK;
        case PPM_BUF_FLUX_RADIUS: memcpy(dst, o->qflux_r, Q * 16); return PPM_OK;
        case PPM_BUF_EMISSION_COUNT: memcpy(dst, o->qemi_n, Q * 16); return PPM_OK;
        case PPM_BUF_HIT_TRIANGLE:
            for (size_t i = 0; i < Q; i++) {
                ((int32_t*)dst)[2 * i] = o->qtri[i];
                ((int32_t*)dst)[2 * i + 1] = o->qtri[i] < 0 ? -1 : (int32_t)o->qmat[i];
            }
            return PPM_OK;
        case PPM_BUF_HIT_T: memcpy(dst, o->qt, Q * 4); return PPM_OK;
        case PPM_BUF_PHOTON_POSITION: memcpy(dst, o->ppos, P * 16); return PPM_OK;
        case PPM_BUF_PHOTON_FLUX: memcpy(dst, o->pflux, P * 16); return PPM_OK;
        case PPM_BUF_PHOTON_DIRECTION: memcpy(dst, o->pdir, P * 16); return PPM_OK;
        case PPM_BUF_PHOTON_RNG: memcpy(dst, o->prng, P * 16); return PPM_OK;
        case PPM_BUF_PHOTON_STATE:
            for (size_t i = 0; i < P; i++) {
                ((int32_t*)dst)[3 * i] = o->ptri[i];
                ((int32_t*)dst)[3 * i + 1] = o->pbounce[i];
                ((int32_t*)dst)[3 * i + 2] = o->pemitted[i];
            }
            return PPM_OK;
        case PPM_BUF_PHOTON_SLOT: memcpy(dst, o->pslot, P * 4); return PPM_OK;
        case PPM_BUF_SLOT_COUNT:
            for (size_t i = 0; i < H; i++) ((uint32_t*)dst)[i] = (uint32_t)o->hcount[i];
            return PPM_OK;
        case PPM_BUF_SLOT_REP: memcpy(dst, o->hrep, H * 4); return PPM_OK;
        case PPM_BUF_COUNTERS: {
            ppm_counters k;
            memset(&k, 0, sizeof(k));
            k.frame_count = o->frame;
            k.eye_samples = o->eye_samples;
            k.photons_emitted = o->photons_emitted;
            k.photon_segments = (uint64_t)o->P * o->frame;
            k.num_photons_f = o->num_photons;
            memcpy(dst, &k, sizeof(k));
            return PPM_OK;
        }
        case PPM_BUF_GRID: {
            ppm_grid g;
            for (int k = 0; k < 3; k++) { g.bbmin[k] = o->bbmin[k]; g.bbmax[k] = o->bbmax[k]; }
            g.initial_radius = o->ir;
            g.grid_scale = o->grid_scale;
            g.hash_scale = o->hash_scale;
            memcpy(dst, &g, sizeof(g));
            return PPM_OK;
        }
        default: return PPM_E_INVAL;
    }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
