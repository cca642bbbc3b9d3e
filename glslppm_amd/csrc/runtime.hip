// Context, HBM allocation, pass sequencing and readback behind include/ppm.h.
//
// ppm_create replaces the texture/FBO set-up of main() (glslppm.cpp:664-829) and the scene
// packing of CGPURT::CreateTextures (gpurt.cpp:47-552); ppm_pass replaces m_display()
// (glslppm.cpp:214-536); ppm_readback replaces glReadPixels and reading the state textures.
// Everything a pass needs (grid scalars, photon counters, frame count, eye cadence, AAOffset) stays
// in HBM, so ppm_pass only queues kernels on the context's stream and never waits for the device.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "context.h"

// Scratch of ppm_pass_static: the eye kernel's per-block min/max partials (24 bytes per tile of
// 256 pixels: images up to about 11M pixels).  Each GPU holds its own copy for the life of the process, and every
// ppm_pass_static call on that GPU uses it.  Host code uses the array only through the device
// address taken once per device at set-up (static_scratch_address), never by its name.
constexpr size_t kStaticScratchBytes = 1u << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

thread_local std::string g_create_error;

// Device addresses of `scratch`, one per device, taken by the first context created there.
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_taken[kScratchDevices] = {};
void* g_scratch_ptr[kScratchDevices] = {};
hipError_t g_scratch_err[kScratchDevices] = {};

// With `device` current: takes the address of this device's copy of `scratch` once.
void static_scratch_address(int device) {
    if (device < 0 || device >= kScratchDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_taken[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    g_scratch_taken[device] = true;
}

int set_err(ppm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_err((ctx), e_ == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP,         \
                           std::string(#call) + ": " + hipGetErrorString(e_));               \
    } while (0)

// The address static_scratch_address took for the context's device.
int static_scratch(ppm_ctx* c, void** out) {
    const int d = c->device;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (d < 0 || d >= kScratchDevices || !g_scratch_taken[d])
        return set_err(c, PPM_E_INVAL, "ppm_pass_static: no static scratch on device " + std::to_string(d));
    if (!g_scratch_ptr[d])
        return set_err(c, PPM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[d]));
    *out = g_scratch_ptr[d];
    return PPM_OK;
}

template <typename T>
int dev_alloc(ppm_ctx* c, T** out, size_t count) {
    void* p = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 16;
    hipError_t e = c->arena ? c->arena->alloc(&p, bytes) : hipMalloc(&p, bytes);
    if (e != hipSuccess)
        return set_err(c, e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP,
                       std::string("device allocation of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    c->allocs.push_back(p);
    c->device_bytes += c->arena ? ppm_internal::Arena::round_size(bytes) : bytes;
    *out = (T*)p;
    return PPM_OK;
}

inline float as_float(int v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// XORShift::m_frand (glslppm.cpp:77-91)
inline float xs_frand(uint32_t* s) {
    const uint32_t t = s[0] ^ (s[0] << 11);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = s[3];
    s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    return (float)s[3] * (1.0f / 4294967295.0f);
}

// The seed grids of RandomizeTextures (glslppm.cpp:113-140) take four draws of the XORShift
// stream per texel, the eye grid's Q texels and then the photon grid's P.  k_seed draws them on the
// device, one thread per chunk of kSeedChunk texels, each from its chunk's start state.  The state
// update of xs_frand is linear over GF(2): the state after n steps is M^n s for a 128 x 128 bit
// matrix M, so the start of chunk c + 1 is J s_c with J = M^(4 kSeedChunk), without stepping
// through the draws between (drawing a 1024^2 job's 2M texels one after another took 39 ms of host
// time, most of the server's admission work).
constexpr int kSeedChunk = 1024;

struct Bits128 {
    uint32_t w[4];
};

// A 128 x 128 bit matrix by columns: column k is the image of the state with only bit k set
// (bit k is bit k % 32 of word k / 32, words in XORShift state order).
struct BitMatrix {
    Bits128 col[128];
};

Bits128 xs_step_bits(Bits128 s) {
    const uint32_t t = s.w[0] ^ (s.w[0] << 11);
    return Bits128{{s.w[1], s.w[2], s.w[3], (s.w[3] ^ (s.w[3] >> 19)) ^ (t ^ (t >> 8))}};
}

Bits128 apply_bits(const BitMatrix& a, const Bits128& v) {
    Bits128 r = {{0, 0, 0, 0}};
    for (int k = 0; k < 128; k++)
        if ((v.w[k >> 5] >> (k & 31)) & 1u)
            for (int i = 0; i < 4; i++) r.w[i] ^= a.col[k].w[i];
    return r;
}

// J = M^(4 kSeedChunk), by squaring M (4 kSeedChunk is a power of two), as 16 tables of 256
// entries: entry v of table b is J applied to byte b of the state being v, so that J s is the XOR
// of 16 entries; built once.
struct JumpTable {
    Bits128 t[16][256];
};

const JumpTable& seed_chunk_jump() {
    static const JumpTable* table = [] {
        static_assert((4 * kSeedChunk & (4 * kSeedChunk - 1)) == 0, "the chunk's draws must be a power of two");
        static BitMatrix m;
        for (int k = 0; k < 128; k++) {
            Bits128 e = {{0, 0, 0, 0}};
            e.w[k >> 5] = 1u << (k & 31);
            m.col[k] = xs_step_bits(e);
        }
        for (int d = 1; d < 4 * kSeedChunk; d *= 2) {
            static BitMatrix sq;
            for (int k = 0; k < 128; k++) sq.col[k] = apply_bits(m, m.col[k]);
            m = sq;
        }
        static JumpTable jt;
        for (int b = 0; b < 16; b++) {
            jt.t[b][0] = Bits128{{0, 0, 0, 0}};
            for (int v = 1; v < 256; v++) {
                const int low = __builtin_ctz((unsigned)v);
                for (int i = 0; i < 4; i++) jt.t[b][v].w[i] = jt.t[b][v & (v - 1)].w[i] ^ m.col[8 * b + low].w[i];
            }
        }
        return &jt;
    }();
    return *table;
}

Bits128 jump_bits(const JumpTable& jt, const Bits128& v) {
    Bits128 r = {{0, 0, 0, 0}};
    for (int b = 0; b < 16; b++) {
        const Bits128& e = jt.t[b][(v.w[b >> 2] >> ((b & 3) * 8)) & 255u];
        for (int i = 0; i < 4; i++) r.w[i] ^= e.w[i];
    }
    return r;
}

// The start state of every chunk of n texels drawn from state s, and s advanced past all of them.
void seed_starts(uint32_t* s, size_t n, std::vector<uint4>& starts) {
    const JumpTable& jump = seed_chunk_jump();
    const size_t chunks = (n + kSeedChunk - 1) / kSeedChunk;
    starts.resize(chunks);
    Bits128 b = {{s[0], s[1], s[2], s[3]}};
    for (size_t c = 0; c < chunks; c++) {
        starts[c] = make_uint4(b.w[0], b.w[1], b.w[2], b.w[3]);
        if (c + 1 < chunks) b = jump_bits(jump, b);
    }
    // the last chunk's draws, one by one
    for (size_t i = (chunks - 1) * kSeedChunk; i < n; i++)
        for (int d = 0; d < 4; d++) b = xs_step_bits(b);
    for (int i = 0; i < 4; i++) s[i] = b.w[i];
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int check_scene(const ppm_scene* s) {
    if (!s) return set_err(nullptr, PPM_E_INVAL, "null scene");
    if (s->num_triangles <= 0 || !s->positions || !s->normals || !s->material)
        return set_err(nullptr, PPM_E_INVAL, "scene has no triangles");
    if (s->num_nodes != 2 * s->num_triangles - 1 || !s->node_bounds || !s->node_triangle || !s->node_links)
        return set_err(nullptr, PPM_E_INVAL, "scene BVH must have 2T-1 nodes");
    if (s->num_materials <= 0 || !s->materials) return set_err(nullptr, PPM_E_INVAL, "scene has no materials");
    if (s->num_lights < 0 || (s->num_lights > 0 && (!s->light_triangle || !s->light_cdf)))
        return set_err(nullptr, PPM_E_INVAL, "bad light table");
    for (int i = 0; i < s->num_triangles; i++)
        if (s->material[i] < 0 || s->material[i] >= s->num_materials)
            return set_err(nullptr, PPM_E_INVAL, "triangle material out of range");
    std::vector<char> in_leaf(s->num_triangles, 0);
    for (int k = 0; k < s->num_nodes; k++) {
        const int t = s->node_triangle[k];
        if (t < -1 || t >= s->num_triangles) return set_err(nullptr, PPM_E_INVAL, "leaf triangle out of range");
        if (t >= 0 && in_leaf[t]++) return set_err(nullptr, PPM_E_INVAL, "triangle in two leaves");
    }
    for (int t = 0; t < s->num_triangles; t++)
        if (!in_leaf[t]) return set_err(nullptr, PPM_E_INVAL, "triangle in no leaf");
    for (size_t k = 0; k < (size_t)6 * s->num_nodes * 2; k++)
        if (s->node_links[k] < -1 || s->node_links[k] >= s->num_nodes)
            return set_err(nullptr, PPM_E_INVAL, "BVH link out of range");
    for (int l = 0; l < s->num_lights; l++)
        if (s->light_triangle[l] < 0 || s->light_triangle[l] >= s->num_triangles)
            return set_err(nullptr, PPM_E_INVAL, "light triangle out of range");
    if (s->texcoords && (!s->textures || s->texture_size <= 0 || s->texture_size > 8192))
        return set_err(nullptr, PPM_E_INVAL, "texture coordinates need a texture volume of size 1..8192");
    return PPM_OK;
}

uint16_t half_bits(_Float16 h) {
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

float half_value(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (float)h;
}

// The largest binary16 value <= x (down) or the smallest >= x (up), never a subnormal (those
// step further outward, to a signed zero or the smallest normal, so that the bounds hold whether
// or not a unit flushes binary16 subnormals); the sign of x is kept, so a positive bound never
// rounds below +0 and a negative one never above -0.
uint16_t half_outward(float x, bool down) {
    uint16_t b = half_bits((_Float16)x);
    for (;;) {
        const float v = half_value(b);
        if (down ? !(v > x) : !(v < x)) break;
        const bool neg = (b & 0x8000u) != 0;
        if (down) b = (b == 0x0000u) ? 0x8001u : (neg ? b + 1 : b - 1);
        else b = (b == 0x8000u) ? 0x0001u : (neg ? b - 1 : b + 1);
    }
    if ((b & 0x7c00u) == 0 && (b & 0x03ffu) != 0) {
        const bool neg = (b & 0x8000u) != 0;
        if (down) b = neg ? 0x8400u : 0x0000u;
        else b = neg ? 0x8000u : 0x0400u;
    }
    return b;
}

// Per-face pre-order node records, 16 bytes each: the box as binary16 bounds rounded outward and
// ordered by the face (for a face whose rays go mainly along +axis M, or -M: x = near | far << 16
// on M, i.e. lo | hi, or hi | lo; y = lo | hi << 16 on the lower other axis; z = the same on the
// higher one), and w = the miss position of an interior node or ~triangle for a leaf, whose miss
// position is always the next one.  The hit link of position i is i + 1.  Each face array has N + 1
// records: the last is padding that the walk's look-ahead fetch may read but never visits.
//
// The same trees, four-wide: each face's pre-order is read as a tree (the children of position i
// are i + 1, then the miss position of each child in turn up to the miss position of i), and
// every record holds a cut of up to four descendants of one interior node, found by replacing the
// interior member of largest surface area by its children while they fit, in pre-order.  Record 0
// holds the cut below the top of the tree.  A record is four slots in the node format above whose
// w is the member's record (interior), ~triangle (a leaf) or kWideEmpty; records are numbered in
// pre-order.  An empty slot's near bound on the major axis is infinite (+inf for a + face, -inf
// for a - face), so that its near slab value is +inf for every ray the record tests serve (finite
// nonzero reciprocals of the face's sign, finite t0) and the slot never passes.  wide_pos holds each record's four members' pre-order positions (N for
// an empty slot), where a walk that runs out of stack continues on the pre-order records.
// A node record of face f from the six outward-rounded bounds (lo xyz, hi xyz) and w.
uint4 face_box(const uint16_t* h, int f, uint32_t w) {
    const int M = f >> 1, A = M == 0 ? 1 : 0, B = M == 2 ? 1 : 2;
    const uint32_t lo = h[M], hi = h[3 + M];
    return make_uint4((f & 1) ? hi | lo << 16 : lo | hi << 16, h[A] | (uint32_t)h[3 + A] << 16,
                      h[B] | (uint32_t)h[3 + B] << 16, w);
}

// An unused member slot of a four-wide record of face f.
uint4 empty_slot(int f) { return make_uint4((f & 1) ? 0xFC00FC00u : 0x7C007C00u, 0, 0, kWideEmpty); }

int pack_nodes(const ppm_scene* s, std::vector<uint4>& out, std::vector<uint4>& wide, std::vector<uint4>& wide_pos,
               int* wide_stride) {
    const int N = s->num_nodes;
    out.assign((size_t)6 * (N + 1), make_uint4(0, 0, 0, 0));
    std::vector<int> order(N), pos(N), mpos(N), tri(N);
    std::vector<std::array<uint16_t, 6>> hb(N);
    std::vector<std::vector<uint4>> fw(6), fp(6);
    int stride = 1;
    for (int f = 0; f < 6; f++) {
        const int32_t* L = s->node_links + (size_t)f * N * 2;
        std::fill(pos.begin(), pos.end(), -1);
        int k = 0, c = 0;
        while (k >= 0) {
            if (c >= N || pos[k] >= 0) return set_err(nullptr, PPM_E_INVAL, "BVH hit links do not form one chain");
            order[c] = k;
            pos[k] = c++;
            k = L[2 * k];
        }
        if (c != N) return set_err(nullptr, PPM_E_INVAL, "BVH hit chain misses nodes");
        for (int i = 0; i < N; i++) {
            const int nk = order[i];
            const int miss = L[2 * nk + 1];
            mpos[i] = miss < 0 ? N : pos[miss];
            if (miss >= 0 && mpos[i] <= i) return set_err(nullptr, PPM_E_INVAL, "BVH miss link points backwards");
            const float* b = s->node_bounds + (size_t)nk * 6;
            tri[i] = s->node_triangle[nk];
            if (tri[i] >= 0 && mpos[i] != i + 1) return set_err(nullptr, PPM_E_INVAL, "BVH leaf miss link is not the next node");
            uint16_t* h = hb[i].data();
            for (int k = 0; k < 6; k++) h[k] = half_outward(b[k], k < 3);
            out[(size_t)f * (N + 1) + i] = face_box(h, f, tri[i] >= 0 ? (uint32_t)~tri[i] : (uint32_t)mpos[i]);
        }
        out[(size_t)f * (N + 1) + N] = make_uint4(0, 0, 0, (uint32_t)(N + 1));

        // children of position x (x = -1: the top of the tree), checked to tile x's range
        auto children = [&](int x, std::vector<int>& ch) -> bool {
            ch.clear();
            const int end = x < 0 ? N : mpos[x];
            int c = x + 1;
            while (c < end) {
                ch.push_back(c);
                if (ch.size() > 4) return false;
                c = mpos[c];
            }
            return c == end && (x < 0 || !ch.empty());
        };
        auto area = [&](int i) {
            const float* b = s->node_bounds + (size_t)order[i] * 6;
            const double dx = (double)b[3] - b[0], dy = (double)b[4] - b[1], dz = (double)b[5] - b[2];
            return dx * dy + dy * dz + dz * dx;
        };
        std::vector<int> rec_of(N, -1), todo{-1}, ch, cut;
        std::vector<std::array<int, 4>> cuts;
        size_t members = 0;
        while (!todo.empty()) {
            const int x = todo.back();
            todo.pop_back();
            if (!children(x, cut)) return set_err(nullptr, PPM_E_INVAL, "BVH links do not form a binary tree");
            for (;;) {
                int best = -1;
                double ba = -1.0;
                for (size_t j = 0; j < cut.size(); j++)
                    if (tri[cut[j]] < 0 && area(cut[j]) > ba) {
                        ba = area(cut[j]);
                        best = (int)j;
                    }
                if (best < 0) break;
                if (!children(cut[best], ch)) return set_err(nullptr, PPM_E_INVAL, "BVH links do not form a binary tree");
                if (cut.size() - 1 + ch.size() > 4) break;
                cut.erase(cut.begin() + best);
                cut.insert(cut.begin() + best, ch.begin(), ch.end());
            }
            if (x >= 0) rec_of[x] = (int)cuts.size();
            std::array<int, 4> a = {-1, -1, -1, -1};
            for (size_t j = 0; j < cut.size(); j++) a[j] = cut[j];
            cuts.push_back(a);
            members += cut.size();
            if (members > (size_t)N) return set_err(nullptr, PPM_E_INVAL, "BVH links do not form a binary tree");
            for (size_t j = cut.size(); j-- > 0;)
                if (tri[cut[j]] < 0) todo.push_back(cut[j]);
        }
        std::vector<uint4>& W = fw[f];
        std::vector<uint4>& WP = fp[f];
        W.resize(cuts.size() * 4);
        WP.resize(cuts.size());
        for (size_t r = 0; r < cuts.size(); r++) {
            uint32_t p4[4];
            for (int j = 0; j < 4; j++) {
                const int m = cuts[r][j];
                uint4 slot = empty_slot(f);
                if (m >= 0) slot = face_box(hb[m].data(), f, tri[m] >= 0 ? (uint32_t)~tri[m] : (uint32_t)rec_of[m]);
                W[r * 4 + j] = slot;
                p4[j] = m >= 0 ? (uint32_t)m : (uint32_t)N;
            }
            WP[r] = make_uint4(p4[0], p4[1], p4[2], p4[3]);
        }
        stride = std::max(stride, (int)cuts.size());
    }
    wide.resize((size_t)6 * stride * 4);
    wide_pos.assign((size_t)6 * stride, make_uint4(N, N, N, N));
    for (int f = 0; f < 6; f++) {
        std::fill(wide.begin() + (size_t)f * stride * 4, wide.begin() + (size_t)(f + 1) * stride * 4, empty_slot(f));
        std::copy(fw[f].begin(), fw[f].end(), wide.begin() + (size_t)f * stride * 4);
        std::copy(fp[f].begin(), fp[f].end(), wide_pos.begin() + (size_t)f * stride);
    }
    *wide_stride = stride;
    return PPM_OK;
}

// The smallest and largest of three coordinates, as the walk takes them (fminf / fmaxf).
float min3f(float a, float b, float c) { return std::fmin(std::fmin(a, b), c); }
float max3f(float a, float b, float c) { return std::fmax(std::fmax(a, b), c); }

// Triangle records of textureTriangles (gpurt.cpp:115-162) in the form the kernels read:
// normal z as the shaders rebuild it, sign(sgn(n.z)(mat+1)) sqrt(|1 - x^2 - y^2|)
// (eyeraytrace.fs:268-270), and the geometric normal normalize(cross(v1-v0, v2-v0)).  The
// intersection record holds the three vertices (the walk forms the edges as the shaders do,
// eyeraytrace.fs:125-126) and the exact box of the triangle's leaf.  Returns 1 when some leaf's
// box is not the box of its triangle's vertices, so that the walk must read the stored boxes
// (equal as floats: bounds that differ only in the sign of a zero give the same comparisons).
int pack_triangles(const ppm_scene* s, std::vector<float4>& isect, std::vector<float4>& shade) {
    const int T = s->num_triangles;
    isect.resize((size_t)T * 4);
    shade.resize((size_t)T * 4);
    std::vector<int> leaf(T, -1);
    for (int k = 0; k < s->num_nodes; k++)
        if (s->node_triangle[k] >= 0) leaf[s->node_triangle[k]] = k;
    int stored = 0;
    for (int t = 0; t < T; t++) {
        const float* p = s->positions + (size_t)t * 9;
        const float* n = s->normals + (size_t)t * 9;
        const int mat = s->material[t];
        const int brdf = s->materials[mat].brdf;
        float e0[3], e1[3], rn[3][3];
        for (int k = 0; k < 3; k++) {
            e0[k] = p[3 + k] - p[k];
            e1[k] = p[6 + k] - p[k];
        }
        for (int v = 0; v < 3; v++) {
            const float nx = n[v * 3], ny = n[v * 3 + 1], nz = n[v * 3 + 2];
            rn[v][0] = nx;
            rn[v][1] = ny;
            rn[v][2] = (nz > 0.0f ? 1.0f : -1.0f) * std::sqrt(std::fabs(1.0f - nx * nx - ny * ny));
        }
        float c[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
        const float inv = 1.0f / std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const float gn[3] = {c[0] * inv, c[1] * inv, c[2] * inv};
        const float* lb = s->node_bounds + (size_t)leaf[t] * 6;
        for (int k = 0; k < 3; k++)
            if (!(lb[k] == min3f(p[k], p[3 + k], p[6 + k]) && lb[3 + k] == max3f(p[k], p[3 + k], p[6 + k])))
                stored = 1;
        isect[(size_t)t * 4 + 0] = make_float4(p[0], p[1], p[2], as_float(brdf));
        isect[(size_t)t * 4 + 1] = make_float4(p[3], p[4], p[5], lb[0]);
        isect[(size_t)t * 4 + 2] = make_float4(p[6], p[7], p[8], lb[1]);
        isect[(size_t)t * 4 + 3] = make_float4(lb[2], lb[3], lb[4], lb[5]);
        float4* o = &shade[(size_t)t * 4];
        o[0] = make_float4(rn[0][0], rn[0][1], rn[0][2], 0.0f);
        o[1] = make_float4(rn[1][0], rn[1][1], rn[1][2], 0.0f);
        o[2] = make_float4(rn[2][0], rn[2][1], rn[2][2], 0.0f);
        o[3] = make_float4(gn[0], gn[1], gn[2], as_float(mat));
    }
    return stored;
}

hipEvent_t pooled_event(ppm_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records an event before a launch when profiling; returns it (or null).
hipEvent_t prof_begin(ppm_ctx* c, int kind, hipStream_t st) {
    if (!c->profiling || !(c->prof_kinds >> kind & 1u)) return nullptr;
    hipEvent_t e = pooled_event(c);
    if (e) (void)hipEventRecord(e, st);
    return e;
}

void prof_end(ppm_ctx* c, int kind, hipEvent_t start, hipStream_t st) {
    if (!c->profiling || !start) return;
    hipEvent_t e = pooled_event(c);
    if (!e) return;
    (void)hipEventRecord(e, st);
    c->ev_pending.push_back({kind, {start, e}});
    // bound the number of outstanding events: wait for both streams (pairs of the eye stage are on
    // the second one), then fold every pair in
    if (c->ev_pending.size() >= 4096) {
        (void)sync_streams(c);
        for (auto& pe : c->ev_pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
                c->prof_ms[pe.first] += ms;
                c->prof_n[pe.first] += 1;
            }
            c->ev_pool.push_back(pe.second.first);
            c->ev_pool.push_back(pe.second.second);
        }
        c->ev_pending.clear();
    }
}

int check_params(const ppm_params* p) {
    if (!p) return set_err(nullptr, PPM_E_INVAL, "null params");
    if (p->image_width <= 0 || p->image_height <= 0 || p->photon_width <= 0 || p->photon_height <= 0 ||
        p->hash_width <= 0 || p->hash_height <= 0)
        return set_err(nullptr, PPM_E_INVAL, "sizes must be positive");
    const long long Q = (long long)p->image_width * p->image_height;
    const long long P = (long long)p->photon_width * p->photon_height;
    const long long H = (long long)p->hash_width * p->hash_height;
    if (Q > (1ll << 26) || P > (1ll << 26)) return set_err(nullptr, PPM_E_INVAL, "Q and P are limited to 2^26");
    if (H > (1ll << 24)) return set_err(nullptr, PPM_E_INVAL, "H is limited to 2^24 (slot numbers are exact floats)");
    if (!is_pow2(p->hash_width)) return set_err(nullptr, PPM_E_INVAL, "hash_width must be a power of two");
    if (p->max_path_length <= 0 || p->max_path_length > 0xffff)
        return set_err(nullptr, PPM_E_INVAL, "max_path_length out of range");
    if (!(p->fixed_radius >= 0.0f) || !(p->initial_footprint > 0.0f) || !(p->alpha > 0.0f && p->alpha <= 1.0f))
        return set_err(nullptr, PPM_E_INVAL, "bad radius or alpha");
    if (p->flags & ~(PPM_FLAG_NO_DRAW | PPM_FLAG_GRAPH)) return set_err(nullptr, PPM_E_INVAL, "unknown flag");
    return PPM_OK;
}

// scatter.vs: depth z = s/Pw + t; points with z >= 1 fail the depth test, which removes the
// columns i >= Pw^2 / (2 Ph) - 1/2 of the last photon row from the representatives
int rep_limit(const ppm_params& p) {
    const long long P = (long long)p.photon_width * p.photon_height;
    const double first = std::ceil((double)p.photon_width * p.photon_width / (2.0 * p.photon_height) - 0.5);
    const long long ex = (long long)p.photon_width - (long long)std::max(0.0, first);
    return (int)(P - std::max(0ll, ex));
}

// An eye stage of job J on stream es: AAOffset and NumEyeSamples (k_eye_begin), the eye trace
// and the grid set-up.  `ahead` begins the next eye pass's stage instead of this pass's.
void queue_eye_stage(ppm_ctx* c, const Job& J, hipStream_t es, int ahead) {
    hipEvent_t e0 = prof_begin(c, PPM_KERNEL_EYE, es);
    const int mode = ahead ? kEyeAhead : kEyeNow;
    hipLaunchKernelGGL(k_eye_begin, dim3(1), dim3(64), 0, es, J, mode, kAnyPass, 0);
    hipLaunchKernelGGL(k_eye, dim3(J.tile_blocks), dim3(256), walk_lds_bytes(J.walk_stack, 256), es, J, mode);
    hipLaunchKernelGGL(k_grid, dim3(1), dim3(1024), 0, es, J, mode);
    prof_end(c, PPM_KERNEL_EYE, e0, es);
}

// Launches one pass of job J on stream st: m_display's stage order (glslppm.cpp:214-536), except
// that the pass's gather (progressive.fs, re.fs) runs in the photon-stage launch of the pass after
// it, beside that pass's walks, or at the end of the sequence (queue_flush); it reads only what
// the pass it gathers left (its hash table of three, its half of the photon records, the uniforms
// k_pass_begin saved for it) and nothing the passes after it read, so every result is unchanged.
// The eye stage (eye trace and grid set-up) and the photon trace do not read each other's
// results: only the deposits need the grid.  So in an eye pass the eye stage runs on the side
// stream beside the photon stage, whose deposits wait (kSlotPending) for k_hash after the join;
// the deep eye paths, chains of dependent walks that keep few lanes busy, then share the GPU with
// the photon walk.  An eye pass whose stage ran ahead (kEyePassReady) deposits at once.
// kEyeStageAhead is the stage of the next eye pass alone, for the side stream after an eye pass: the
// passes in between read the other half of the double-buffered hit points.  `fork_ahead` starts
// that stage from inside a photon pass instead, after k_pass_begin, beside the photon stage.
void queue_pass(ppm_ctx* c, const Job& J, hipStream_t st, int kind, bool fork_ahead = false) {
    if (kind == kEyeStageAhead) {
        queue_eye_stage(c, J, st, 1);
        return;
    }
    const bool stage = kind == kEyePass;
    hipEvent_t e1 = prof_begin(c, PPM_KERNEL_CLEAR, st);
    hipLaunchKernelGGL(k_pass_begin, dim3(J.begin_blocks), dim3(256), 0, st, J, 0u);
    prof_end(c, PPM_KERNEL_CLEAR, e1, st);
    const bool fork = stage && c->side && c->eye_overlap;
    hipStream_t es = fork ? c->side : st;
    if (fork) {
        (void)hipEventRecord(c->ev_fork, st);
        (void)hipStreamWaitEvent(es, c->ev_fork, 0);
    }
    if (stage) queue_eye_stage(c, J, es, 0);
    if (fork) (void)hipEventRecord(c->ev_join, es);
    hipEvent_t e2 = prof_begin(c, PPM_KERNEL_PHOTON, st);
    if (fork_ahead) {
        (void)hipEventRecord(c->ev_fork, st);
        (void)hipStreamWaitEvent(c->side, c->ev_fork, 0);
        queue_eye_stage(c, J, c->side, 1);
        (void)hipEventRecord(c->ev_ahead, c->side);
    }
    hipLaunchKernelGGL(k_photon_stage, dim3(J.trace_blocks + J.gather_blocks), dim3(c->trace_block),
                       walk_lds_bytes(J.walk_stack, c->trace_block), st, J);
    prof_end(c, PPM_KERNEL_PHOTON, e2, st);
    if (fork) (void)hipStreamWaitEvent(st, c->ev_join, 0);
    if (stage) hipLaunchKernelGGL(k_hash, dim3(J.photon_blocks), dim3(256), 0, st, J);
}

// The gather of the last pass of a sequence (the other passes' gathers run in the photon-stage
// launch of the pass after them).
void queue_flush(ppm_ctx* c, const Job& J, hipStream_t st) {
    hipEvent_t e = prof_begin(c, PPM_KERNEL_GATHER, st);
    hipLaunchKernelGGL(k_gather, dim3(J.tile_blocks), dim3(256), 0, st, J);
    prof_end(c, PPM_KERNEL_GATHER, e, st);
}

// One captured pass of a kind (queue_pass) for PPM_FLAG_GRAPH.
int capture_pass(ppm_ctx* c, int kind) {
    if (c->graph_pass[kind]) return PPM_OK;
    hipGraph_t g = nullptr;
    hipStream_t st = kind == kEyeStageAhead ? c->side : c->stream;
    HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const bool prof = c->profiling;
    c->profiling = false;
    queue_pass(c, c->job, st, kind);
    c->profiling = prof;
    HIP_TRY(c, hipStreamEndCapture(st, &g));
    const hipError_t e = hipGraphInstantiate(&c->graph_pass[kind], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_TRY(c, e);
    return PPM_OK;
}

// The pass loop of ppm_pass and ppm_pass_static.  J is the job record the launches take: the
// context's own, or a copy whose partial results live in the static scratch (then no graphs,
// which are captured with the context's own record).
int run_passes(ppm_ctx* c, const Job& J, uint32_t npasses, bool allow_graph) {
    const uint64_t L = (uint64_t)c->prm.max_path_length;
    const bool graph = allow_graph && c->use_graph && !c->profiling;
    // after each eye pass the next eye pass's stage starts on `side`, beside the photon passes in
    // between, and that eye pass waits for it; with one pass per eye pass there is nothing between.
    // Launched stream by stream, it starts inside the first photon pass (queue_pass); a captured
    // pass cannot fork mid-way, so with graphs it starts right after the eye pass.
    const bool ahead = c->side && c->eye_overlap && L >= 2;
    if (graph) {
        int rc;
        for (int k = 0; k < kPassKinds; k++)
            if ((k != kEyeStageAhead || c->side) && (rc = capture_pass(c, k))) return rc;
    }
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    auto launch_ahead = [&]() -> int {
        HIP_TRY(c, hipEventRecord(c->ev_fork, c->stream));
        HIP_TRY(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
        if (graph) HIP_TRY(c, hipGraphLaunch(c->graph_pass[kEyeStageAhead], c->side));
        else queue_pass(c, J, c->side, kEyeStageAhead);
        HIP_TRY(c, hipEventRecord(c->ev_ahead, c->side));
        c->ahead_due = false;
        c->ahead_pending = true;
        return PPM_OK;
    };
    for (uint32_t it = 0; it < npasses; it++) {
        const bool eye = c->frame % L == 0;
        int rc;
        if (c->ahead_due && !ahead) c->ahead_due = false;  // the eye pass runs its own stage
        if (c->ahead_due && (graph || eye) && (rc = launch_ahead())) return rc;
        int kind = kPhotonPass;
        if (eye && c->ahead_pending) {
            HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_ahead, 0));
            c->ahead_pending = false;
            kind = kEyePassReady;
        } else if (eye) {
            kind = kEyePass;
        }
        if (graph) {
            HIP_TRY(c, hipGraphLaunch(c->graph_pass[kind], c->stream));
        } else {
            const bool fork = c->ahead_due && kind == kPhotonPass;
            queue_pass(c, J, c->stream, kind, fork);
            if (fork) {
                c->ahead_due = false;
                c->ahead_pending = true;
            }
        }
        if (eye && ahead) {
            c->ahead_due = true;
            if (graph && (rc = launch_ahead())) return rc;
        }
        c->frame++;
    }
    queue_flush(c, J, c->stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
    c->timed = true;
    return PPM_OK;
}


}  // namespace

namespace ppm_internal {

size_t Arena::round_size(size_t bytes) {
    if (bytes <= 65536) return 65536;
    size_t top = 65536;
    while (top < bytes) top <<= 1;
    const size_t q = top >> 3;  // eighths of the power of two above
    return (bytes + q - 1) / q * q;
}

Arena::~Arena() {
    (void)hipSetDevice(device_);
    for (auto& kv : free_)
        for (auto& b : kv.second) {
            if (b.ev) {
                (void)hipEventSynchronize(b.ev);
                (void)hipEventDestroy(b.ev);
            }
            (void)hipFree(b.p);
        }
}

hipError_t Arena::alloc(void** out, size_t bytes) {
    const size_t sz = round_size(bytes);
    auto it = free_.find(sz);
    if (it != free_.end()) {
        auto& v = it->second;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k].ev && hipEventQuery(v[k].ev) != hipSuccess) continue;
            if (v[k].ev) (void)hipEventDestroy(v[k].ev);
            *out = v[k].p;
            v.erase(v.begin() + (long)k);
            in_use_ += sz;
            return hipSuccess;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        trim(0);
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) return e;
    }
    size_of_[p] = sz;
    held_ += sz;
    in_use_ += sz;
    *out = p;
    return hipSuccess;
}

void Arena::release(void* p, hipStream_t last_use) {
    auto it = size_of_.find(p);
    if (it == size_of_.end()) return;
    hipEvent_t ev = nullptr;
    if (last_use && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(ev, last_use) != hipSuccess) {
            (void)hipEventDestroy(ev);
            ev = nullptr;
        }
    }
    free_[it->second].push_back(Cached{p, ev});
    in_use_ -= it->second;
}

void Arena::trim(size_t target) {
    for (auto& kv : free_) {
        auto& v = kv.second;
        for (size_t k = 0; k < v.size() && held_ > target;) {
            if (v[k].ev && hipEventQuery(v[k].ev) != hipSuccess) {
                k++;
                continue;
            }
            if (v[k].ev) (void)hipEventDestroy(v[k].ev);
            (void)hipFree(v[k].p);
            held_ -= kv.first;
            size_of_.erase(v[k].p);
            v.erase(v.begin() + (long)k);
        }
    }
}

int set_error(ppm_ctx* c, int code, const std::string& msg) { return set_err(c, code, msg); }
const std::string& create_error() { return g_create_error; }

int validate(const ppm_scene* s, const ppm_params* p) {
    int rc;
    if ((rc = check_params(p)) || (rc = check_scene(s))) return rc;
    return PPM_OK;
}

int validate_params(const ppm_params* p) { return check_params(p); }
int check_scene_public(const ppm_scene* s) { return check_scene(s); }

int tile_blocks(int width, int height, int wave_shift) {
    const int tw = 2 << wave_shift, th = 2 * (64 >> wave_shift);
    return ((width + tw - 1) / tw) * ((height + th - 1) / th);
}

int max_tile_blocks(const ppm_params* p) {
    int m = 0;
    for (int sh = 0; sh <= 6; sh++) m = std::max(m, tile_blocks(p->image_width, p->image_height, sh));
    return m;
}

int check_tuning(const ppm_tuning* t, int* wave_shift) {
    if (!t) return PPM_E_INVAL;
    if (t->trace_block < 64 || t->trace_block > kTraceBlockMax || t->trace_block % 64) return PPM_E_INVAL;
    int sh = -1;
    for (int k = 0; k <= 6; k++)
        if (t->tile_wave_width == (1 << k)) sh = k;
    if (sh < 0) return PPM_E_INVAL;
    if (t->passes_per_graph > 64) return PPM_E_INVAL;
    if (t->walk_stack < 0 || t->walk_stack > kWalkStackMax) return PPM_E_INVAL;
    if (t->eye_overlap != 0 && t->eye_overlap != 1) return PPM_E_INVAL;
    if (wave_shift) *wave_shift = sh;
    return PPM_OK;
}

void set_launch_shape(ppm_ctx* c, int trace_block, int wave_shift, int walk_stack) {
    Job& J = c->job;
    J.walk_stack = walk_stack;
    J.tl.wave_shift = wave_shift;
    J.tl.tiles_x = (J.tl.width + (2 << wave_shift) - 1) / (2 << wave_shift);
    J.tile_blocks = tile_blocks(J.tl.width, J.tl.height, wave_shift);
    c->trace_block = trace_block;
    J.trace_blocks = (c->P + trace_block - 1) / trace_block;
    J.gather_blocks = gather_blocks(J.tile_blocks, trace_block);
    for (int k = 0; k < kPassKinds; k++)
        if (c->graph_pass[k]) {
            // a graph in flight keeps running; destroy it once the streams are idle
            (void)sync_streams(c);
            (void)hipGraphExecDestroy(c->graph_pass[k]);
            c->graph_pass[k] = nullptr;
        }
}
int photon_blocks(const ppm_params* p) { return (p->photon_width * p->photon_height + 255) / 256; }

// Blocks of `block` threads that the photon-stage launch adds for the gather of the pass before:
// one wave per quarter of each of the tiles.
int gather_blocks(int tiles, int block) {
    const int waves = block / 64;
    return (4 * tiles + waves - 1) / waves;
}

size_t state_footprint(const ppm_params* p) {
    const size_t Q = (size_t)p->image_width * p->image_height, P = (size_t)p->photon_width * p->photon_height;
    const size_t H = (size_t)p->hash_width * p->hash_height;
    const size_t sizes[] = {
        // hit points: pos, refl, nrm, rng, emission (two buffers each), (tau, R), (E, N), tri, (mat, t)
        // (two buffers each), rgba8, eye partials
        Q * 32, Q * 32, Q * 32, Q * 32, Q * 32, Q * 16, Q * 16, Q * 8, Q * 16, Q * 4, (size_t)max_tile_blocks(p) * 24,
        (size_t)max_tile_blocks(p) * 4, (size_t)max_tile_blocks(p) * 4,  // tile costs and order
        // photons: pos, flux, dir (two halves each), rng, state, slot
        P * 32, P * 32, P * 32, P * 16, P * 8, P * 4,
        // hash (count, representative) records, three tables; globals; resolve targets; the seed
        // chunks' start states
        H * 24, sizeof(DevGlobals), Q * 16, Q * 16, (Q + P + 1023) / 1024 * 16};
    size_t tot = 0;
    for (size_t b : sizes) tot += Arena::round_size(b ? b : 16);
    return tot;
}

size_t scene_footprint(int T, int N, int M, int NL, int tex_size, int wide_stride) {
    const size_t S = (size_t)tex_size, R = (size_t)wide_stride;
    const size_t sizes[] = {(size_t)6 * (N + 1) * 16, 6 * R * 64, 6 * R * 16, (size_t)T * 64, (size_t)T * 64,
                            (size_t)M * 32, (size_t)NL * 4, (size_t)NL * 4, S ? (size_t)T * 32 : 0, (size_t)M * S * S * 4};
    size_t tot = 0;
    for (size_t b : sizes) tot += Arena::round_size(b ? b : 16);
    return tot;
}

size_t PackedScene::device_bytes() const { return scene_footprint(T, N, M, NL, tex_size, wide_stride); }

int wide_records(const ppm_scene* s) {
    std::vector<uint4> nodes, wide, wide_pos;
    int stride = 0;
    if (pack_nodes(s, nodes, wide, wide_pos, &stride)) return std::max(1, s->num_triangles);
    return stride;
}

SceneDev::~SceneDev() {
    for (void* p : allocs) {
        if (arena) arena->release(p, last_use);
        else (void)hipFree(p);
    }
}

int pack_scene(const ppm_scene* s, PackedScene* out) {
    int rc;
    if ((rc = pack_nodes(s, out->nodes, out->wide, out->wide_pos, &out->wide_stride))) return rc;
    out->leaf_box_stored = pack_triangles(s, out->isect, out->shade);
    out->T = s->num_triangles;
    out->N = s->num_nodes;
    out->M = s->num_materials;
    out->NL = s->num_lights;
    out->mats.resize((size_t)out->M * 2);
    for (int m = 0; m < out->M; m++) {
        const ppm_material& mt = s->materials[m];
        out->mats[2 * m] = make_float4(mt.color[0], mt.color[1], mt.color[2], mt.eta);
        out->mats[2 * m + 1] = make_float4((float)mt.brdf, mt.g, 0.0f, 0.0f);
    }
    out->light_tri.assign(s->light_triangle, s->light_triangle + out->NL);
    out->light_cdf.assign(s->light_cdf, s->light_cdf + out->NL);
    out->light_area = s->light_area;
    for (int k = 0; k < 4; k++) out->bsphere[k] = s->bsphere[k];
    // the largest node bound magnitude, capped at 2^17: a bound beyond it packs as an infinity or
    // as +-65504, more than half its magnitude inward of it (outward for the box)
    float bmax = 0.0f;
    for (size_t k = 0; k < (size_t)s->num_nodes * 6; k++) {
        const float a = std::fabs(s->node_bounds[k]);
        if (a > bmax) bmax = std::fmin(a, 0x1p17f);
    }
    out->slab_eps = std::fmax(std::ldexp(bmax, -22), 0x1p-126f);
    out->camera = s->camera;
    out->tex.clear();
    out->volume.clear();
    out->tex_size = 0;
    if (s->texcoords && s->textures && s->texture_size > 0) {
        const int T = s->num_triangles, S = s->texture_size;
        out->tex_size = S;
        out->tex.resize((size_t)T * 2);
        for (int t = 0; t < T; t++) {
            const float* u = s->texcoords + (size_t)t * 6;
            out->tex[(size_t)t * 2] = make_float4(u[0], u[1], u[2], u[3]);
            out->tex[(size_t)t * 2 + 1] = make_float4(u[4], u[5], 0.0f, 0.0f);
        }
        const size_t n = (size_t)out->M * S * S;
        out->volume.resize(n);
        for (size_t i = 0; i < n; i++)
            out->volume[i] = make_uchar4(s->textures[3 * i], s->textures[3 * i + 1], s->textures[3 * i + 2], 255);
    }
    return PPM_OK;
}

int upload_scene(const PackedScene& ps, Arena* arena, hipStream_t up, hipStream_t last_use,
                 std::shared_ptr<SceneDev>* out, std::string* err) {
    auto sd = std::make_shared<SceneDev>();
    sd->arena = arena;
    sd->last_use = last_use;
    sd->T = ps.T;
    sd->N = ps.N;
    sd->M = ps.M;
    sd->NL = ps.NL;
    auto put = [&](const void* src, size_t bytes, void** dst) -> int {
        const size_t n = bytes ? bytes : 16;
        hipError_t e = arena ? arena->alloc(dst, n) : hipMalloc(dst, n);
        if (e == hipSuccess) {
            sd->allocs.push_back(*dst);
            sd->bytes += arena ? Arena::round_size(n) : n;
            if (bytes) e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, up);
        }
        if (e != hipSuccess) {
            *err = std::string("scene upload: ") + hipGetErrorString(e);
            return e == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP;
        }
        return PPM_OK;
    };
    void *d_nodes, *d_wide, *d_wpos, *d_isect, *d_shade, *d_mats, *d_ltri, *d_lcdf, *d_tex = nullptr, *d_vol = nullptr;
    int rc;
    if ((rc = put(ps.nodes.data(), ps.nodes.size() * 16, &d_nodes)) ||
        (rc = put(ps.wide.data(), ps.wide.size() * 16, &d_wide)) ||
        (rc = put(ps.wide_pos.data(), ps.wide_pos.size() * 16, &d_wpos)) ||
        (rc = put(ps.isect.data(), ps.isect.size() * 16, &d_isect)) ||
        (rc = put(ps.shade.data(), ps.shade.size() * 16, &d_shade)) ||
        (rc = put(ps.mats.data(), ps.mats.size() * 16, &d_mats)) ||
        (rc = put(ps.light_tri.data(), ps.light_tri.size() * 4, &d_ltri)) ||
        (rc = put(ps.light_cdf.data(), ps.light_cdf.size() * 4, &d_lcdf)))
        return rc;
    if (ps.tex_size > 0 && ((rc = put(ps.tex.data(), ps.tex.size() * 16, &d_tex)) ||
                            (rc = put(ps.volume.data(), ps.volume.size() * 4, &d_vol))))
        return rc;
    DevScene& sc = sd->s;
    std::memset(&sc, 0, sizeof(sc));
    sc.nodes = (const uint4*)d_nodes;
    sc.wide = (const uint4*)d_wide;
    sc.wide_pos = (const uint4*)d_wpos;
    sc.wide_stride = ps.wide_stride;
    sc.tri_isect = (const float4*)d_isect;
    sc.tri_shade = (const float4*)d_shade;
    sc.mats = (const float4*)d_mats;
    sc.light_tri = (const int*)d_ltri;
    sc.light_cdf = (const float*)d_lcdf;
    sc.num_nodes = ps.N;
    sc.num_lights = ps.NL;
    sc.light_area = ps.light_area;
    sc.bsphere_x = ps.bsphere[0];
    sc.bsphere_y = ps.bsphere[1];
    sc.bsphere_z = ps.bsphere[2];
    sc.bsphere_r = ps.bsphere[3];
    sc.tri_tex = (const float4*)d_tex;
    sc.volume = (const uchar4*)d_vol;
    sc.tex_size = ps.tex_size;
    sc.num_materials = ps.M;
    sc.material_rcp = 1.0f / (float)ps.M;
    sc.slab_eps = ps.slab_eps;
    sc.leaf_box_stored = ps.leaf_box_stored;
    *out = sd;
    return PPM_OK;
}

int create_context(ppm_ctx* c, std::shared_ptr<SceneDev> scene, const ppm_camera& k, const ppm_params* p,
                   hipStream_t stream, hipStream_t upload_st, Arena* arena, uint32_t passes_left) {
    c->prm = *p;
    c->arena = arena;
    c->scene = scene;
    c->Q = p->image_width * p->image_height;
    c->P = p->photon_width * p->photon_height;
    c->H = p->hash_width * p->hash_height;
    c->T = scene->T;
    c->N = scene->N;
    c->M = scene->M;
    c->NL = scene->NL;
    uint32_t xs[4];
    std::memcpy(xs, p->xorshift, sizeof(xs));

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_err(c, PPM_E_NODEV, "no HIP device");
    if (p->device < 0 || p->device >= ndev) return set_err(c, PPM_E_INVAL, "device ordinal out of range");
    c->device = p->device;
    HIP_TRY(c, hipSetDevice(c->device));
    static_scratch_address(c->device);
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else {
        HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        HIP_TRY(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HIP_TRY(c, hipEventCreateWithFlags(&c->ev_ahead, hipEventDisableTiming));
    }
    hipStream_t up = upload_st ? upload_st : c->stream;
    HIP_TRY(c, hipEventCreate(&c->ev0));
    HIP_TRY(c, hipEventCreate(&c->ev1));

    int rc;
    Job& J = c->job;
    std::memset(&J, 0, sizeof(J));
    J.s = scene->s;
    J.cam = DevCamera{k.origin[0], k.origin[1], k.origin[2], k.u[0], k.u[1], k.u[2], k.v[0], k.v[1], k.v[2],
                      k.w[0],      k.w[1],      k.w[2],      k.half_width, k.half_height, k.distance};
    J.ep.width = p->image_width;
    J.ep.height = p->image_height;
    J.ep.focal_length = p->focal_length;
    J.ep.aperture = p->aperture;
    J.ep.max_path = p->max_path_length;
    J.ep.fixed_radius = p->fixed_radius;
    J.ep.footprint = p->initial_footprint;
    J.ep.footprint_width = (float)p->image_width;
    J.pp.count = c->P;
    J.pp.width = p->photon_width;
    J.pp.max_path = p->max_path_length;
    J.pp.rep_limit = rep_limit(*p);
    J.gp.count = c->Q;
    J.gp.alpha = p->alpha;
    J.gp.draw = (p->flags & PPM_FLAG_NO_DRAW) ? 0 : 1;
    J.tl.width = p->image_width;
    J.tl.height = p->image_height;
    J.photon_blocks = photon_blocks(p);
    J.begin_blocks = 1;
    // one-wave blocks for the photon walk (measured: 1% faster than 256 at 1024^2, 3% on the 2.26M
    // triangle terrain), 8x8-pixel waves; the tuner may change them
    set_launch_shape(c, kTraceBlockDefault, 3, kWalkStack);
    // the eye stage beside the photon stage pays for its stream synchronisation from about 512^2
    // pixels on (measured: +4% at 1024^2, -7% at 256^2)
    c->eye_overlap = c->Q >= (1 << 18);
    c->use_graph = (p->flags & PPM_FLAG_GRAPH) != 0;

    const size_t Q = c->Q, P = c->P, H = c->H;
    HitState& hs = J.hs;
    PhotonState& ps = J.ps;
    hs.q = (int)Q;
    if ((rc = dev_alloc(c, &hs.pos, 2 * Q)) || (rc = dev_alloc(c, &hs.refl, 2 * Q)) ||
        (rc = dev_alloc(c, &hs.nrm, 2 * Q)) || (rc = dev_alloc(c, &hs.rng, 2 * Q)) || (rc = dev_alloc(c, &hs.em, 2 * Q)) ||
        (rc = dev_alloc(c, &hs.tau_r, Q)) || (rc = dev_alloc(c, &hs.e_n, Q)) || (rc = dev_alloc(c, &hs.tri, 2 * Q)) ||
        (rc = dev_alloc(c, &hs.mat_t, 2 * Q)) || (rc = dev_alloc(c, &hs.rgba8, Q)) ||
        (rc = dev_alloc(c, &hs.partial, (size_t)max_tile_blocks(p) * 6)) ||
        (rc = dev_alloc(c, &hs.tile_cost, (size_t)max_tile_blocks(p))) ||
        (rc = dev_alloc(c, &hs.tile_order, (size_t)max_tile_blocks(p))) || (rc = dev_alloc(c, &ps.pos, 2 * P)) ||
        (rc = dev_alloc(c, &ps.flux, 2 * P)) || (rc = dev_alloc(c, &ps.dir, 2 * P)) || (rc = dev_alloc(c, &ps.rng, P)) ||
        (rc = dev_alloc(c, &ps.state, P)) || (rc = dev_alloc(c, &ps.slot, P)) ||

        (rc = dev_alloc(c, &J.h.slots, 6 * H)) || (rc = dev_alloc(c, &J.g, 1)) ||
        (rc = dev_alloc(c, &c->resolve_lin, Q)) || (rc = dev_alloc(c, &c->resolve_map, Q)))
        return rc;
    J.h.size = c->H;
    J.h.width = p->hash_width;
    J.h.height = p->hash_height;
    J.h.num_f = (float)c->H;

    // per-pixel and per-photon generator states (the first eye stage, buffer 0, reads its states
    // from buffer 1), then the stream's state for the AAOffset draws
    std::vector<uint4> starts;
    seed_starts(xs, Q + P, starts);
    uint4* d_starts = nullptr;
    if ((rc = dev_alloc(c, &d_starts, starts.size()))) return rc;
    HIP_TRY(c, hipMemcpyAsync(d_starts, starts.data(), starts.size() * sizeof(uint4), hipMemcpyHostToDevice, up));
    hipLaunchKernelGGL(k_seed, dim3((unsigned)((starts.size() + 255) / 256)), dim3(256), 0, up, hs.rng + Q, (int)Q,
                       ps.rng, (int)P, (const uint4*)d_starts, kSeedChunk);
    HIP_TRY(c, hipGetLastError());
    DevGlobals g0;
    std::memset(&g0, 0, sizeof(g0));
    std::memcpy(g0.xs, xs, sizeof(xs));
    g0.passes_left = passes_left;
    HIP_TRY(c, hipMemcpyAsync(J.g, &g0, sizeof(g0), hipMemcpyHostToDevice, up));
    const int n = (int)std::max(std::max(Q, P), H);
    hipLaunchKernelGGL(k_init_state, dim3((n + 255) / 256), dim3(256), 0, up, J);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(up));
    return PPM_OK;
}

void release_context(ppm_ctx* c) {
    (void)hipSetDevice(c->device);
    for (auto& pe : c->ev_pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    c->ev_pending.clear();
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    for (int k = 0; k < kPassKinds; k++)
        if (c->graph_pass[k]) {
            (void)hipGraphExecDestroy(c->graph_pass[k]);
            c->graph_pass[k] = nullptr;
        }
    for (void* p : c->allocs) {
        if (c->arena) c->arena->release(p, c->stream);
        else (void)hipFree(p);
    }
    c->allocs.clear();
    c->device_bytes = 0;
    c->scene.reset();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_ahead) (void)hipEventDestroy(c->ev_ahead);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    c->ev0 = c->ev1 = c->ev_fork = c->ev_join = c->ev_ahead = nullptr;
    c->side = nullptr;
    c->stream = nullptr;
}

}  // namespace ppm_internal

extern "C" {

void ppm_params_default(ppm_params* p, int32_t res) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->image_width = p->image_height = res;
    p->photon_width = p->photon_height = res;
    p->hash_width = p->hash_height = res;
    p->max_path_length = 10;
    p->alpha = 0.7f;
    p->initial_footprint = 2.5f;
    p->fixed_radius = 0.0f;
    p->focal_length = 13.0f;
    p->aperture = 0.0f;
    p->xorshift[0] = 123456789u;
    p->xorshift[1] = 362436069u;
    p->xorshift[2] = 521288629u;
    p->xorshift[3] = 88675123u;
    p->device = 0;
    p->flags = 0;
}

int ppm_create(const ppm_scene* scene, const ppm_params* params, ppm_ctx** out) {
    if (!out) return set_err(nullptr, PPM_E_INVAL, "null output");
    *out = nullptr;
    int rc;
    if ((rc = ppm_internal::validate(scene, params))) return rc;
    ppm_ctx* c = new (std::nothrow) ppm_ctx();
    if (!c) return set_err(nullptr, PPM_E_OOM, "out of host memory");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete c;
        return set_err(nullptr, PPM_E_NODEV, "no HIP device");
    }
    if (params->device >= ndev || hipSetDevice(params->device) != hipSuccess) {
        delete c;
        return set_err(nullptr, PPM_E_INVAL, "device ordinal out of range");
    }
    ppm_internal::PackedScene ps;
    if ((rc = ppm_internal::pack_scene(scene, &ps))) {
        delete c;
        return rc;
    }
    std::shared_ptr<ppm_internal::SceneDev> sd;
    std::string err;
    rc = ppm_internal::upload_scene(ps, nullptr, nullptr, nullptr, &sd, &err);
    if (rc == PPM_OK && hipStreamSynchronize(nullptr) != hipSuccess) {
        rc = PPM_E_HIP;
        err = "scene upload failed";
    }
    if (rc != PPM_OK) {
        delete c;
        return set_err(nullptr, rc, err);
    }
    rc = ppm_internal::create_context(c, sd, scene->camera, params, nullptr, nullptr, nullptr, 0xffffffffu);
    if (rc != PPM_OK) {
        g_create_error = c->err.empty() ? g_create_error : c->err;
        ppm_internal::release_context(c);
        delete c;
        return rc;
    }
    *out = c;
    return PPM_OK;
}

int ppm_pass(ppm_ctx* c, uint32_t npasses) {
    if (!c) return PPM_E_INVAL;
    if (!c->own_stream) return set_err(c, PPM_E_INVAL, "this context belongs to a batch; use ppm_batch_pass");
    HIP_TRY(c, hipSetDevice(c->device));
    if (npasses == 0) return PPM_OK;
    return run_passes(c, c->job, npasses, true);
}

int ppm_pass_static(ppm_ctx* c, uint32_t npasses) {
    if (!c) return PPM_E_INVAL;
    if (!c->own_stream) return set_err(c, PPM_E_INVAL, "this context belongs to a batch; use ppm_batch_pass");
    HIP_TRY(c, hipSetDevice(c->device));
    // the eye kernel's per-block min/max of hit positions
    const size_t eye_bytes = (size_t)c->job.tile_blocks * 6 * 4;
    if (eye_bytes > kStaticScratchBytes)
        return set_err(c, PPM_E_INVAL,
                       "ppm_pass_static: this context's partial results need " + std::to_string(eye_bytes) +
                           " bytes of scratch; the static scratch holds " + std::to_string(kStaticScratchBytes) +
                           " (use ppm_pass for images above about 11M pixels)");
    void* base = nullptr;
    int rc = static_scratch(c, &base);
    if (rc) return rc;
    if (npasses == 0) return PPM_OK;
    Job J = c->job;
    J.hs.partial = static_cast<float*>(base);
    // per-kernel timing would create events: it stays off for these passes
    const bool prof = c->profiling;
    c->profiling = false;
    rc = run_passes(c, J, npasses, false);
    c->profiling = prof;
    return rc;
}

int ppm_sync(ppm_ctx* c) {
    if (!c) return PPM_E_INVAL;
    HIP_TRY(c, sync_streams(c));
    return PPM_OK;
}

float ppm_last_pass_ms(ppm_ctx* c) {
    if (!c || !c->timed) return -1.0f;
    if (hipEventSynchronize(c->ev1) != hipSuccess) return -1.0f;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) return -1.0f;
    return ms;
}

size_t ppm_buffer_bytes(const ppm_ctx* c, ppm_buf which) {
    if (!c) return 0;
    const size_t Q = c->Q, P = c->P, H = c->H;
    switch (which) {
        case PPM_BUF_RADIANCE:
        case PPM_BUF_TONEMAPPED:
        case PPM_BUF_HIT_POSITION:
        case PPM_BUF_HIT_REFLECTANCE:
        case PPM_BUF_HIT_NORMAL:
        case PPM_BUF_EYE_RNG:
        case PPM_BUF_FLUX_RADIUS:
        case PPM_BUF_EMISSION_COUNT: return Q * 16;
        case PPM_BUF_RGBA8: return Q * 4;
        case PPM_BUF_HIT_TRIANGLE: return Q * 8;
        case PPM_BUF_HIT_T: return Q * 4;
        case PPM_BUF_PHOTON_POSITION:
        case PPM_BUF_PHOTON_FLUX:
        case PPM_BUF_PHOTON_DIRECTION:
        case PPM_BUF_PHOTON_RNG: return P * 16;
        case PPM_BUF_PHOTON_STATE: return P * 12;
        case PPM_BUF_PHOTON_SLOT: return P * 4;
        case PPM_BUF_SLOT_COUNT:
        case PPM_BUF_SLOT_REP: return H * 4;
        case PPM_BUF_COUNTERS: return sizeof(ppm_counters);
        case PPM_BUF_GRID: return sizeof(ppm_grid);
        default: return 0;
    }
}

int ppm_readback(ppm_ctx* c, ppm_buf which, void* dst, size_t bytes) {
    if (!c || !dst) return PPM_E_INVAL;
    const size_t need = ppm_buffer_bytes(c, which);
    if (need == 0) return set_err(c, PPM_E_INVAL, "unknown buffer");
    if (bytes != need) return set_err(c, PPM_E_INVAL, "byte count does not match the buffer size");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, sync_streams(c));
    const size_t Q = c->Q, P = c->P, H = c->H;
    const Job& J = c->job;
    auto d2h = [&](const void* src, size_t n) -> int {
        HIP_TRY(c, hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
        return PPM_OK;
    };
    DevGlobals g;
    HIP_TRY(c, hipMemcpy(&g, J.g, sizeof(g), hipMemcpyDeviceToHost));
    // the hit points of the latest eye pass (before any: buffer 1, which holds the seeds)
    const size_t cb = (size_t)((g.eye_samples + 1) & 1) * Q;
    switch (which) {
        case PPM_BUF_RADIANCE:
        case PPM_BUF_TONEMAPPED:
        case PPM_BUF_RGBA8: {
            if (g.frame == 0) return set_err(c, PPM_E_INVAL, "no pass has run");
            hipLaunchKernelGGL(k_resolve, dim3((c->Q + 255) / 256), dim3(256), 0, c->stream, c->Q, J.hs, J.g,
                               c->resolve_lin, c->resolve_map);
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, sync_streams(c));
            if (which == PPM_BUF_RADIANCE) return d2h(c->resolve_lin, Q * 16);
            if (which == PPM_BUF_TONEMAPPED) return d2h(c->resolve_map, Q * 16);
            return d2h(J.hs.rgba8, Q * 4);
        }
        case PPM_BUF_HIT_POSITION: return d2h(J.hs.pos + cb, Q * 16);
        case PPM_BUF_HIT_REFLECTANCE: return d2h(J.hs.refl + cb, Q * 16);
        case PPM_BUF_HIT_NORMAL: return d2h(J.hs.nrm + cb, Q * 16);
        case PPM_BUF_EYE_RNG: return d2h(J.hs.rng + cb, Q * 16);
        case PPM_BUF_FLUX_RADIUS: return d2h(J.hs.tau_r, Q * 16);
        case PPM_BUF_EMISSION_COUNT: return d2h(J.hs.e_n, Q * 16);
        case PPM_BUF_HIT_TRIANGLE:
        case PPM_BUF_HIT_T: {
            std::vector<int> tri(Q);
            std::vector<float2> mt(Q);
            HIP_TRY(c, hipMemcpy(tri.data(), J.hs.tri + cb, Q * 4, hipMemcpyDeviceToHost));
            HIP_TRY(c, hipMemcpy(mt.data(), J.hs.mat_t + cb, Q * 8, hipMemcpyDeviceToHost));
            if (which == PPM_BUF_HIT_T) {
                float* o = (float*)dst;
                for (size_t i = 0; i < Q; i++) o[i] = mt[i].y;
            } else {
                int32_t* o = (int32_t*)dst;
                for (size_t i = 0; i < Q; i++) {
                    o[2 * i] = tri[i];
                    o[2 * i + 1] = tri[i] < 0 ? -1 : (int32_t)mt[i].x;
                }
            }
            return PPM_OK;
        }
        // the photon half of the last pass that ran (both hold the initial records before any)
        case PPM_BUF_PHOTON_POSITION: return d2h(J.ps.pos + (size_t)g.half * P, P * 16);
        case PPM_BUF_PHOTON_FLUX: return d2h(J.ps.flux + (size_t)g.half * P, P * 16);
        case PPM_BUF_PHOTON_DIRECTION: return d2h(J.ps.dir + (size_t)g.half * P, P * 16);
        case PPM_BUF_PHOTON_RNG: return d2h(J.ps.rng, P * 16);
        case PPM_BUF_PHOTON_STATE: {
            std::vector<int2> st(P);
            HIP_TRY(c, hipMemcpy(st.data(), J.ps.state, P * 8, hipMemcpyDeviceToHost));
            int32_t* o = (int32_t*)dst;
            for (size_t i = 0; i < P; i++) {
                o[3 * i] = st[i].x;
                o[3 * i + 1] = st[i].y & 0xffff;
                o[3 * i + 2] = (st[i].y >> 16) & 1;
            }
            return PPM_OK;
        }
        case PPM_BUF_PHOTON_SLOT: return d2h(J.ps.slot, P * 4);
        case PPM_BUF_SLOT_COUNT:
        case PPM_BUF_SLOT_REP: {
            // the table of the last pass that ran (passes take the three in turn)
            std::vector<uint32_t> rec(2 * H);
            const size_t table = (size_t)g.table;
            HIP_TRY(c, hipMemcpy(rec.data(), J.h.slots + table * 2 * H, H * 8, hipMemcpyDeviceToHost));
            uint32_t* o = (uint32_t*)dst;
            const size_t k = which == PPM_BUF_SLOT_COUNT ? 0 : 1;
            for (size_t i = 0; i < H; i++) o[i] = rec[2 * i + k];
            return PPM_OK;
        }
        case PPM_BUF_COUNTERS: {
            ppm_counters k;
            std::memset(&k, 0, sizeof(k));
            k.frame_count = g.frame;
            k.eye_samples = g.eye_samples;
            folded_photon_counts(g, &k.photons_emitted, &k.num_photons_f);
            k.photon_segments = (uint64_t)c->P * g.frame;
            std::memcpy(dst, &k, sizeof(k));
            return PPM_OK;
        }
        case PPM_BUF_GRID: {
            ppm_grid k;
            const GridScalars& gr = g.grid[(g.eye_samples + 1) & 1];
            for (int i = 0; i < 3; i++) {
                k.bbmin[i] = gr.bbmin[i];
                k.bbmax[i] = gr.bbmax[i];
            }
            k.initial_radius = gr.initial_radius;
            k.grid_scale = gr.grid_scale;
            k.hash_scale = gr.hash_scale;
            std::memcpy(dst, &k, sizeof(k));
            return PPM_OK;
        }
        default: return set_err(c, PPM_E_INVAL, "unknown buffer");
    }
}

int ppm_profile_enable(ppm_ctx* c, int enable) {
    if (!c) return PPM_E_INVAL;
    if (enable != 0 && enable != 1 && ((enable & ~(PPM_PROFILE_KINDS | 0xf)) != 0 || !(enable & PPM_PROFILE_KINDS)))
        return set_err(c, PPM_E_INVAL, "profile: 0, 1 or PPM_PROFILE_KINDS with kernel kind bits");
    HIP_TRY(c, sync_streams(c));
    for (auto& pe : c->ev_pending) {
        c->ev_pool.push_back(pe.second.first);
        c->ev_pool.push_back(pe.second.second);
    }
    c->ev_pending.clear();
    for (int k = 0; k < PPM_KERNEL_KINDS; k++) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
    }
    c->profiling = enable != 0;
    c->prof_kinds = enable == 1 ? 0xfu : (unsigned)enable & 0xfu;
    return PPM_OK;
}

int ppm_profile_read(ppm_ctx* c, double* ms, uint64_t* launches) {
    if (!c) return PPM_E_INVAL;
    HIP_TRY(c, sync_streams(c));
    for (auto& pe : c->ev_pending) {
        float t = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&t, pe.second.first, pe.second.second));
        c->prof_ms[pe.first] += t;
        c->prof_n[pe.first] += 1;
        c->ev_pool.push_back(pe.second.first);
        c->ev_pool.push_back(pe.second.second);
    }
    c->ev_pending.clear();
    for (int k = 0; k < PPM_KERNEL_KINDS; k++) {
        if (ms) ms[k] = c->prof_ms[k];
        if (launches) launches[k] = c->prof_n[k];
    }
    return PPM_OK;
}

int ppm_get_tuning(ppm_ctx* c, ppm_tuning* t) {
    if (!c || !t) return PPM_E_INVAL;
    t->trace_block = c->trace_block;
    t->tile_wave_width = 1 << c->job.tl.wave_shift;
    t->passes_per_graph = c->use_graph ? 1u : 0u;
    t->walk_stack = c->job.walk_stack;
    t->eye_overlap = c->eye_overlap ? 1 : 0;
    t->ms_per_pass = 0.0f;
    return PPM_OK;
}

int ppm_set_tuning(ppm_ctx* c, const ppm_tuning* t) {
    if (!c) return PPM_E_INVAL;
    int sh = 0;
    if (ppm_internal::check_tuning(t, &sh) || t->passes_per_graph > 1)
        return set_err(c, PPM_E_INVAL, "tuning out of range");
    ppm_internal::set_launch_shape(c, t->trace_block, sh, t->walk_stack);
    c->use_graph = t->passes_per_graph > 0;
    c->eye_overlap = t->eye_overlap != 0;
    return PPM_OK;
}

int ppm_autotune(ppm_ctx* c, uint32_t passes_per_trial, ppm_tuning* best) {
    if (!c) return PPM_E_INVAL;
    if (!c->own_stream) return set_err(c, PPM_E_INVAL, "this context belongs to a batch; use ppm_batch_autotune");
    // four eye cadences per trial: shorter trials leave the choice to noise of a few percent
    if (passes_per_trial == 0) passes_per_trial = 4u * (uint32_t)c->prm.max_path_length;
    ppm_tuning start;
    ppm_get_tuning(c, &start);
    auto trial = [&](const ppm_tuning& t, float* ms) -> int {
        int rc;
        if ((rc = ppm_set_tuning(c, &t))) return rc;
        if ((rc = ppm_pass(c, 1))) return rc;  // capture and first-launch costs stay out of the timing
        if ((rc = ppm_pass(c, passes_per_trial))) return rc;
        // an eye stage started ahead inside the trial belongs to it: the timed region ends after it
        if (c->ahead_pending) {
            HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_ahead, 0));
            HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        }
        const float v = ppm_last_pass_ms(c);
        if (v < 0.0f) return set_err(c, PPM_E_HIP, "trial timing failed");
        *ms = v / (float)passes_per_trial;
        return PPM_OK;
    };
    return ppm_internal::autotune(start, false, trial, best);
}

void ppm_destroy(ppm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)sync_streams(c);
    ppm_internal::release_context(c);
    delete c;
}

const char* ppm_last_error(const ppm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

}  // extern "C"
