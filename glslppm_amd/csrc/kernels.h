// Device-side data layout and kernels of the SPPM pass loop on gfx950 (wave64).
//
// HBM layout (all arrays 16-byte aligned, structure-of-arrays so that a wave's 64 lanes read
// 1 KiB contiguous per field):
//   scene:  nodes[6][N+1] uint4 per cube face, in that face's pre-order (+1 padding node): the box
//                                  as six binary16 rounded outward (lo.x lo.y | lo.z hi.x |
//                                  hi.y hi.z), then the miss position of an interior node or
//                                  ~triangle of a leaf (whose miss position is the next one); the
//                                  hit link is the next position, so it is not stored.
//           wide[6][R][4] uint4    the same trees four-wide: a record is four member slots (the
//                                  member's box as above, then its record, ~triangle of a leaf, or
//                                  kWideEmpty), members in pre-order; record 0 is the top
//           wide_pos[6][R] uint4   the members' pre-order positions (N: empty)
//           tri_isect[T][4] float4 (v0, brdf), (v1, leaf min.x), (v2, leaf min.y),
//                                  (leaf min.z, leaf max.xyz): one 64-byte record per triangle
//                                  with the exact box of its leaf, whose last 16 bytes a walk
//                                  reads only when some leaf's box is not the box of its
//                                  triangle's vertices (DevScene::leaf_box_stored)
//           tri_shade[T][4] float4 (n0, 0) (n1, 0) (n2, 0) (gn, material): the shading normals,
//                                  the geometric normal and the material (the vertices are
//                                  tri_isect's)
//           mats[M][2] float4      (rgb, eta), (brdf, g, 0, 0)
//           tri_tex[T][2] float4   (u0, v0, u1, v1), (u2, v2, 0, 0)     only with textures
//           volume[M][S][S] uchar4 RGB8 + pad of every material (S = 512)  only with textures
//   hit points (Q): pos, refl, nrm, rng, (tau, R), (E, N) float4; tri int; (mat, t) float2;
//                   per pixel tile: depth of its deepest eye path, the next eye pass's tile order
//   photons (P):    pos, flux, dir [2][P] float4 (.w = hit normal; pass f writes half f & 1 from
//                   half (f + 1) & 1, so the gather of pass f can read its representatives while
//                   the photon stage of pass f + 1 runs), rng float4; state int2 (triangle,
//                   bounce | emitted << 16); slot int
//   hash (H):       three tables of (count u32, representative u32) per slot (0xffffffff = none);
//                   pass f deposits into table f % 3; a photon's slot is kSlotPending between an
//                   eye pass's photon trace and its k_hash
#pragma once

#include "fmath.h"

// Unused slot of a four-wide record.
constexpr unsigned int kWideEmpty = 0x80000000u;

struct DevScene {
    const uint4* __restrict__ nodes;
    const uint4* __restrict__ wide;      // [6][wide_stride][4] four-wide records
    const uint4* __restrict__ wide_pos;  // [6][wide_stride] pre-order positions of their members
    const float4* __restrict__ tri_isect;
    const float4* __restrict__ tri_shade;
    const float4* __restrict__ mats;
    const int* __restrict__ light_tri;
    const float* __restrict__ light_cdf;
    int num_nodes;
    int num_lights;
    int wide_stride;
    float light_area;
    float bsphere_x, bsphere_y, bsphere_z, bsphere_r;
    // textures (null when no triangle is textured): per triangle (u0, v0, u1, v1), (u2, v2, 0, 0);
    // volume [M][S][S] RGBA8 of the materials' images, sampled like a GL_LINEAR / GL_REPEAT 3D
    // texture; material_rcp = 1.0f / M (the MaterialNumRcp uniform)
    const float4* __restrict__ tri_tex;
    const uchar4* __restrict__ volume;
    int tex_size;
    int num_materials;
    float material_rcp;
    // 2^-22 times the largest node bound magnitude capped at 2^17 (at least 2^-126): scales the
    // slack of the four-wide record test's fused slab values
    float slab_eps;
    // 0: every leaf's box is the box of its triangle's three vertices (min and max per axis, as
    // the reference's build makes it, bvh.cpp:32-57 from mesh.cpp:439-446), taken from the
    // vertices; 1: the leaf boxes are read from the records
    int leaf_box_stored;
};

struct DevCamera {
    float ox, oy, oz;
    float ux, uy, uz;
    float vx, vy, vz;
    float wx, wy, wz;
    float half_w, half_h, dist;
};

// Grid scalars, pass counters and the per-pass scalars of one job, kept in HBM so that no pass
// needs a host round trip and the pass sequence is the same launch sequence every pass (which a
// hipGraph can replay, and which many jobs can share).  k_pass_begin writes the per-pass fields
// from the job's frame count and its own XORShift stream (glslppm.cpp:214-287).
// The grid scalars an eye pass sets up (glslppm.cpp:294-311).
struct GridScalars {
    float bbmin[3];
    float bbmax[3];
    float initial_radius;
    float grid_scale;
    float hash_scale;
    float pad;
};

// What the gather of a pass reads of that pass's uniforms: the gather of pass f runs in the launch
// of pass f + 1's photon stage, after k_pass_begin of pass f + 1 has changed the live fields.
// k_pass_begin of pass f writes snap[f & 1]; np (NumPhotons after pass f, for its draw) is filled
// in by the k_pass_begin of pass f + 1, which folds pass f's emitted counts.
struct GatherSnap {
    int first_pass;  // pass f is frame 0
    int eye_now;     // pass f traced eye rays: the gather copies the new emission into e_n
    int cb;          // the hit-point buffer (and grid) of the latest eye pass at pass f
    int table;       // pass f's hash table (0..2)
    int half;        // the photon half pass f wrote (f & 1)
    float np;        // NumPhotons after pass f
    int pad[2];
};

// Gather modes: the gather of the pass before the one in progress, run by the gather blocks of the
// photon-stage launch, or the gather of the last pass begun, run by k_gather at the end of a pass
// sequence.
constexpr int kGatherLagged = 1;
constexpr int kGatherFlush = 2;

// Launch modes of an eye stage (k_eye_begin, k_eye, k_grid): the stage of the eye pass in progress
// (run before that pass's deposits), or the stage of the job's next eye pass, begun ahead of it.
constexpr int kEyeNow = 1;
constexpr int kEyeAhead = 2;
// k_eye_begin's pass argument when the caller orders its ahead stages itself (the plain loop).
constexpr unsigned int kAnyPass = 0xffffffffu;

// Emitted-flag counters of the pass in progress: k_photon_stage adds each block's count to one of
// kEmitShards counters (block b to shard b % kEmitShards), each in a 128-byte line of its own so
// that no line takes many atomics; the next pass's begin folds them into NumPhotons.
constexpr int kEmitShards = 64;
constexpr int kEmitStride = 32;  // unsigned ints between shards

struct DevGlobals {
    // of the latest eye pass of each parity: the passes after eye pass e read grid[e & 1], as they
    // read that eye pass's hit points (HitState buffers)
    GridScalars grid[2];
    unsigned long long photons_emitted;
    unsigned long long frame;        // passes run (FrameCount)
    unsigned long long eye_samples;  // eye passes begun; the latest one's buffers are (eye_samples - 1) & 1
    float num_photons_f;             // NumPhotons up to the pass before the last begun one
    unsigned int pad2;
    unsigned int pad0;
    unsigned int xs[4];              // XORShift state after the seed grids (AAOffset draws)
    unsigned int passes_left;        // passes this job still runs; 0xffffffff: no limit
    int active;                      // this pass advances the job
    int eye_now;                     // this pass traces eye rays (FrameCount % MaxPathLength == 0)
    int first_pass;                  // this pass is frame 0
    int num_eye_samples;             // the NumEyeSamples uniform of the eye stage in progress
    float aa_x, aa_y;                // AAOffset of the eye stage in progress
    int order_tiles;                 // the tile count tile_order was built for (0: none yet)
    unsigned int eye_stages;         // eye stages begun (k_eye_begin)
    int eye_stage;                   // the eye stage in progress; it writes buffers eye_stage & 1
    int eye_go;                      // the stage in progress and its launches' mode (kEyeNow or
                                     // kEyeAhead; 0: none), set by k_eye_begin, cleared by k_grid
    int defer;                       // this pass's deposits wait for k_hash (its grid is not set up)
    unsigned int eye_stages_done;    // eye stages whose grid is set up (k_grid)
    unsigned int begin_pass;         // the launch sequence's pass number at this pass's k_pass_begin
    int ahead_frames;                // passes from this one to the next eye pass (k_pass_begin)
    int ahead_ok;                    // the job's budget reaches that eye pass (k_pass_begin)
    int table;                       // this pass's hash table (frame f: f % 3)
    int half;                        // the photon half this pass writes (f & 1)
    int lag_pending;                 // the pass before this one is still to be gathered (k_pass_begin)
    int lag_idx;                     // its snap index
    unsigned long long gathered;     // `frame` when k_gather last ran: the pass before it is gathered
    GatherSnap snap[2];              // per pass parity
    // emitted flags of the last begun pass, not yet in num_photons_f / photons_emitted
    unsigned int emit_shard[kEmitShards * kEmitStride];
};

// The eye pass's outputs are double-buffered by eye-pass parity: eye pass e writes buffer e & 1
// (pos, refl, nrm, tri, mat_t, rng, em at [b * q + pixel]) from buffer (e - 1) & 1 (rng, em), and
// the passes until the next eye pass read buffer e & 1, so an eye stage can run while the passes
// before it still read the previous one.  tau_r and e_n (progressive statistics) are single.
struct HitState {
    int q;         // pixels (Q)
    float4* pos;   // [2][Q]
    float4* refl;  // [2][Q]
    float4* nrm;   // [2][Q]
    float4* rng;   // [2][Q]
    float4* em;    // [2][Q] emission average (E of (E, N)); the gather copies it into e_n.xyz
    float4* tau_r;
    float4* e_n;
    int* tri;      // [2][Q]
    float2* mat_t; // [2][Q]
    uchar4* rgba8;
    float* partial;  // [blocks][6] min xyz, max xyz of the eye kernel
    int* tile_cost;  // [tiles] segments of the deepest eye path of each tile in the last eye pass
    int* tile_order; // [tiles] the order k_eye takes the tiles in: deepest paths first
};

// PhotonState::slot of a photon whose deposit waits for k_hash (eye passes).
constexpr int kSlotPending = -2;

struct PhotonState {
    float4* pos;   // [2][P]
    float4* flux;  // [2][P]
    float4* dir;   // [2][P]
    float4* rng;
    int2* state;
    int* slot;
};

struct HashState {
    unsigned int* slots;   // [3][H][2]: count, representative (0xffffffff = none), one 8-byte record
                           // per slot so that a gather reads one line per cell; three tables used in
                           // turn: a pass's photon stage clears the table of the next pass beside its
                           // walks while the gather of the pass before reads its own (hash_table,
                           // next_table)
    int size;          // H
    int width;         // hash_width (power of two)
    int height;
    float num_f;       // H as float (HashNum)
};

struct EyeParams {
    int width, height;
    float focal_length, aperture;
    int max_path;
    float fixed_radius;
    float footprint;
    float footprint_width;
};

struct PhotonParams {
    int count;         // P
    int width;         // photon_width
    int max_path;
    int rep_limit;     // photons with index >= rep_limit are never representatives
};

struct GatherParams {
    int count;         // Q
    float alpha;
    int draw;
};

// Pixels are processed in tiles of 2x2 waves, each wave a (2^wave_shift) x (64 / 2^wave_shift)
// rectangle (8x8 by default), so that a wave's eye rays and hash cells are neighbours; one
// 256-thread block per tile.  The shape changes which pixels share a wave, never a pixel's result.
struct Tiles {
    int width, height;   // image
    int tiles_x;         // ceil(width / tile width)
    int wave_shift;      // log2 of a wave's width in pixels (0..6)
};

// Launch shapes that leave every result unchanged, chosen by the autotuner.
constexpr int kTraceBlockMax = 512;
constexpr int kTraceBlockDefault = 64;

// Stack entries per lane of the four-wide walks (default, most: 32 x 512 lanes x 4 bytes = 64 KiB
// of LDS); deeper walks continue on the pre-order records.
constexpr int kWalkStack = 16;
constexpr int kWalkStackMax = 32;
// Dynamic LDS of a walk launch (k_eye, k_photon_stage): the lanes' stacks.
inline size_t walk_lds_bytes(int walk_stack, int block) { return (size_t)walk_stack * (size_t)block * 4; }

// Everything the kernels need about one job (one scene, camera and state set).  A plain pass loop
// passes it by value; a batch keeps an array of them in HBM and every launch covers all jobs.
struct Job {
    DevScene s;
    DevCamera cam;
    EyeParams ep;
    PhotonParams pp;
    GatherParams gp;
    Tiles tl;
    HitState hs;
    PhotonState ps;
    HashState h;
    DevGlobals* g;
    int tile_blocks;     // pixel tiles (k_eye, k_gather)
    int photon_blocks;   // 256 photon slots per block (k_hash)
    int begin_blocks;    // blocks of k_pass_begin (one)
    int trace_blocks;    // photon slots per block = the launch's block size (k_photon_stage)
    int gather_blocks;   // blocks of k_photon_stage after its trace blocks: the gather of the pass
                         // before, one wave per quarter tile
    int walk_stack;      // stack entries per lane of the walks (LDS: walk_stack x block x 4 bytes)
};

// Block map of a batched launch: entry b = (job, block of that job) for grid block b.
struct BlockRef {
    int job;
    int block;
};

// Kernels of one job: grid = that job's blocks.  The _batch forms take the job array and, for
// the kernels with several blocks per job, the block map of that kernel kind; k_eye_begin_batch
// runs one thread per job, k_grid_batch one block per job.  k_photon_stage runs with any block size
// up to kTraceBlockMax (a multiple of 64) over trace_blocks + gather_blocks blocks per job (its
// batch map lists every job's trace blocks, then every job's gather blocks, so that the gathers are
// dispatched last and fill the tail of the walks); the others with 256 (1024 for k_grid).  k_gather
// is the gather of the last pass begun, launched once at the end of a pass sequence.
__global__ void k_pass_begin(Job J, unsigned int pass);
__global__ void k_eye_begin(Job J, int mode, unsigned int pass, int lag);
__global__ void k_eye(Job J, int mode);
__global__ void k_grid(Job J, int mode);
__global__ void k_photon_stage(Job J);
__global__ void k_hash(Job J);
__global__ void k_gather(Job J);
__global__ void k_pass_begin_batch(const Job* jobs, const BlockRef* map, unsigned int pass, int defer_all);
__global__ void k_eye_begin_batch(const Job* jobs, int n_jobs, int mode, unsigned int pass, int lag);
__global__ void k_eye_batch(const Job* jobs, const BlockRef* map, int mode);
__global__ void k_grid_batch(const Job* jobs, int mode);
__global__ void k_photon_stage_batch(const Job* jobs, const BlockRef* map);
__global__ void k_hash_batch(const Job* jobs, const BlockRef* map);
__global__ void k_gather_batch(const Job* jobs, const BlockRef* map);
__global__ void k_batch_status(const Job* jobs, int n_jobs, unsigned long long* out);
__global__ void k_resolve(int count, HitState hs, const DevGlobals* g, float4* linear, float4* mapped);
__global__ void k_init_state(Job J);
__global__ void k_seed(float4* eye, int q, float4* photon, int p, const uint4* starts, int chunk);
