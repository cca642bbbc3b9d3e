// Internal: the per-job context shared by the plain pass loop (runtime.hip) and the batched and
// served forms (batch.hip).  One ppm_ctx owns one scene, one camera and one set of pass state in
// HBM; its kernels see it as a Job record.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ppm.h"
#include "kernels.h"

namespace ppm_internal {

// Caching device allocator for contexts that come and go while others run (batches, serving).
// hipFree waits for the whole device, so freed blocks are kept and reused; a block is handed out
// again only after the event recorded when it was released has completed, i.e. after every
// launch queued before the release has finished with it.  Sizes are rounded up to a multiple of
// an eighth of the next power of two, so that blocks of similar jobs are interchangeable.
class Arena {
public:
    explicit Arena(int device) : device_(device) {}
    ~Arena();
    hipError_t alloc(void** out, size_t bytes);
    // Returns a block whose last use is queued on `last_use` (null: no use pending); the block is
    // handed out again once everything queued there so far has finished.
    void release(void* p, hipStream_t last_use);
    size_t held_bytes() const { return held_; }      // in use + cached
    size_t in_use_bytes() const { return in_use_; }
    // Frees cached blocks whose release event has completed until at most `target` bytes are held.
    void trim(size_t target);
    static size_t round_size(size_t bytes);

private:
    struct Cached {
        void* p;
        hipEvent_t ev;
    };
    int device_;
    size_t held_ = 0, in_use_ = 0;
    std::map<size_t, std::vector<Cached>> free_;
    std::map<void*, size_t> size_of_;
};

// A scene in the form the kernels read (CGPURT::CreateTextures, gpurt.cpp:47-552), on the host.
struct PackedScene {
    std::vector<uint4> nodes;
    std::vector<uint4> wide, wide_pos;  // four-wide records [6][wide_stride][4] and their members' positions
    int wide_stride = 1;
    std::vector<float4> isect, shade, mats, tex;
    std::vector<uchar4> volume;
    int tex_size = 0;
    std::vector<int> light_tri;
    std::vector<float> light_cdf;
    int T = 0, N = 0, M = 0, NL = 0;
    float light_area = 0.0f;
    float bsphere[4] = {0, 0, 0, 0};
    float slab_eps = 0x1p-126f;  // DevScene::slab_eps
    int leaf_box_stored = 0;     // DevScene::leaf_box_stored
    ppm_camera camera;
    size_t device_bytes() const;
};

// A packed scene in HBM, shared by every job that renders it (jobs differ in camera and
// parameters).  Its blocks go back to the arena (or are freed) when the last job lets go of it.
struct SceneDev {
    DevScene s;
    int T = 0, N = 0, M = 0, NL = 0;
    Arena* arena = nullptr;
    hipStream_t last_use = nullptr;  // the stream whose passes read the scene
    std::vector<void*> allocs;
    size_t bytes = 0;
    ~SceneDev();
};

}  // namespace ppm_internal

// What one launch sequence of the plain loop does (queue_pass): a photon pass, an eye pass that
// runs its own eye stage, an eye pass whose eye stage ran ahead, and that ahead stage alone.
enum PassKind { kPhotonPass = 0, kEyePass = 1, kEyePassReady = 2, kEyeStageAhead = 3, kPassKinds = 4 };

struct ppm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // the stream passes run on
    bool own_stream = true;         // false when the stream belongs to a batch
    // an eye pass's eye stage runs on `side` while the photons fly on `stream` (joined before the
    // deposits): only contexts that own their stream have one
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // the eye stage of the next eye pass, started on `side` after an eye pass (ppm_pass)
    hipEvent_t ev_ahead = nullptr;
    bool ahead_due = false;      // an eye pass has passed and the next one's stage is not queued yet
    bool ahead_pending = false;  // that stage is queued on `side`; the next eye pass waits for it
    ppm_internal::Arena* arena = nullptr;  // null: hipMalloc / hipFree
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    ppm_params prm;
    int Q = 0, P = 0, H = 0, N = 0, T = 0, M = 0, NL = 0;
    Job job;
    std::shared_ptr<ppm_internal::SceneDev> scene;
    float4* resolve_lin = nullptr;
    float4* resolve_map = nullptr;
    std::vector<void*> allocs;
    size_t device_bytes = 0;
    // per-kernel event timing (ppm_profile_enable)
    bool profiling = false;
    unsigned prof_kinds = 0;  // kernel kinds timed while profiling (bit per PPM_KERNEL_*)
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[PPM_KERNEL_KINDS] = {0, 0, 0, 0};
    uint64_t prof_n[PPM_KERNEL_KINDS] = {0, 0, 0, 0};
    uint64_t frame = 0;             // passes queued by ppm_pass (the plain loop's eye cadence)
    // a batch job's frame count and pass budget as its k_pass_begin advances them, pass by queued
    // pass (the batch launches eye-stage work only in passes where some job can have it)
    uint64_t batch_frame = 0;
    uint32_t batch_left = 0;
    int trace_block = kTraceBlockDefault;  // launch shape (ppm_set_tuning)
    bool use_graph = false;         // replay passes as hipGraphs (PPM_FLAG_GRAPH)
    bool eye_overlap = false;       // eye stages on `side`, ahead of their eye passes (ppm_set_tuning)
    // hipGraphs of one pass of each kind (PPM_FLAG_GRAPH)
    hipGraphExec_t graph_pass[kPassKinds] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
};

// NumPhotons and the exact emitted count as a reader sees them after the passes queued so far: the
// last pass's per-wave counts are still in the job's counters, and are folded in here exactly as
// the next pass's k_pass_begin folds them (NumPhotons += int(sum) in float, glslppm.cpp:396-397).
inline void folded_photon_counts(const DevGlobals& g, uint64_t* emitted, float* num_photons) {
    unsigned int sum = 0;
    for (int k = 0; k < kEmitShards; k++) sum += g.emit_shard[k * kEmitStride];
    *emitted = g.photons_emitted + sum;
    *num_photons = g.num_photons_f + (float)(int)sum;
}

// Waits for everything queued on a context's streams.
inline hipError_t sync_streams(ppm_ctx* c) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && c->side) e = hipStreamSynchronize(c->side);
    return e;
}

// The batch (ppm_batch_*): resident jobs advanced by one launch per stage per pass.
struct ppm_batch {
    int device = 0;
    hipStream_t stream = nullptr;   // passes
    hipStream_t side = nullptr;     // the eye stage of every pass, beside the photon stage
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // plain launches with eye_overlap: eye stages begun ahead of their eye passes run on `ahead`;
    // ring[p % kAheadRing] is recorded after the ahead launches of pass p, and a pass p with eye
    // passes waits (before its k_hash) for those of pass p - lag (lag + 1 = the shortest eye cadence
    // of the jobs)
    hipStream_t ahead = nullptr;
    static constexpr int kAheadRing = 32;
    hipEvent_t ring[kAheadRing] = {};
    uint32_t pass_no = 0;           // ahead-mode passes queued
    uint32_t ring_from = 0;         // the first pass whose ring event may be waited for
    int lag = 0;
    hipStream_t upload = nullptr;   // scene uploads and job initialisation
    ppm_internal::Arena* arena = nullptr;
    uint32_t passes_per_graph = 1;  // 0: plain launches
    int trace_block = kTraceBlockDefault;  // launch shape shared by every job (ppm_batch_set_tuning)
    int wave_shift = 3;
    int walk_stack = kWalkStack;
    bool eye_overlap = true;        // eye stage on `side` every pass
    std::vector<ppm_ctx*> slots;    // job id -> context (null once removed)
    // device tables of the current job set
    bool dirty = true;
    int n_jobs = 0;
    Job* d_jobs = nullptr;
    BlockRef* d_map[4] = {nullptr, nullptr, nullptr, nullptr};  // tiles, photon slots, hash clears, rays
    int n_blocks[4] = {0, 0, 0, 0};
    hipGraphExec_t graph = nullptr;
    struct Retired {
        hipGraphExec_t g;
        hipEvent_t ev;
    };
    std::vector<Retired> retired;
    std::vector<int32_t> table_ids;  // job id of each entry of d_jobs
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // recorded after the first pass of the latest ppm_batch_pass call: its k_pass_begin_batch has
    // folded every job's last emitted counts, so a job that has stopped no longer changes them
    hipEvent_t ev_begun = nullptr;
    bool timed = false;
    std::string err;
};


namespace ppm_internal {

// Adds a job that renders an already resident scene (the server's shared scenes).
int batch_add_resident(ppm_batch* b, std::shared_ptr<SceneDev> scene, const ppm_camera& cam, const ppm_params* p,
                       uint32_t passes, int32_t* job);
// Rebuilds the batch tables if jobs were added or removed.
int batch_prepare(ppm_batch* b);
// Whether a job still has passes of its budget left after the passes queued so far.
bool batch_job_has_passes(const ppm_batch* b, int32_t job);
// Queues, after the passes queued so far, a copy of every table job's (frame, passes_left) into
// host_out[2 i], host_out[2 i + 1] (pinned, >= 2 x jobs entries; dev_tmp the same in HBM) and
// records `done` after it.  *ids receives the job id of each entry.
int batch_snapshot(ppm_batch* b, unsigned long long* dev_tmp, unsigned long long* host_out, hipEvent_t done,
                   std::vector<int32_t>* ids);
// Copies a finished job's results on the upload stream while passes of other jobs run: linear
// and tone-mapped radiance are resolved there first.  Every pointer may be null.
int batch_fetch(ppm_batch* b, int32_t job, float4* radiance, float4* flux_radius, float4* emission_count,
                uchar4* rgba8, ppm_counters* counters);

int validate(const ppm_scene* s, const ppm_params* p);
int check_scene_public(const ppm_scene* s);
int validate_params(const ppm_params* p);
// Host packing of a scene (checks the link table).  Error text in create_error().
int pack_scene(const ppm_scene* s, PackedScene* out);
// Uploads a packed scene; complete when `up` has been synchronised (the caller keeps `ps` alive
// until then).  last_use: the stream whose passes will read it.
int upload_scene(const PackedScene& ps, Arena* arena, hipStream_t up, hipStream_t last_use,
                 std::shared_ptr<SceneDev>* out, std::string* err);
// Creates a context's pass state for a resident scene and a camera.  stream: where passes will
// run (null: a stream of its own).  Uploads and initialisation run on `upload` (null: the pass
// stream) and are complete on return.  passes_left: the device-side pass budget (0xffffffff:
// unlimited).
int create_context(ppm_ctx* c, std::shared_ptr<SceneDev> scene, const ppm_camera& cam, const ppm_params* p,
                   hipStream_t stream, hipStream_t upload, Arena* arena, uint32_t passes_left);
// Releases the device state.  Arena blocks become reusable once the work queued on the
// context's stream so far has finished; other blocks are freed at once.
void release_context(ppm_ctx* c);
// HBM bytes of a context's pass state, and of a scene, as the arena allocates them.
size_t state_footprint(const ppm_params* p);
size_t scene_footprint(int T, int N, int M, int NL, int tex_size, int wide_stride);
// Four-wide records per face of a scene's tree (the packing's wide_stride; T if the tree is invalid).
int wide_records(const ppm_scene* s);
int set_error(ppm_ctx* c, int code, const std::string& msg);
const std::string& create_error();
// Blocks of each kind for one job.
int tile_blocks(int width, int height, int wave_shift);
int max_tile_blocks(const ppm_params* p);  // over every tile shape
int photon_blocks(const ppm_params* p);
int gather_blocks(int tiles, int block);
// Checks a tuning and converts it to (trace block, wave shift); PPM_E_INVAL if out of range.
int check_tuning(const ppm_tuning* t, int* wave_shift);
// Sets the launch shape of a context's job (invalidates its captured graphs).
void set_launch_shape(ppm_ctx* c, int trace_block, int wave_shift, int walk_stack);
// Coordinate search over the launch shapes: for each knob in turn, `trial(t)` applies t, runs
// passes and returns their milliseconds; a knob's value replaces the kept one only when it is 1%
// faster.  The winner is applied last and written to *best.
template <class Trial>
int autotune(ppm_tuning start, bool batch, Trial trial, ppm_tuning* best) {
    ppm_tuning cur = start;
    float ms = 0.0f;
    int rc = trial(cur, &ms);
    if (rc) return rc;
    cur.ms_per_pass = ms;
    const int blocks[] = {64, 128, 256, 512};
    const int widths[] = {4, 8, 16, 32, 64};
    const uint32_t graphs_batch[] = {0, 1, 4};
    const uint32_t graphs_plain[] = {0, 1};
    const int stacks[] = {8, 16, 32};
    for (int knob = 0; knob < 5; knob++) {
        const int n = knob == 0 ? 4 : knob == 1 ? 5 : knob == 2 ? (batch ? 3 : 2) : knob == 3 ? 3 : 2;
        const ppm_tuning base = cur;  // the value this knob starts from is already timed
        for (int k = 0; k < n; k++) {
            ppm_tuning t = cur;
            if (knob == 0) t.trace_block = blocks[k];
            if (knob == 1) t.tile_wave_width = widths[k];
            if (knob == 2) t.passes_per_graph = batch ? graphs_batch[k] : graphs_plain[k];
            if (knob == 3) t.walk_stack = stacks[k];
            if (knob == 4) t.eye_overlap = k;
            if ((knob == 0 && t.trace_block == base.trace_block) ||
                (knob == 1 && t.tile_wave_width == base.tile_wave_width) ||
                (knob == 2 && t.passes_per_graph == base.passes_per_graph) ||
                (knob == 3 && t.walk_stack == base.walk_stack) || (knob == 4 && t.eye_overlap == base.eye_overlap))
                continue;
            if ((rc = trial(t, &ms))) return rc;
            // a shape must beat the kept one by 1%: trials of a few eye cadences vary by about that
            if (ms < cur.ms_per_pass * 0.99f) {
                cur = t;
                cur.ms_per_pass = ms;
            }
        }
    }
    rc = trial(cur, &ms);
    if (best) *best = cur;
    return rc;
}

}  // namespace ppm_internal
