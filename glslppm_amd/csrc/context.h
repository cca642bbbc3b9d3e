// Internal: the per-job context shared by the plain pass loop (runtime.hip) and the batched and
// served forms (batch.hip).  One ppm_ctx owns one scene, one camera and one set of pass state in
// HBM; its kernels see it as a Job record.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
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

}  // namespace ppm_internal

struct ppm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // the stream passes run on
    bool own_stream = true;         // false when the stream belongs to a batch
    ppm_internal::Arena* arena = nullptr;  // null: hipMalloc / hipFree
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    ppm_params prm;
    int Q = 0, P = 0, H = 0, N = 0, T = 0, M = 0, NL = 0;
    Job job;
    float4* resolve_lin = nullptr;
    float4* resolve_map = nullptr;
    std::vector<void*> allocs;
    size_t device_bytes = 0;
    // per-kernel event timing (ppm_profile_enable)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double prof_ms[PPM_KERNEL_KINDS] = {0, 0, 0, 0};
    uint64_t prof_n[PPM_KERNEL_KINDS] = {0, 0, 0, 0};
    uint64_t frame = 0;             // passes queued by ppm_pass (the plain loop's eye cadence)
    // hipGraphs of one photon-only pass and one eye pass (PPM_FLAG_GRAPH)
    hipGraphExec_t graph_pass[2] = {nullptr, nullptr};
    std::string err;
};

namespace ppm_internal {

int validate(const ppm_scene* s, const ppm_params* p);
// Creates the context's device state.  stream: where passes will run (null: a stream of its
// own).  Uploads and initialisation run on `upload` (null: the pass stream) and are complete on
// return.  passes_left: the device-side pass budget (0xffffffff: unlimited).
int create_context(ppm_ctx* c, const ppm_scene* s, const ppm_params* p, hipStream_t stream, hipStream_t upload,
                   Arena* arena, uint32_t passes_left);
// Releases the device state.  Arena blocks become reusable once the work queued on the
// context's stream so far has finished; other blocks are freed at once.
void release_context(ppm_ctx* c);
// HBM bytes a context of this scene and these parameters allocates.
size_t footprint(const ppm_scene* s, const ppm_params* p);
int set_error(ppm_ctx* c, int code, const std::string& msg);
const std::string& create_error();
// Blocks of each kind for one job.
int tile_blocks(const ppm_params* p);
int photon_blocks(const ppm_params* p);
int clear_blocks(const ppm_params* p);

}  // namespace ppm_internal
