// Batched pass loop behind include/ppm.h (ppm_batch_*): many jobs -- each a scene, a camera and a
// set of SPPM state -- advance together, one launch per stage per pass for all of them.
//
// The reference renders one scene per process, one GL pass after another (m_display,
// glslppm.cpp:214-536); a small scene there leaves the GPU idle between tiny draws.  Here every
// kernel of the pass is launched once over the concatenated blocks of all resident jobs: a block
// map gives each grid block its (job, block of that job), and the job records (Job, kernels.h)
// live in HBM.  Per-pass decisions (eye cadence, AAOffset, first pass, pass budget) are taken on
// the device by k_pass_begin_batch from each job's own counters, so jobs in different phases share
// one launch sequence, and that sequence is captured once as a hipGraph and replayed every pass.
// The graph is captured again only when a job is added or removed.
//
// Each job's kernels touch only that job's memory and its hash sees only commutative atomics, so
// every job's result is bit-identical to the same job run alone through ppm_create / ppm_pass.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.h"

struct ppm_batch {
    int device = 0;
    hipStream_t stream = nullptr;   // passes
    hipStream_t upload = nullptr;   // scene uploads and job initialisation
    ppm_internal::Arena* arena = nullptr;
    uint32_t passes_per_graph = 1;  // 0: plain launches
    std::vector<ppm_ctx*> slots;    // job id -> context (null once removed)
    // device tables of the current job set
    bool dirty = true;
    int n_jobs = 0;
    Job* d_jobs = nullptr;
    BlockRef* d_map[3] = {nullptr, nullptr, nullptr};  // tiles, photon slots, hash clears
    int n_blocks[3] = {0, 0, 0};
    hipGraphExec_t graph = nullptr;
    struct Retired {
        hipGraphExec_t g;
        hipEvent_t ev;
    };
    std::vector<Retired> retired;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    std::string err;
};

namespace {

thread_local std::string g_batch_error;

int berr(ppm_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_error = msg;
    return code;
}

#define BHIP(b, call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return berr((b), e_ == hipErrorOutOfMemory ? PPM_E_OOM : PPM_E_HIP,              \
                        std::string(#call) + ": " + hipGetErrorString(e_));                  \
    } while (0)

enum { kMapTile = 0, kMapPhoton = 1, kMapClear = 2 };

// Destroys retired graphs whose last launch has finished (all of them when `wait`).
void reap(ppm_batch* b, bool wait) {
    for (size_t k = 0; k < b->retired.size();) {
        auto& r = b->retired[k];
        if (wait) (void)hipEventSynchronize(r.ev);
        if (hipEventQuery(r.ev) == hipSuccess) {
            (void)hipGraphExecDestroy(r.g);
            (void)hipEventDestroy(r.ev);
            b->retired.erase(b->retired.begin() + (long)k);
        } else {
            k++;
        }
    }
}

void retire_graph(ppm_batch* b) {
    if (!b->graph) return;
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
        hipEventRecord(ev, b->stream) == hipSuccess) {
        b->retired.push_back({b->graph, ev});
    } else {
        if (ev) (void)hipEventDestroy(ev);
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(b->graph);
    }
    b->graph = nullptr;
}

// One batched pass over the current tables: m_display's stage order for every job.
void queue_batch_pass(ppm_batch* b, hipStream_t st) {
    const int n = b->n_jobs;
    const Job* J = b->d_jobs;
    hipLaunchKernelGGL(k_pass_begin_batch, dim3((n + 63) / 64), dim3(64), 0, st, J, n);
    hipLaunchKernelGGL(k_eye_batch, dim3(b->n_blocks[kMapTile]), dim3(256), 0, st, J, (const BlockRef*)b->d_map[kMapTile]);
    hipLaunchKernelGGL(k_grid_batch, dim3(n), dim3(1024), 0, st, J);
    hipLaunchKernelGGL(k_clear_slots_batch, dim3(b->n_blocks[kMapClear]), dim3(256), 0, st, J,
                       (const BlockRef*)b->d_map[kMapClear]);
    hipLaunchKernelGGL(k_photon_gen_batch, dim3(b->n_blocks[kMapPhoton]), dim3(256), 0, st, J,
                       (const BlockRef*)b->d_map[kMapPhoton]);
    hipLaunchKernelGGL(k_bucket_scan_batch, dim3(n), dim3(1024), 0, st, J);
    hipLaunchKernelGGL(k_bucket_scatter_batch, dim3(b->n_blocks[kMapPhoton]), dim3(256), 0, st, J,
                       (const BlockRef*)b->d_map[kMapPhoton]);
    hipLaunchKernelGGL(k_photon_trace_batch, dim3(b->n_blocks[kMapPhoton]), dim3(256), 0, st, J,
                       (const BlockRef*)b->d_map[kMapPhoton]);
    hipLaunchKernelGGL(k_gather_batch, dim3(b->n_blocks[kMapTile]), dim3(256), 0, st, J, (const BlockRef*)b->d_map[kMapTile]);
}

// Rebuilds the job array, the block maps and the graph after jobs were added or removed.
int rebuild(ppm_batch* b) {
    if (!b->dirty) return PPM_OK;
    std::vector<Job> jobs;
    std::vector<BlockRef> maps[3];
    for (ppm_ctx* c : b->slots) {
        if (!c) continue;
        const int j = (int)jobs.size();
        jobs.push_back(c->job);
        const int per[3] = {c->job.tile_blocks, c->job.photon_blocks, c->job.clear_blocks};
        for (int k = 0; k < 3; k++)
            for (int i = 0; i < per[k]; i++) maps[k].push_back(BlockRef{j, i});
    }
    // the previous tables may still be read by queued passes: hand them back to the arena, which
    // reuses them only after the stream has passed this point
    if (b->d_jobs) b->arena->release(b->d_jobs, b->stream);
    for (int k = 0; k < 3; k++)
        if (b->d_map[k]) b->arena->release(b->d_map[k], b->stream);
    b->d_jobs = nullptr;
    for (int k = 0; k < 3; k++) b->d_map[k] = nullptr;
    retire_graph(b);
    b->n_jobs = (int)jobs.size();
    for (int k = 0; k < 3; k++) b->n_blocks[k] = (int)maps[k].size();
    if (b->n_jobs == 0) {
        b->dirty = false;
        return PPM_OK;
    }
    BHIP(b, b->arena->alloc((void**)&b->d_jobs, jobs.size() * sizeof(Job)));
    BHIP(b, hipMemcpyAsync(b->d_jobs, jobs.data(), jobs.size() * sizeof(Job), hipMemcpyHostToDevice, b->upload));
    for (int k = 0; k < 3; k++) {
        BHIP(b, b->arena->alloc((void**)&b->d_map[k], maps[k].size() * sizeof(BlockRef)));
        BHIP(b, hipMemcpyAsync(b->d_map[k], maps[k].data(), maps[k].size() * sizeof(BlockRef), hipMemcpyHostToDevice,
                               b->upload));
    }
    BHIP(b, hipStreamSynchronize(b->upload));
    if (b->passes_per_graph > 0) {
        hipGraph_t g = nullptr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < b->passes_per_graph; k++) queue_batch_pass(b, b->stream);
        BHIP(b, hipStreamEndCapture(b->stream, &g));
        const hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BHIP(b, e);
    }
    b->dirty = false;
    return PPM_OK;
}

ppm_ctx* job_ctx(ppm_batch* b, int32_t job) {
    if (!b || job < 0 || (size_t)job >= b->slots.size()) return nullptr;
    return b->slots[(size_t)job];
}

}  // namespace

extern "C" {

void ppm_batch_config_default(ppm_batch_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->passes_per_graph = 1;
}

int ppm_batch_create(const ppm_batch_config* cfg, ppm_batch** out) {
    if (!out) return berr(nullptr, PPM_E_INVAL, "null output");
    *out = nullptr;
    ppm_batch_config d;
    ppm_batch_config_default(&d);
    if (!cfg) cfg = &d;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return berr(nullptr, PPM_E_NODEV, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return berr(nullptr, PPM_E_INVAL, "device ordinal out of range");
    if (cfg->passes_per_graph > 64) return berr(nullptr, PPM_E_INVAL, "passes_per_graph is limited to 64");
    ppm_batch* b = new (std::nothrow) ppm_batch();
    if (!b) return berr(nullptr, PPM_E_OOM, "out of host memory");
    b->device = cfg->device;
    b->passes_per_graph = cfg->passes_per_graph;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->upload, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        g_batch_error = std::string("batch streams: ") + hipGetErrorString(e);
        ppm_batch_destroy(b);
        return PPM_E_HIP;
    }
    b->arena = new (std::nothrow) ppm_internal::Arena(b->device);
    if (!b->arena) {
        ppm_batch_destroy(b);
        return berr(nullptr, PPM_E_OOM, "out of host memory");
    }
    *out = b;
    return PPM_OK;
}

int ppm_batch_add(ppm_batch* b, const ppm_scene* scene, const ppm_params* params, uint32_t passes, int32_t* job) {
    if (!b || !job) return berr(b, PPM_E_INVAL, "null argument");
    *job = -1;
    int rc;
    if ((rc = ppm_internal::validate(scene, params))) return berr(b, rc, ppm_internal::create_error());
    if (params->device != b->device) return berr(b, PPM_E_INVAL, "params.device differs from the batch's device");
    ppm_ctx* c = new (std::nothrow) ppm_ctx();
    if (!c) return berr(b, PPM_E_OOM, "out of host memory");
    rc = ppm_internal::create_context(c, scene, params, b->stream, b->upload, b->arena, passes);
    if (rc != PPM_OK) {
        const std::string m = c->err.empty() ? ppm_internal::create_error() : c->err;
        ppm_internal::release_context(c);
        delete c;
        return berr(b, rc, m);
    }
    b->slots.push_back(c);
    *job = (int32_t)(b->slots.size() - 1);
    b->dirty = true;
    return PPM_OK;
}

int ppm_batch_remove(ppm_batch* b, int32_t job) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    // blocks return to the arena once the passes queued so far are done with them
    ppm_internal::release_context(c);
    delete c;
    b->slots[(size_t)job] = nullptr;
    b->dirty = true;
    return PPM_OK;
}

int ppm_batch_set_passes(ppm_batch* b, int32_t job, uint32_t passes) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    BHIP(b, hipSetDevice(b->device));
    // ordered after the passes queued so far; a plain 4-byte store into the job's counters
    BHIP(b, hipMemcpyAsync((char*)c->job.g + offsetof(DevGlobals, passes_left), &passes, sizeof(passes),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    return PPM_OK;
}

int ppm_batch_pass(ppm_batch* b, uint32_t npasses) {
    if (!b) return PPM_E_INVAL;
    BHIP(b, hipSetDevice(b->device));
    int rc;
    if ((rc = rebuild(b))) return rc;
    reap(b, false);
    if (npasses == 0 || b->n_jobs == 0) return PPM_OK;
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    uint32_t left = npasses;
    if (b->graph) {
        while (left >= b->passes_per_graph) {
            BHIP(b, hipGraphLaunch(b->graph, b->stream));
            left -= b->passes_per_graph;
        }
    }
    for (; left > 0; left--) queue_batch_pass(b, b->stream);
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    b->timed = true;
    return PPM_OK;
}

int ppm_batch_sync(ppm_batch* b) {
    if (!b) return PPM_E_INVAL;
    BHIP(b, hipStreamSynchronize(b->stream));
    reap(b, false);
    return PPM_OK;
}

float ppm_batch_last_pass_ms(ppm_batch* b) {
    if (!b || !b->timed) return -1.0f;
    if (hipEventSynchronize(b->ev1) != hipSuccess) return -1.0f;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) return -1.0f;
    return ms;
}

int ppm_batch_progress(ppm_batch* b, int32_t job, uint64_t* frames, uint32_t* passes_left) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    DevGlobals g;
    BHIP(b, hipMemcpy(&g, c->job.g, sizeof(g), hipMemcpyDeviceToHost));
    if (frames) *frames = g.frame;
    if (passes_left) *passes_left = g.passes_left;
    return PPM_OK;
}

int ppm_batch_readback(ppm_batch* b, int32_t job, ppm_buf which, void* dst, size_t bytes) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    const int rc = ppm_readback(c, which, dst, bytes);
    if (rc != PPM_OK) return berr(b, rc, c->err);
    return PPM_OK;
}

size_t ppm_batch_buffer_bytes(ppm_batch* b, int32_t job, ppm_buf which) {
    ppm_ctx* c = job_ctx(b, job);
    return c ? ppm_buffer_bytes(c, which) : 0;
}

int32_t ppm_batch_num_jobs(const ppm_batch* b) {
    if (!b) return 0;
    int32_t n = 0;
    for (ppm_ctx* c : b->slots) n += c ? 1 : 0;
    return n;
}

uint64_t ppm_batch_device_bytes(const ppm_batch* b) { return b && b->arena ? b->arena->held_bytes() : 0; }

uint64_t ppm_scene_footprint(const ppm_scene* scene, const ppm_params* params) {
    if (!scene || !params) return 0;
    return ppm_internal::footprint(scene, params);
}

void ppm_batch_destroy(ppm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->upload) (void)hipStreamSynchronize(b->upload);
    for (ppm_ctx* c : b->slots)
        if (c) {
            ppm_internal::release_context(c);
            delete c;
        }
    b->slots.clear();
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    reap(b, true);
    if (b->arena) {
        if (b->d_jobs) b->arena->release(b->d_jobs, nullptr);
        for (int k = 0; k < 3; k++)
            if (b->d_map[k]) b->arena->release(b->d_map[k], nullptr);
        delete b->arena;
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->upload) (void)hipStreamDestroy(b->upload);
    delete b;
}

const char* ppm_batch_last_error(const ppm_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

}  // extern "C"
