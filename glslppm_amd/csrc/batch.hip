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
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.h"

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

enum { kMapTile = 0, kMapPhoton = 1, kMapBegin = 2, kMapTrace = 3, kMaps = 4 };

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

// One batched pass over the current tables: m_display's stage order for every job, each job's
// gather of the pass before in the blocks the photon-stage launch runs after every trace block.  With
// eye_overlap the eye stages run beside the photon stage: a job's stage that has not begun when its
// eye pass comes runs on `side` and that pass's deposits wait for k_hash after the join (kEyeNow);
// with `run_ahead` (plain launches), a job's next eye stage begins on `ahead` in the pass after
// each of its eye passes and runs beside the photon passes until the next one (kEyeAhead), so the
// eye stages of many jobs whose eye passes coincide spread over the passes in between instead of
// lengthening the eye pass.  An eye pass p then holds its deposits back (defer_all) and waits for
// the ahead launches of pass p - lag after its photon stage, before k_hash, which covers every
// ahead stage it needs (eye_begin starts one only with at least `lag` passes to go); the photon
// stage itself waits for nothing on `ahead`.  Without eye_overlap everything runs in order on one
// stream.
// The eye-stage work a batched pass can have: an eye pass of some job (its stage, if not begun
// ahead, and its held-back deposits), or some job that may begin its next eye stage ahead.
struct PassNeeds {
    bool eye_pass;
    bool ahead;
};

// Advances every job's frame and budget mirror by one queued batched pass, by k_pass_begin's rule
// (a job with budget left runs the pass, which is an eye pass when its frame count is a multiple
// of its cadence), and returns what the pass can need: eye_begin begins a kEyeNow stage only in a
// job's eye pass, k_hash deposits only there, and a kEyeAhead stage begins only in a job's photon
// pass with at least `lag` passes to its next eye pass and the budget to reach it.  A launch the
// pass cannot need would exit in every block, so leaving it out changes nothing.
PassNeeds advance_mirrors(ppm_batch* b) {
    PassNeeds n{false, false};
    for (ppm_ctx* c : b->slots) {
        if (!c || c->batch_left == 0u) continue;
        const uint64_t L = (uint64_t)c->prm.max_path_length;
        const uint64_t in = c->batch_frame % L;
        if (in == 0) {
            n.eye_pass = true;
        } else {
            const uint64_t to_next = L - in;
            const bool budget = c->batch_left == 0xffffffffu || (uint64_t)(c->batch_left - 1u) >= to_next;
            if (budget && to_next >= (uint64_t)b->lag) n.ahead = true;
        }
        c->batch_frame++;
        if (c->batch_left != 0xffffffffu) c->batch_left--;
    }
    return n;
}

void queue_batch_pass(ppm_batch* b, hipStream_t st, bool run_ahead, PassNeeds need) {
    const int n = b->n_jobs;
    const Job* J = b->d_jobs;
    const unsigned int pass = run_ahead ? b->pass_no : 0u;
    const bool now = need.eye_pass, ahead = run_ahead && need.ahead;
    // with stages begun ahead, an eye pass's deposits always wait for k_hash, which runs after the
    // pass has waited for those stages: the photon stage need not wait for them
    hipLaunchKernelGGL(k_pass_begin_batch, dim3(b->n_blocks[kMapBegin]), dim3(256), 0, st, J,
                       (const BlockRef*)b->d_map[kMapBegin], pass, run_ahead ? 1 : 0);
    hipStream_t es = b->eye_overlap ? b->side : st;
    if (b->eye_overlap && (now || ahead)) (void)hipEventRecord(b->ev_fork, st);
    auto eye_stage = [&](hipStream_t s, int mode, int lag) {
        hipLaunchKernelGGL(k_eye_begin_batch, dim3((n + 63) / 64), dim3(64), 0, s, J, n, mode, pass, lag);
        hipLaunchKernelGGL(k_eye_batch, dim3(b->n_blocks[kMapTile]), dim3(256), walk_lds_bytes(b->walk_stack, 256), s,
                           J, (const BlockRef*)b->d_map[kMapTile], mode);
        hipLaunchKernelGGL(k_grid_batch, dim3(n), dim3(1024), 0, s, J, mode);
    };
    if (now) {
        if (b->eye_overlap) (void)hipStreamWaitEvent(es, b->ev_fork, 0);
        eye_stage(es, kEyeNow, 0);
        if (b->eye_overlap) (void)hipEventRecord(b->ev_join, es);
    }
    if (run_ahead) {
        if (ahead) {
            (void)hipStreamWaitEvent(b->ahead, b->ev_fork, 0);
            eye_stage(b->ahead, kEyeAhead, b->lag);
        }
        (void)hipEventRecord(b->ring[pass % ppm_batch::kAheadRing], b->ahead);
        b->pass_no++;
    }
    hipLaunchKernelGGL(k_photon_stage_batch, dim3(b->n_blocks[kMapTrace]), dim3(b->trace_block),
                       walk_lds_bytes(b->walk_stack, b->trace_block), st, J, (const BlockRef*)b->d_map[kMapTrace]);
    if (now) {
        if (b->eye_overlap) (void)hipStreamWaitEvent(st, b->ev_join, 0);
        // the ahead launches of pass - lag and before: every stage an eye pass of this pass needs
        // began there (eye_begin), so its grid and hit points are complete before k_hash
        if (run_ahead && pass >= b->ring_from + (uint32_t)b->lag)
            (void)hipStreamWaitEvent(st, b->ring[(pass - (uint32_t)b->lag) % ppm_batch::kAheadRing], 0);
        hipLaunchKernelGGL(k_hash_batch, dim3(b->n_blocks[kMapPhoton]), dim3(256), 0, st, J,
                           (const BlockRef*)b->d_map[kMapPhoton]);
    }
}

// The gathers of the last batched pass of a sequence; every other pass's gathers run in the
// photon-stage launch of the pass after it.
void queue_batch_flush(ppm_batch* b, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_batch, dim3(b->n_blocks[kMapTile]), dim3(256), 0, st, (const Job*)b->d_jobs,
                       (const BlockRef*)b->d_map[kMapTile]);
}

// Makes the pass stream wait for every eye stage queued on `ahead` so far, and starts a new ring
// epoch (later passes need not wait for ring events recorded before this point).  Called before
// the job tables or a job's memory change hands and before the host reads job state.
void join_ahead(ppm_batch* b) {
    if (b->pass_no > b->ring_from)
        (void)hipStreamWaitEvent(b->stream, b->ring[(b->pass_no - 1) % ppm_batch::kAheadRing], 0);
    b->ring_from = b->pass_no;
}

bool ahead_mode(const ppm_batch* b) { return b->eye_overlap && b->passes_per_graph == 0; }

// Rebuilds the job array, the block maps and the graph after jobs were added or removed.
int rebuild(ppm_batch* b) {
    if (!b->dirty) return PPM_OK;
    std::vector<Job> jobs;
    std::vector<int> tris;
    std::vector<BlockRef> maps[kMaps];
    b->table_ids.clear();
    for (size_t id = 0; id < b->slots.size(); id++) {
        ppm_ctx* c = b->slots[id];
        if (!c) continue;
        const int j = (int)jobs.size();
        jobs.push_back(c->job);
        tris.push_back(c->T);
        b->table_ids.push_back((int32_t)id);
        const int per[kMaps] = {0, c->job.photon_blocks, c->job.begin_blocks, 0};
        for (int k = 0; k < kMaps; k++)
            for (int i = 0; i < per[k]; i++) maps[k].push_back(BlockRef{j, i});
    }
    // The walk launches (eye tiles, photon trace blocks) take the jobs of the largest scenes first:
    // their walks are the longest, and a launch ends with the blocks dispatched last.  The
    // photon-stage launch's gather blocks follow every job's trace blocks, so that they are
    // dispatched last of all.
    std::vector<int> order(jobs.size());
    for (size_t j = 0; j < jobs.size(); j++) order[j] = (int)j;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return tris[x] > tris[y]; });
    for (int j : order) {
        for (int i = 0; i < jobs[j].tile_blocks; i++) maps[kMapTile].push_back(BlockRef{j, i});
        for (int i = 0; i < jobs[j].trace_blocks; i++) maps[kMapTrace].push_back(BlockRef{j, i});
    }
    for (int j : order)
        for (int i = 0; i < jobs[j].gather_blocks; i++)
            maps[kMapTrace].push_back(BlockRef{j, jobs[j].trace_blocks + i});
    // the previous tables may still be read by queued passes (eye stages on `ahead` included): hand
    // them back to the arena, which reuses them only after the stream has passed this point
    join_ahead(b);
    if (b->d_jobs) b->arena->release(b->d_jobs, b->stream);
    for (int k = 0; k < kMaps; k++)
        if (b->d_map[k]) b->arena->release(b->d_map[k], b->stream);
    b->d_jobs = nullptr;
    for (int k = 0; k < kMaps; k++) b->d_map[k] = nullptr;
    retire_graph(b);
    b->n_jobs = (int)jobs.size();
    for (int k = 0; k < kMaps; k++) b->n_blocks[k] = (int)maps[k].size();
    int shortest = 0x7fffffff;
    for (const Job& j : jobs) shortest = std::min(shortest, j.ep.max_path);
    b->lag = jobs.empty() ? 0 : std::min(shortest - 1, ppm_batch::kAheadRing - 1);
    if (b->n_jobs == 0) {
        b->dirty = false;
        return PPM_OK;
    }
    BHIP(b, b->arena->alloc((void**)&b->d_jobs, jobs.size() * sizeof(Job)));
    BHIP(b, hipMemcpyAsync(b->d_jobs, jobs.data(), jobs.size() * sizeof(Job), hipMemcpyHostToDevice, b->upload));
    for (int k = 0; k < kMaps; k++) {
        BHIP(b, b->arena->alloc((void**)&b->d_map[k], maps[k].size() * sizeof(BlockRef)));
        BHIP(b, hipMemcpyAsync(b->d_map[k], maps[k].data(), maps[k].size() * sizeof(BlockRef), hipMemcpyHostToDevice,
                               b->upload));
    }
    BHIP(b, hipStreamSynchronize(b->upload));
    if (b->passes_per_graph > 0) {
        hipGraph_t g = nullptr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        // a captured pass is replayed whatever the jobs' phases: it carries every launch
        for (uint32_t k = 0; k < b->passes_per_graph; k++) queue_batch_pass(b, b->stream, false, PassNeeds{true, true});
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

namespace ppm_internal {

int batch_add_resident(ppm_batch* b, std::shared_ptr<SceneDev> scene, const ppm_camera& cam, const ppm_params* p,
                       uint32_t passes, int32_t* job) {
    *job = -1;
    ppm_ctx* c = new (std::nothrow) ppm_ctx();
    if (!c) return berr(b, PPM_E_OOM, "out of host memory");
    const int rc = create_context(c, scene, cam, p, b->stream, b->upload, b->arena, passes);
    if (rc != PPM_OK) {
        const std::string m = c->err.empty() ? create_error() : c->err;
        release_context(c);
        delete c;
        return berr(b, rc, m);
    }
    set_launch_shape(c, b->trace_block, b->wave_shift, b->walk_stack);
    c->batch_frame = 0;
    c->batch_left = passes;
    b->slots.push_back(c);
    *job = (int32_t)(b->slots.size() - 1);
    b->dirty = true;
    return PPM_OK;
}

bool batch_job_has_passes(const ppm_batch* b, int32_t job) {
    if (!b || job < 0 || (size_t)job >= b->slots.size() || !b->slots[(size_t)job]) return false;
    return b->slots[(size_t)job]->batch_left != 0u;
}

int batch_prepare(ppm_batch* b) {
    BHIP(b, hipSetDevice(b->device));
    return rebuild(b);
}

int batch_snapshot(ppm_batch* b, unsigned long long* dev_tmp, unsigned long long* host_out, hipEvent_t done,
                   std::vector<int32_t>* ids) {
    *ids = b->table_ids;
    const int n = b->n_jobs;
    if (n > 0) {
        hipLaunchKernelGGL(k_batch_status, dim3((n + 63) / 64), dim3(64), 0, b->stream, (const Job*)b->d_jobs, n, dev_tmp);
        BHIP(b, hipGetLastError());
        BHIP(b, hipMemcpyAsync(host_out, dev_tmp, (size_t)n * 16, hipMemcpyDeviceToHost, b->stream));
    }
    BHIP(b, hipEventRecord(done, b->stream));
    return PPM_OK;
}

int batch_fetch(ppm_batch* b, int32_t job, float4* radiance, float4* flux_radius, float4* emission_count,
                uchar4* rgba8, ppm_counters* counters) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    BHIP(b, hipSetDevice(b->device));
    const Job& J = c->job;
    const size_t Q = (size_t)c->Q;
    hipStream_t up = b->upload;
    // the job has stopped, but the first pass queued after its last one folds its last emitted
    // counts into NumPhotons, which the resolve and the counters read: wait for that fold
    BHIP(b, hipStreamWaitEvent(up, b->ev_begun, 0));
    if (radiance || rgba8) {
        hipLaunchKernelGGL(k_resolve, dim3((c->Q + 255) / 256), dim3(256), 0, up, c->Q, J.hs, (const DevGlobals*)J.g,
                           c->resolve_lin, c->resolve_map);
        BHIP(b, hipGetLastError());
    }
    if (radiance) BHIP(b, hipMemcpyAsync(radiance, c->resolve_lin, Q * 16, hipMemcpyDeviceToHost, up));
    if (flux_radius) BHIP(b, hipMemcpyAsync(flux_radius, J.hs.tau_r, Q * 16, hipMemcpyDeviceToHost, up));
    if (emission_count) BHIP(b, hipMemcpyAsync(emission_count, J.hs.e_n, Q * 16, hipMemcpyDeviceToHost, up));
    if (rgba8) BHIP(b, hipMemcpyAsync(rgba8, J.hs.rgba8, Q * 4, hipMemcpyDeviceToHost, up));
    DevGlobals g;
    BHIP(b, hipMemcpyAsync(&g, J.g, sizeof(g), hipMemcpyDeviceToHost, up));
    BHIP(b, hipStreamSynchronize(up));
    if (counters) {
        std::memset(counters, 0, sizeof(*counters));
        counters->frame_count = g.frame;
        counters->eye_samples = g.eye_samples;
        folded_photon_counts(g, &counters->photons_emitted, &counters->num_photons_f);
        counters->photon_segments = (uint64_t)c->P * g.frame;
    }
    return PPM_OK;
}

}  // namespace ppm_internal

extern "C" {

void ppm_batch_config_default(ppm_batch_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->passes_per_graph = 0;
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
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->ahead, hipStreamNonBlocking);
    for (int k = 0; k < ppm_batch::kAheadRing && e == hipSuccess; k++)
        e = hipEventCreateWithFlags(&b->ring[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_begun, hipEventDisableTiming);
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
    BHIP(b, hipSetDevice(b->device));
    ppm_internal::PackedScene ps;
    if ((rc = ppm_internal::pack_scene(scene, &ps))) return berr(b, rc, ppm_internal::create_error());
    std::shared_ptr<ppm_internal::SceneDev> sd;
    std::string err;
    if ((rc = ppm_internal::upload_scene(ps, b->arena, b->upload, b->stream, &sd, &err))) {
        (void)hipStreamSynchronize(b->upload);
        return berr(b, rc, err);
    }
    BHIP(b, hipStreamSynchronize(b->upload));  // the packed host copy is released on return
    return ppm_internal::batch_add_resident(b, sd, scene->camera, params, passes, job);
}

int ppm_batch_remove(ppm_batch* b, int32_t job) {
    ppm_ctx* c = job_ctx(b, job);
    if (!c) return berr(b, PPM_E_INVAL, "no such job");
    // blocks return to the arena once the passes queued so far (eye stages on `ahead` included)
    // are done with them
    join_ahead(b);
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
    c->batch_left = passes;
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
    bool begun = false;
    if (b->graph) {
        while (left >= b->passes_per_graph) {
            BHIP(b, hipGraphLaunch(b->graph, b->stream));
            for (uint32_t k = 0; k < b->passes_per_graph; k++) (void)advance_mirrors(b);
            left -= b->passes_per_graph;
            if (!begun) BHIP(b, hipEventRecord(b->ev_begun, b->stream));
            begun = true;
        }
    }
    const bool run_ahead = ahead_mode(b);
    for (; left > 0; left--) {
        const PassNeeds need = advance_mirrors(b);
        queue_batch_pass(b, b->stream, run_ahead, need);
        if (!begun) BHIP(b, hipEventRecord(b->ev_begun, b->stream));
        begun = true;
    }
    queue_batch_flush(b, b->stream);
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    b->timed = true;
    return PPM_OK;
}

int ppm_batch_sync(ppm_batch* b) {
    if (!b) return PPM_E_INVAL;
    BHIP(b, hipSetDevice(b->device));
    join_ahead(b);
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
    join_ahead(b);
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
    BHIP(b, hipSetDevice(b->device));
    join_ahead(b);
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
    const int tex = (scene->texcoords && scene->textures) ? scene->texture_size : 0;
    return ppm_internal::state_footprint(params) +
           ppm_internal::scene_footprint(scene->num_triangles, scene->num_nodes, scene->num_materials, scene->num_lights,
                                         tex, ppm_internal::wide_records(scene));
}

int ppm_batch_get_tuning(ppm_batch* b, ppm_tuning* t) {
    if (!b || !t) return PPM_E_INVAL;
    t->trace_block = b->trace_block;
    t->tile_wave_width = 1 << b->wave_shift;
    t->passes_per_graph = b->passes_per_graph;
    t->walk_stack = b->walk_stack;
    t->eye_overlap = b->eye_overlap ? 1 : 0;
    t->ms_per_pass = 0.0f;
    return PPM_OK;
}

int ppm_batch_set_tuning(ppm_batch* b, const ppm_tuning* t) {
    if (!b) return PPM_E_INVAL;
    int sh = 0;
    if (ppm_internal::check_tuning(t, &sh)) return berr(b, PPM_E_INVAL, "tuning out of range");
    b->trace_block = t->trace_block;
    b->wave_shift = sh;
    b->passes_per_graph = t->passes_per_graph;
    b->walk_stack = t->walk_stack;
    b->eye_overlap = t->eye_overlap != 0;
    for (ppm_ctx* c : b->slots)
        if (c) ppm_internal::set_launch_shape(c, b->trace_block, b->wave_shift, b->walk_stack);
    b->dirty = true;
    return PPM_OK;
}

int ppm_batch_autotune(ppm_batch* b, uint32_t passes_per_trial, ppm_tuning* best) {
    if (!b) return PPM_E_INVAL;
    if (ppm_batch_num_jobs(b) == 0) return berr(b, PPM_E_INVAL, "the batch has no job to tune on");
    if (passes_per_trial == 0) {
        int L = 1;
        for (ppm_ctx* c : b->slots)
            if (c) L = std::max(L, c->prm.max_path_length);
        passes_per_trial = 4u * (uint32_t)L;
    }
    ppm_tuning start;
    ppm_batch_get_tuning(b, &start);
    auto trial = [&](const ppm_tuning& t, float* ms) -> int {
        int rc;
        if ((rc = ppm_batch_set_tuning(b, &t))) return rc;
        if ((rc = ppm_batch_pass(b, 1))) return rc;  // table upload and graph capture stay out of the timing
        if ((rc = ppm_batch_pass(b, passes_per_trial))) return rc;
        // eye stages begun ahead inside the trial belong to it: the timed region ends after them
        join_ahead(b);
        BHIP(b, hipEventRecord(b->ev1, b->stream));
        const float v = ppm_batch_last_pass_ms(b);
        if (v < 0.0f) return berr(b, PPM_E_HIP, "trial timing failed");
        *ms = v / (float)passes_per_trial;
        return PPM_OK;
    };
    return ppm_internal::autotune(start, true, trial, best);
}

uint64_t ppm_state_footprint(const ppm_params* params) {
    return params ? ppm_internal::state_footprint(params) : 0;
}

void ppm_batch_destroy(ppm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->ahead) (void)hipStreamSynchronize(b->ahead);
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
        for (int k = 0; k < kMaps; k++)
            if (b->d_map[k]) b->arena->release(b->d_map[k], nullptr);
        delete b->arena;
    }
    if (b->side) (void)hipStreamSynchronize(b->side);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev_begun) (void)hipEventDestroy(b->ev_begun);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    for (int k = 0; k < ppm_batch::kAheadRing; k++)
        if (b->ring[k]) (void)hipEventDestroy(b->ring[k]);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->ahead) (void)hipStreamDestroy(b->ahead);
    if (b->upload) (void)hipStreamDestroy(b->upload);
    delete b;
}

const char* ppm_batch_last_error(const ppm_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

}  // extern "C"
