// Serving loop behind include/ppm.h (ppm_server_*): continuous batching of render jobs.
//
// A worker thread owns one batch (batch.hip) and repeats:
//   1. admit queued jobs, first come first served, skipping a job that does not fit the HBM
//      budget or the job limit (its scene is uploaded with its first job; the uploads and the
//      job's initialisation run on the batch's upload stream, overlapping the step in flight);
//   2. queue the next step (passes_per_step batched passes and its status snapshot: frame and pass
//      budget of every job, copied to pinned memory behind the step's passes) while fewer than two
//      steps are in flight and some job has passes left, so that one step always waits behind the
//      one running;
//   3. read the oldest step's status once two are in flight (or nothing more was queued);
//   4. copy out the results of the jobs that reached their target (on the upload stream, while
//      the new step runs; a finished job's kernels do nothing once the new step's first pass has
//      folded its last emitted counts, which the copy waits for, so its state is final) and remove
//      them from the batch, which frees their memory for later admissions.
// A job's target pass count is its device-side pass budget, so a step may run past it for other
// jobs without changing it.  Jobs never share state, so a job's results are those of the same
// job run alone.

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "context.h"

namespace {

struct SJob {
    int64_t id = 0, scene_id = 0;
    ppm_camera cam;
    ppm_params prm;
    uint32_t target = 0;
    int state = PPM_JOB_QUEUED;
    int error = PPM_OK;
    std::string err;
    int32_t bjob = -1;
    uint64_t passes = 0;
    size_t footprint = 0;
    double t_submit = 0.0, t_admit = -1.0, t_first = -1.0, t_done = -1.0;
    std::vector<float4> radiance, flux_radius, emission;
    std::vector<uchar4> rgba8;
    ppm_counters counters;
};

struct SScene {
    std::shared_ptr<ppm_internal::PackedScene> host;
    std::shared_ptr<ppm_internal::SceneDev> dev;  // touched by the worker only
    size_t dev_bytes = 0;
    int users = 0;    // queued and running jobs that name the scene
    int running = 0;  // resident jobs that render it
    bool removed = false;
};

thread_local std::string g_server_error;

}  // namespace

struct ppm_server {
    ppm_server_config cfg;
    ppm_batch* b = nullptr;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::map<int64_t, SScene> scenes;
    std::map<int64_t, std::shared_ptr<SJob>> jobs;
    std::deque<int64_t> queue;
    int64_t next_scene = 1, next_job = 1;
    bool stop = false;
    std::thread worker;
    // worker state (device_bytes is also read by ppm_server_stats_get, under mu)
    std::vector<int64_t> resident;
    std::map<int32_t, int64_t> by_bjob;
    size_t device_bytes = 0;
    // status snapshots of the steps in flight (kSteps at most), one slot per step
    static constexpr int kSteps = 2;
    unsigned long long* d_status[kSteps] = {nullptr, nullptr};
    unsigned long long* h_status[kSteps] = {nullptr, nullptr};
    hipEvent_t ev_status[kSteps] = {nullptr, nullptr};
    ppm_server_stats st;
    std::string err;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int serr(ppm_server* s, int code, const std::string& msg) {
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = msg;
    } else {
        g_server_error = msg;
    }
    return code;
}

// Marks a job finished (done or failed) and lets its scene go; caller holds mu.
void finish_locked(ppm_server* S, SJob& J, int state, int error, const std::string& msg) {
    J.state = state;
    J.error = error;
    J.err = msg;
    J.t_done = S->now();
    auto it = S->scenes.find(J.scene_id);
    if (it != S->scenes.end()) {
        it->second.users--;
        if (J.bjob >= 0) it->second.running--;
    }
    if (state == PPM_JOB_DONE) {
        S->st.done++;
        S->st.photon_segments += (uint64_t)J.prm.photon_width * J.prm.photon_height * J.target;
    } else {
        S->st.failed++;
    }
    S->cv_done.notify_all();
}

// Drops the device copy of scenes that no resident job renders when they are removed or when
// `need` more bytes are wanted, and forgets removed scenes no job names.  Caller holds mu.
void drop_scenes_locked(ppm_server* S, size_t need) {
    for (auto it = S->scenes.begin(); it != S->scenes.end();) {
        SScene& sc = it->second;
        if (sc.dev && sc.running == 0 && (sc.removed || S->device_bytes + need > S->cfg.hbm_budget)) {
            S->device_bytes -= sc.dev_bytes;
            sc.dev.reset();
            sc.dev_bytes = 0;
        }
        if (sc.removed && sc.users == 0 && !sc.dev) it = S->scenes.erase(it);
        else ++it;
    }
}

void admit(ppm_server* S) {
    for (;;) {
        std::shared_ptr<SJob> J;
        std::shared_ptr<ppm_internal::PackedScene> host;
        bool need_upload = false;
        size_t need = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if ((int)S->resident.size() >= S->cfg.max_jobs || S->queue.empty()) return;
            // first come first served among the jobs that fit; if none does, scenes that no
            // resident job renders give their memory back and the search runs once more
            for (int attempt = 0; attempt < 2 && !J; attempt++) {
                size_t head_need = 0;
                for (auto q = S->queue.begin(); q != S->queue.end(); ++q) {
                    auto& cand = S->jobs[*q];
                    SScene& sc = S->scenes[cand->scene_id];
                    const size_t n = cand->footprint + (sc.dev ? 0 : sc.host->device_bytes());
                    if (q == S->queue.begin()) head_need = n;
                    if (S->device_bytes + n > S->cfg.hbm_budget) continue;
                    J = cand;
                    host = sc.host;
                    need_upload = !sc.dev;
                    need = n;
                    S->queue.erase(q);
                    break;
                }
                if (!J) drop_scenes_locked(S, head_need);
            }
            if (J) S->scenes[J->scene_id].running++;
        }
        if (!J) return;
        // the pool may hold cached blocks of departed jobs: keep what it holds within the budget
        if (S->b->arena->held_bytes() + need > S->cfg.hbm_budget)
            S->b->arena->trim(S->cfg.hbm_budget > need ? S->cfg.hbm_budget - need : 0);
        int rc = PPM_OK;
        std::string msg;
        std::shared_ptr<ppm_internal::SceneDev> dev;
        if (need_upload) {
            rc = ppm_internal::upload_scene(*host, S->b->arena, S->b->upload, S->b->stream, &dev, &msg);
            if (hipStreamSynchronize(S->b->upload) != hipSuccess && rc == PPM_OK) {
                rc = PPM_E_HIP;
                msg = "scene upload failed";
            }
            if (rc == PPM_OK) {
                std::lock_guard<std::mutex> lk(S->mu);
                SScene& sc = S->scenes[J->scene_id];
                sc.dev = dev;
                sc.dev_bytes = dev->bytes;
                S->device_bytes += dev->bytes;
            }
        } else {
            std::lock_guard<std::mutex> lk(S->mu);
            dev = S->scenes[J->scene_id].dev;
        }
        int32_t bjob = -1;
        if (rc == PPM_OK) {
            rc = ppm_internal::batch_add_resident(S->b, dev, J->cam, &J->prm, J->target, &bjob);
            if (rc != PPM_OK) msg = S->b->err;
        }
        std::lock_guard<std::mutex> lk(S->mu);
        if (rc != PPM_OK) {
            S->scenes[J->scene_id].running--;
            finish_locked(S, *J, PPM_JOB_FAILED, rc, msg);
            continue;
        }
        J->state = PPM_JOB_RUNNING;
        J->bjob = bjob;
        J->t_admit = S->now();
        S->device_bytes += J->footprint;
        S->resident.push_back(J->id);
        S->by_bjob[bjob] = J->id;
    }
}

// Results of a finished job, then its departure from the batch.
void retire(ppm_server* S, int64_t id) {
    std::shared_ptr<SJob> J;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J = S->jobs[id];
    }
    const size_t Q = (size_t)J->prm.image_width * J->prm.image_height;
    std::vector<float4> rad(Q), fr(Q), en(Q);
    std::vector<uchar4> px(Q);
    ppm_counters k;
    int rc = ppm_internal::batch_fetch(S->b, J->bjob, rad.data(), fr.data(), en.data(), px.data(), &k);
    const std::string msg = rc ? S->b->err : std::string();
    (void)ppm_batch_remove(S->b, J->bjob);
    std::lock_guard<std::mutex> lk(S->mu);
    S->by_bjob.erase(J->bjob);
    for (size_t i = 0; i < S->resident.size(); i++)
        if (S->resident[i] == id) {
            S->resident.erase(S->resident.begin() + (long)i);
            break;
        }
    S->device_bytes -= J->footprint;
    J->radiance.swap(rad);
    J->flux_radius.swap(fr);
    J->emission.swap(en);
    J->rgba8.swap(px);
    J->counters = k;
    J->passes = k.frame_count;
    finish_locked(S, *J, rc ? PPM_JOB_FAILED : PPM_JOB_DONE, rc, msg);
}

void fail_all(ppm_server* S, int code, const std::string& msg) {
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& kv : S->jobs) {
        SJob& J = *kv.second;
        if (J.state == PPM_JOB_QUEUED || J.state == PPM_JOB_RUNNING) finish_locked(S, J, PPM_JOB_FAILED, code, msg);
    }
    S->queue.clear();
}

void worker_main(ppm_server* S) {
    (void)hipSetDevice(S->cfg.device);
    constexpr int kSteps = ppm_server::kSteps;
    int head = 0, in_flight = 0;  // the oldest step's status slot, steps queued and not yet read
    std::vector<int32_t> snap_ids[kSteps];
    double busy_mark = -1.0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            while (!S->stop && S->queue.empty() && S->resident.empty() && in_flight == 0) S->cv_work.wait(lk);
            if (S->stop) break;
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            drop_scenes_locked(S, 0);
        }
        admit(S);
        // queue the next step while a status slot is free and some resident job has passes left
        // (the batch's mirror of the queued passes), so that a step is always queued behind the
        // one running while the host admits, reads statuses and copies results out
        bool queued = false;
        if (in_flight < kSteps) {
            bool work_left = false;
            {
                std::lock_guard<std::mutex> lk(S->mu);
                for (int64_t id : S->resident)
                    if (ppm_internal::batch_job_has_passes(S->b, S->jobs[id]->bjob)) {
                        work_left = true;
                        break;
                    }
                const double t = S->now();
                if (busy_mark >= 0.0) S->st.busy_seconds += t - busy_mark;
                busy_mark = S->resident.empty() ? -1.0 : t;
            }
            if (work_left) {
                const int slot = (head + in_flight) % kSteps;
                int rc = ppm_batch_pass(S->b, S->cfg.passes_per_step);
                if (rc == PPM_OK)
                    rc = ppm_internal::batch_snapshot(S->b, S->d_status[slot], S->h_status[slot], S->ev_status[slot],
                                                      &snap_ids[slot]);
                if (rc != PPM_OK) {
                    fail_all(S, rc, S->b->err);
                    break;
                }
                in_flight++;
                queued = true;
                std::lock_guard<std::mutex> lk(S->mu);
                S->st.steps++;
            }
        }
        // read the oldest step's status once the pipeline is full or nothing more was queued
        std::vector<int64_t> finished;
        if (in_flight > 0 && (in_flight == kSteps || !queued)) {
            if (hipEventSynchronize(S->ev_status[head]) != hipSuccess) {
                fail_all(S, PPM_E_HIP, "device error while running the batch");
                break;
            }
            std::lock_guard<std::mutex> lk(S->mu);
            const double t = S->now();
            const std::vector<int32_t>& ids = snap_ids[head];
            const unsigned long long* st = S->h_status[head];
            for (size_t i = 0; i < ids.size(); i++) {
                auto it = S->by_bjob.find(ids[i]);
                if (it == S->by_bjob.end()) continue;
                SJob& J = *S->jobs[it->second];
                const unsigned long long frames = st[2 * i], left = st[2 * i + 1];
                J.passes = frames;
                if (frames >= 1 && J.t_first < 0.0) J.t_first = t;
                if (left == 0ull && frames >= J.target && J.state == PPM_JOB_RUNNING) finished.push_back(J.id);
            }
            head = (head + 1) % kSteps;
            in_flight--;
        }
        for (int64_t id : finished) retire(S, id);
    }
    for (int k = 0; k < in_flight; k++) (void)hipEventSynchronize(S->ev_status[(head + k) % kSteps]);
}

}  // namespace

extern "C" {

void ppm_server_config_default(ppm_server_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->hbm_budget = 200ull << 30;
    cfg->max_jobs = 256;
    cfg->passes_per_step = 8;
    cfg->passes_per_graph = 0;
}

int ppm_server_create(const ppm_server_config* cfg, ppm_server** out) {
    if (!out) return serr(nullptr, PPM_E_INVAL, "null output");
    *out = nullptr;
    ppm_server_config d;
    ppm_server_config_default(&d);
    if (!cfg) cfg = &d;
    if (cfg->max_jobs <= 0 || cfg->passes_per_step == 0 || cfg->hbm_budget == 0)
        return serr(nullptr, PPM_E_INVAL, "max_jobs, passes_per_step and hbm_budget must be positive");
    ppm_batch_config bc;
    ppm_batch_config_default(&bc);
    bc.device = cfg->device;
    bc.passes_per_graph = cfg->passes_per_graph;
    ppm_batch* b = nullptr;
    int rc = ppm_batch_create(&bc, &b);
    if (rc != PPM_OK) return serr(nullptr, rc, ppm_batch_last_error(nullptr));
    ppm_server* S = new (std::nothrow) ppm_server();
    if (!S) {
        ppm_batch_destroy(b);
        return serr(nullptr, PPM_E_OOM, "out of host memory");
    }
    S->cfg = *cfg;
    S->b = b;
    std::memset(&S->st, 0, sizeof(S->st));
    S->t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSuccess;
    for (int k = 0; k < ppm_server::kSteps && e == hipSuccess; k++) {
        e = hipMalloc(&S->d_status[k], (size_t)cfg->max_jobs * 16);
        if (e == hipSuccess) e = hipHostMalloc(&S->h_status[k], (size_t)cfg->max_jobs * 16, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&S->ev_status[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        g_server_error = std::string("server buffers: ") + hipGetErrorString(e);
        for (int k = 0; k < ppm_server::kSteps; k++) {
            if (S->ev_status[k]) (void)hipEventDestroy(S->ev_status[k]);
            if (S->d_status[k]) (void)hipFree(S->d_status[k]);
            if (S->h_status[k]) (void)hipHostFree(S->h_status[k]);
        }
        ppm_batch_destroy(b);
        delete S;
        return PPM_E_HIP;
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return PPM_OK;
}

int ppm_server_add_scene(ppm_server* S, const ppm_scene* scene, int64_t* scene_id) {
    if (!S || !scene || !scene_id) return serr(S, PPM_E_INVAL, "null argument");
    int rc;
    if ((rc = ppm_internal::check_scene_public(scene))) return serr(S, rc, ppm_internal::create_error());
    auto host = std::make_shared<ppm_internal::PackedScene>();
    if ((rc = ppm_internal::pack_scene(scene, host.get()))) return serr(S, rc, ppm_internal::create_error());
    std::lock_guard<std::mutex> lk(S->mu);
    const int64_t id = S->next_scene++;
    SScene& sc = S->scenes[id];
    sc.host = host;
    *scene_id = id;
    return PPM_OK;
}

int ppm_server_remove_scene(ppm_server* S, int64_t scene_id) {
    if (!S) return PPM_E_INVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->scenes.find(scene_id);
    if (it == S->scenes.end() || it->second.removed) {
        S->err = "no such scene";
        return PPM_E_INVAL;
    }
    it->second.removed = true;  // the worker drops it once no job uses it
    S->cv_work.notify_all();
    return PPM_OK;
}

int ppm_server_submit(ppm_server* S, int64_t scene_id, const ppm_camera* camera, const ppm_params* params,
                      uint32_t target_passes, int64_t* job) {
    if (!S || !params || !job) return serr(S, PPM_E_INVAL, "null argument");
    *job = -1;
    int rc;
    if ((rc = ppm_internal::validate_params(params))) return serr(S, rc, ppm_internal::create_error());
    if (params->device != S->cfg.device) return serr(S, PPM_E_INVAL, "params.device differs from the server's");
    if (target_passes == 0 || target_passes == PPM_PASSES_UNLIMITED)
        return serr(S, PPM_E_INVAL, "target_passes must be finite and positive");
    auto J = std::make_shared<SJob>();
    J->scene_id = scene_id;
    J->prm = *params;
    J->target = target_passes;
    J->footprint = ppm_internal::state_footprint(params);
    std::memset(&J->counters, 0, sizeof(J->counters));
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->scenes.find(scene_id);
    if (it == S->scenes.end() || it->second.removed) {
        S->err = "no such scene";
        return PPM_E_INVAL;
    }
    if (J->footprint + it->second.host->device_bytes() > S->cfg.hbm_budget) {
        S->err = "the job does not fit the server's HBM budget";
        return PPM_E_OOM;
    }
    J->cam = camera ? *camera : it->second.host->camera;
    it->second.users++;
    J->id = S->next_job++;
    J->t_submit = S->now();
    S->jobs[J->id] = J;
    S->queue.push_back(J->id);
    *job = J->id;
    S->cv_work.notify_all();
    return PPM_OK;
}

int ppm_server_wait(ppm_server* S, int64_t job, double timeout_s) {
    if (!S) return PPM_E_INVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) {
        S->err = "no such job";
        return PPM_E_INVAL;
    }
    std::shared_ptr<SJob> J = it->second;
    auto finished = [&] { return J->state == PPM_JOB_DONE || J->state == PPM_JOB_FAILED; };
    if (timeout_s < 0.0) {
        S->cv_done.wait(lk, finished);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) {
        return 1;
    }
    if (J->state == PPM_JOB_FAILED) {
        S->err = J->err;
        return J->error ? J->error : PPM_E_HIP;
    }
    return PPM_OK;
}

int ppm_server_info(ppm_server* S, int64_t job, ppm_job_info* info) {
    if (!S || !info) return PPM_E_INVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) {
        S->err = "no such job";
        return PPM_E_INVAL;
    }
    const SJob& J = *it->second;
    info->state = J.state;
    info->target_passes = J.target;
    info->passes_done = J.passes;
    info->footprint = J.footprint;
    info->t_submit = J.t_submit;
    info->t_admit = J.t_admit;
    info->t_first_pass = J.t_first;
    info->t_done = J.t_done;
    return PPM_OK;
}

int ppm_server_result(ppm_server* S, int64_t job, ppm_buf which, void* dst, size_t bytes) {
    if (!S || !dst) return PPM_E_INVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end() || it->second->state != PPM_JOB_DONE) {
        S->err = "no finished job with this id";
        return PPM_E_INVAL;
    }
    const SJob& J = *it->second;
    const void* src = nullptr;
    size_t n = 0;
    switch (which) {
        case PPM_BUF_RADIANCE: src = J.radiance.data(); n = J.radiance.size() * 16; break;
        case PPM_BUF_FLUX_RADIUS: src = J.flux_radius.data(); n = J.flux_radius.size() * 16; break;
        case PPM_BUF_EMISSION_COUNT: src = J.emission.data(); n = J.emission.size() * 16; break;
        case PPM_BUF_RGBA8: src = J.rgba8.data(); n = J.rgba8.size() * 4; break;
        case PPM_BUF_COUNTERS: src = &J.counters; n = sizeof(J.counters); break;
        default: S->err = "the server keeps radiance, rgba8, flux_radius, emission_count and counters"; return PPM_E_INVAL;
    }
    if (bytes != n) {
        S->err = "byte count does not match the buffer size";
        return PPM_E_INVAL;
    }
    std::memcpy(dst, src, n);
    return PPM_OK;
}

int ppm_server_release(ppm_server* S, int64_t job) {
    if (!S) return PPM_E_INVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end() || (it->second->state != PPM_JOB_DONE && it->second->state != PPM_JOB_FAILED)) {
        S->err = "no finished job with this id";
        return PPM_E_INVAL;
    }
    S->jobs.erase(it);
    return PPM_OK;
}

int ppm_server_stats_get(ppm_server* S, ppm_server_stats* out) {
    if (!S || !out) return PPM_E_INVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->queued = (int32_t)S->queue.size();
    out->resident = (int32_t)S->resident.size();
    out->device_bytes = S->device_bytes;
    return PPM_OK;
}

double ppm_server_now(const ppm_server* S) { return S ? S->now() : -1.0; }

void ppm_server_destroy(ppm_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    fail_all(S, PPM_E_INVAL, "server stopped");
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto& kv : S->scenes) kv.second.dev.reset();
    }
    ppm_batch_destroy(S->b);
    for (int k = 0; k < ppm_server::kSteps; k++) {
        if (S->ev_status[k]) (void)hipEventDestroy(S->ev_status[k]);
        if (S->d_status[k]) (void)hipFree(S->d_status[k]);
        if (S->h_status[k]) (void)hipHostFree(S->h_status[k]);
    }
    delete S;
}

const char* ppm_server_last_error(const ppm_server* S) { return S ? S->err.c_str() : g_server_error.c_str(); }

}  // extern "C"
