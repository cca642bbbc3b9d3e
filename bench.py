"""Throughput of the SPPM pass loop on one MI355X.

A step is one pass (m_display, glslppm.cpp:214-536) over synthetic input: photon trace of every
photon slot, hash build, gather and progressive update of every hit point, draw, and the eye
trace every max_path_length passes.  The default workload is the largest single-GPU one, the
caustic-heavy terrain: a 1024x1024-cell value-noise heightfield plus 8 glass icospheres (2,260,994
triangles), a small area light, 1920x1080 eye rays, 2048^2 photons per pass, 2^24 hash slots, 12
bounces, alpha 0.7.  spheres-1024 (the Cornell room plus 40 icospheres, ~200k triangles, 1024^2
everything, 10 bounces) and cornell-256 (36 triangles, 256^2, 8 bounces) are the smaller ones.

Prints one JSON line.  value = photon segments per second (P x passes / s) with all state resident
in HBM; eye rays/s, passes/s, emitted photons/s and the reference's "M paths/s" are reported beside
it.  The roofline entry is the dominant kernel's algorithmic bytes per launch over its average
launch time, measured with HIP events on the context stream inside the timed region.  The CPU
baseline is the oracle (strict fp32 restatement, OpenMP) timed on this host on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photons/s + eye-rays/s per pass; p50/p99 time-to-target-passes under load"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of work: the compulsory state traffic of one pass, each array read and
# written once in its minimal fp32 payload (B_pass = 216 P + 16 H + 120 Q + 104 Q / L):
#   photon segment (photon stage): state in 72 B (pos, dir, flux, hit normal 4 x 12 B, generator
#     16 B, triangle 4 B, bounce 4 B) + state out 72 B + hash and bin 36 B + its representative
#     record 36 B = 216 B
#   hit point per pass (k_gather, with the draw fused): gather and update 80 B (60 read, 20
#     written) + draw 40 B = 120 B
#   eye ray (eye stage, every L passes): 104 B
#   hash slot: clear plus read of count and representative, 16 B.  The photon stage clears the
#   next pass's table (8 B per slot, written); the gather reads the slots it needs.
# The photon-stage launch of a pass also runs the gather of the pass before it (every launch of a
# ppm_pass call but the first, whose pass before was gathered at the end of the call before), so
# its bytes per launch are 216 P + 8 H + 120 Q x (launches with a gather) / launches; k_gather
# runs the gather of the call's last pass.
BYTES_PER = {"photon": 216.0, "gather": 120.0, "eye": 104.0, "clear": 0.0}
SLOT_BYTES_PER = {"photon": 8.0, "gather": 0.0, "eye": 0.0, "clear": 0.0}
KERNEL_NAMES = {"photon": "photon stage with the gather of the pass before (k_photon_stage)",
                "gather": "k_gather (the last pass of a call)",
                "eye": "eye stage (k_eye_begin + k_eye + k_grid)", "clear": "k_pass_begin (the pass's uniforms)"}

WORKLOADS = {
    "spheres-1024": dict(kind=1, seed=0x67736C70, count=40, detail=4, res=1024, L=10, camera=((0, 0, 1.667), (0, 0, 0), 45.0),
                         cpu_passes=50, desc="Cornell room + 40 icospheres (subdiv 4): Q=P=H=1024^2, 10 bounces, alpha 0.7"),
    "cornell-256": dict(kind=0, seed=0, count=0, detail=0, res=256, L=8, camera=((0, 0, 1.667), (0, 0, 0), 45.0),
                        fixed_radius=0.02, cpu_passes=1500,
                        desc="Cornell room, 36 triangles: Q=P=H=256^2, 8 bounces, radius 1/50 of the room"),
    "terrain-c2": dict(kind=2, seed=0x67736C72, count=8, detail=1024, res=2048, L=12,
                       camera=((0.0, 0.55, 1.9), (0.0, -0.25, 0.0), 50.0), image=(1920, 1080), hash=(4096, 4096),
                       cpu_passes=12,
                       desc="1024^2-cell value-noise heightfield + 8 glass icospheres (subdiv 5), 2,260,994 "
                            "triangles, small area light: Q=1920x1080, P=2048^2, H=2^24, 12 bounces"),
}


BATCH_WORKLOADS = {
    "batch-mix64": "64 seeded scenes (room + icospheres), triangles log-uniform 1k-500k, resolution uniform "
                   "over 128^2/256^2/512^2/1024^2, P = Q, H = Q, 10 bounces, passes uniform 64-256",
    "batch-cornell64": "64 Cornell rooms of 36 triangles with distinct generator seeds: Q=P=H=256^2, 8 bounces, "
                       "radius 1/50 of the room, 64 passes each",
}


def batch_jobs(name, n_jobs=64):
    """(mesh, params, passes) of every job of a batch workload; scenes built on a thread pool."""
    import concurrent.futures as cf

    import numpy as np

    import glslppm_amd as g
    specs = []
    for k in range(n_jobs):
        if name == "batch-cornell64":
            specs.append(dict(kind=0, seed=0, count=0, detail=0, res=256, L=8, passes=64, fixed_radius=0.02,
                              xs=k))
        else:
            rs = np.random.RandomState(0x67736C70 + k)
            T = float(np.exp(rs.uniform(np.log(1000.0), np.log(500000.0))))
            detail = 3 if T < 20000 else 4
            count = max(1, int(round((T - 36) / (20 * 4 ** detail))))
            res = int(rs.choice([128, 256, 512, 1024]))
            specs.append(dict(kind=1, seed=0x67736C70 + k, count=count, detail=detail, res=res, L=10,
                              passes=int(rs.randint(64, 257)), fixed_radius=0.0, xs=k))

    def make(sp):
        m = g.Mesh.generate(sp["kind"], seed=sp["seed"], count=sp["count"], detail=sp["detail"])
        m.set_camera((0, 0, 1.667), (0, 0, 0), sp["res"], sp["res"], 45.0).build()
        kw = dict(max_path_length=sp["L"])
        if sp["fixed_radius"]:
            kw["fixed_radius"] = sp["fixed_radius"]
        p = g.default_params(sp["res"], **kw)
        p.xorshift[0] = (123456789 + 7919 * sp["xs"]) & 0xFFFFFFFF
        return m, p, sp["passes"]

    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(make, specs))


def amdahl_bound(plain_s, plain_by_res, work_by_res):
    """Upper bound on the batch's speedup over the plain loop: the jobs with the largest image in
    the mix (1024^2 here, a million photons per pass) already keep the whole GPU busy in the plain
    loop, so their photon segments per second there is the rate the GPU sustains on these scenes;
    batching hides the small jobs' idle launch gaps but must still do their work, so the batch takes
    at least (all photon segments) / (that rate).  None when no job is that large: smaller jobs do
    not fill the GPU alone, so their plain-loop rate bounds nothing."""
    big = max(plain_by_res)
    if int(big) < 1024 or plain_by_res[big] <= 0 or work_by_res.get(big, 0) <= 0:
        return None
    rate = work_by_res[big] / plain_by_res[big]
    return plain_s / (sum(work_by_res.values()) / rate)


def run_batch_workload(args):
    """Aggregate photon segments/s of one batch against the plain one-scene loop, job after job,
    over the same jobs and passes; scene upload and state creation are outside both timings."""
    import glslppm_amd as g
    t = time.time()
    jobs = batch_jobs(args.workload)
    build_s = time.time() - t
    work = sum(p.photon_width * p.photon_height * n for (_, p, n) in jobs)
    eye = sum(p.image_width * p.image_height * ((n + p.max_path_length - 1) // p.max_path_length) for (_, p, n) in jobs)
    tris = [m.num_triangles for (m, _, _) in jobs]

    # plain loop: one Renderer per job, its passes, synchronised
    plain_s = 0.0
    plain_by_res, work_by_res = {}, {}
    for (m, p, n) in jobs:
        r = g.Renderer(m, p)
        r.sync()
        t0 = time.perf_counter()
        r.run(n)
        r.sync()
        dt = time.perf_counter() - t0
        plain_s += dt
        plain_by_res[p.image_width] = plain_by_res.get(p.image_width, 0.0) + dt
        work_by_res[p.image_width] = work_by_res.get(p.image_width, 0) + p.photon_width * p.photon_height * n
        r.close()

    b = g.Batch(passes_per_graph=args.passes_per_graph)
    if args.tuning:
        b.set_tuning(**json.loads(args.tuning))
    ids = [b.add(m, p, n) for (m, p, n) in jobs]
    most = max(n for (_, _, n) in jobs)
    b.run(0)  # builds the tables and the graph
    b.sync()
    t0 = time.perf_counter()
    b.run(most)
    b.sync()
    batch_s = time.perf_counter() - t0
    for j, (_, _, n) in zip(ids, jobs):
        assert b.progress(j) == (n, 0)
    # spot check: the largest job equals itself run alone
    big = max(range(len(jobs)), key=lambda i: jobs[i][1].photon_width * jobs[i][1].photon_height)
    r = g.Renderer(jobs[big][0], jobs[big][1])
    r.run(jobs[big][2])
    same = bool((r.readback("flux_radius") == b.readback(ids[big], "flux_radius")).all()
                and (r.readback("radiance") == b.readback(ids[big], "radiance")).all())
    r.close()
    hbm = b.device_bytes
    b.close()
    line = {
        "metric": METRIC,
        "value": work / batch_s,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": most,
        "warmup": 0,
        "ms_per_step": batch_s * 1e3 / most,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, xorshift128 seeds of the reference)",
        "config": {"workload": BATCH_WORKLOADS[args.workload], "jobs": len(jobs), "triangles_min": min(tris),
                   "triangles_max": max(tris), "photon_segments": work, "eye_rays": eye,
                   "passes_per_graph": args.passes_per_graph, "scene_build_s": round(build_s, 2),
                   "batch_hbm_bytes": hbm},
        "eye_rays_per_s": eye / batch_s,
        "plain_loop": {"value": work / plain_s, "unit": "photons/s", "seconds": plain_s,
                       "seconds_by_resolution": {str(k): round(v, 4) for k, v in sorted(plain_by_res.items())}},
        "batch_seconds": batch_s,
        "speedup_vs_plain_loop": plain_s / batch_s,
        "amdahl_bound": amdahl_bound(plain_s, plain_by_res, work_by_res),
        "largest_job_equal_alone": same,
    }
    print(json.dumps(line), flush=True)


SERVE_DESC = ("Poisson arrivals at RATE jobs/s for SECONDS s; each job one of 64 seeded scenes (room + icospheres, "
              "triangles log-uniform 1k-500k, resolution 128^2-1024^2, P = H = Q, 10 bounces) with a random camera "
              "(position uniform in the room, FOV uniform 40-70 degrees) and a target pass count log-uniform 16-512")


def run_serve_workload(args):
    """Continuous batching under a Poisson request stream: latency percentiles and throughput,
    beside the plain one-scene loop timed on a sample of the same jobs."""
    import threading

    import numpy as np

    import glslppm_amd as g
    t = time.time()
    pool = batch_jobs("batch-mix64")
    build_s = time.time() - t
    srv = g.Server(hbm_budget=args.hbm_budget << 30, passes_per_step=args.passes_per_step,
                   passes_per_graph=args.passes_per_graph, max_jobs=args.max_jobs)
    sids = [srv.add_scene(m) for (m, _, _) in pool]
    rs = np.random.RandomState(0x5E2E)
    n_jobs = int(rs.poisson(args.rate * args.serve_seconds))
    arrivals = np.sort(rs.uniform(0.0, args.serve_seconds, n_jobs))
    specs = []
    for k in range(n_jobs):
        si = int(rs.randint(len(pool)))
        res = pool[si][1].image_width
        while True:
            o = rs.uniform(-0.45, 0.45, 3)
            if np.linalg.norm(o) > 0.15:
                break
        fov = float(rs.uniform(40.0, 70.0))
        target = int(round(float(np.exp(rs.uniform(np.log(16.0), np.log(512.0))))))
        p = g.default_params(res, max_path_length=10)
        p.xorshift[0] = (123456789 + 104729 * k) & 0xFFFFFFFF
        specs.append((si, (tuple(o), (0.0, 0.0, 0.0), res, res, fov), p, target))
    cams = [g.make_camera(*sp[1]) for sp in specs]
    ids = [None] * n_jobs
    t0 = srv.now()

    def feeder():
        for k in range(n_jobs):
            delay = t0 + arrivals[k] - srv.now()
            if delay > 0:
                time.sleep(delay)
            si, _, p, target = specs[k]
            ids[k] = srv.submit(sids[si], p, target, camera=cams[k])
            if (k + 1) % 500 == 0:
                st = srv.stats()
                print("serve: %d of %d jobs submitted at %.1f s, %d done, %d resident, %d queued" % (
                    k + 1, n_jobs, srv.now() - t0, st["done"], st["resident"], st["queued"]), file=sys.stderr, flush=True)

    # jobs are collected in arrival order as they finish; their host copies are released at once,
    # except those of the three jobs checked against the job alone below
    keep = {0, n_jobs // 2, n_jobs - 1}
    infos = [None] * n_jobs
    th = threading.Thread(target=feeder)
    th.start()
    for k in range(n_jobs):
        while ids[k] is None:
            time.sleep(0.001)
        srv.wait(ids[k])
        infos[k] = srv.info(ids[k])
        if k not in keep:
            srv.release(ids[k])
    th.join()
    st = srv.stats()
    first = np.array([i["t_first_pass"] - i["t_submit"] for i in infos])
    target_t = np.array([i["t_done"] - i["t_submit"] for i in infos])
    work = sum(specs[k][2].photon_width * specs[k][2].photon_height * specs[k][3] for k in range(n_jobs))
    span = max(i["t_done"] for i in infos) - min(i["t_submit"] for i in infos)
    # alone versus in the mix, bit for bit, on three jobs
    same = True
    for k in (0, n_jobs // 2, n_jobs - 1):
        si, cam, p, target = specs[k]
        m = pool[si][0]
        m.set_camera(*cam)
        r = g.Renderer(m, p)
        r.run(target)
        same &= bool((r.readback("radiance") == srv.result(ids[k], "radiance")).all())
        same &= bool((r.readback("flux_radius") == srv.result(ids[k], "flux_radius")).all())
        r.close()
    srv.close()
    # the plain one-scene loop on a sample of the same jobs, one after another
    sample = list(range(0, n_jobs, max(1, n_jobs // args.plain_sample)))
    plain_s, plain_work = 0.0, 0
    plain_by_res, work_by_res = {}, {}
    for k in sample:
        si, cam, p, target = specs[k]
        m = pool[si][0]
        m.set_camera(*cam)
        r = g.Renderer(m, p)
        r.sync()
        t1 = time.perf_counter()
        r.run(target)
        r.sync()
        dt = time.perf_counter() - t1
        plain_s += dt
        plain_by_res[p.image_width] = plain_by_res.get(p.image_width, 0.0) + dt
        work_by_res[p.image_width] = work_by_res.get(p.image_width, 0) + p.photon_width * p.photon_height * target
        plain_work += p.photon_width * p.photon_height * target
        r.close()
    pct = lambda a, q: float(np.percentile(a, q))
    line = {
        "metric": METRIC,
        "value": work / span,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": int(st["steps"]),
        "warmup": 0,
        "ms_per_step": span * 1e3 / max(1, st["steps"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes, cameras, arrivals and targets)",
        "config": {"workload": SERVE_DESC.replace("RATE", str(args.rate)).replace("SECONDS", str(args.serve_seconds)),
                   "jobs": n_jobs, "hbm_budget_gib": args.hbm_budget, "passes_per_step": args.passes_per_step,
                   "passes_per_graph": args.passes_per_graph, "max_jobs": args.max_jobs,
                   "scene_build_s": round(build_s, 2)},
        "time_to_first_pass_s": {"p50": pct(first, 50), "p99": pct(first, 99), "max": float(first.max())},
        "time_to_target_s": {"p50": pct(target_t, 50), "p99": pct(target_t, 99), "max": float(target_t.max())},
        "span_s": span,
        "gpu_busy_fraction": st["busy_seconds"] / span,
        "plain_loop": {"value": plain_work / plain_s, "unit": "photons/s", "jobs": len(sample),
                       "seconds_by_resolution": {str(k): round(v, 4) for k, v in sorted(plain_by_res.items())}},
        "speedup_vs_plain_loop": (work / span) / (plain_work / plain_s),
        "amdahl_bound": amdahl_bound(plain_s, plain_by_res, work_by_res),
        "jobs_equal_alone": same,
    }
    print(json.dumps(line), flush=True)


def build_scene(w):
    import glslppm_amd as g
    m = g.Mesh.generate(w["kind"], seed=w["seed"], count=w["count"], detail=w["detail"])
    (org, at, fov) = w["camera"]
    iw, ih = w.get("image", (w["res"], w["res"]))
    m.set_camera(org, at, iw, ih, fov)
    t = time.time()
    m.build()
    return m, time.time() - t


def params_for(w):
    import glslppm_amd as g
    kw = dict(max_path_length=w["L"])
    if w.get("fixed_radius"):
        kw["fixed_radius"] = w["fixed_radius"]
    if "image" in w:
        kw["image_width"], kw["image_height"] = w["image"]
    if "hash" in w:
        kw["hash_width"], kw["hash_height"] = w["hash"]
    return g.default_params(w["res"], **kw)


def cpu_baseline(mesh, params, passes):
    """Oracle on this host's cores over the first `passes` passes of the same workload."""
    import oracle
    o = oracle.OracleRenderer(mesh, params)
    t = time.time()
    o.run(passes)
    dt = time.time() - t
    P = params.photon_width * params.photon_height
    return {"value": P * passes / dt, "unit": "photons/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": "first %d passes (%d eye passes, %d photon passes) of the same workload, %.1f s" % (
                passes, (passes + params.max_path_length - 1) // params.max_path_length, passes, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="terrain-c2",
                    choices=sorted(WORKLOADS) + sorted(BATCH_WORKLOADS) + ["serve"])
    ap.add_argument("--rate", type=float, default=20.0, help="serve: job arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="serve: length of the arrival window")
    ap.add_argument("--hbm-budget", type=int, default=200, help="serve: device memory budget in GiB")
    ap.add_argument("--passes-per-step", type=int, default=8, help="serve: batched passes between admissions")
    ap.add_argument("--plain-sample", type=int, default=40, help="serve: jobs timed in the plain loop")
    ap.add_argument("--max-jobs", type=int, default=256, help="serve: resident jobs at most")
    ap.add_argument("--passes-per-graph", type=int, default=0,
                    help="batch and serve workloads: passes per hipGraph (0: plain launches)")
    ap.add_argument("--graph", action="store_true",
                    help="plain workloads: replay passes as hipGraphs (the timed region then carries no timing "
                         "events; the roofline's launch time comes from the stage passes before it)")
    ap.add_argument("--autotune", action="store_true",
                    help="plain workloads: pick launch shapes with ppm_autotune during the warm-up")
    ap.add_argument("--tuning", default=None,
                    help='launch shapes to set first (plain and batch workloads), as JSON, e.g. \'{"walk_stack": 8}\'')
    ap.add_argument("--readback-each-pass", default=None, choices=["rgba8", "radiance"],
                    help="plain workloads: also time --steps passes each followed by a readback of this image "
                         "(ppm_readback synchronises and copies over PCIe); reported as pcie_inclusive")
    ap.add_argument("--cpu-passes", type=int, default=None,
                    help="passes of the CPU baseline sample (0: skip; default: the workload's, about 10-20 s of "
                         "oracle time on 16 cores)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (torch), the known-size reference for "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.workload in BATCH_WORKLOADS:
        run_batch_workload(args)
        return
    if args.workload == "serve":
        run_serve_workload(args)
        return

    import glslppm_amd as g
    w = WORKLOADS[args.workload]
    mesh, build_s = build_scene(w)
    params = params_for(w)
    if args.graph:
        params.flags |= g.FLAG_GRAPH
    r = g.Renderer(mesh, params)
    Q = params.image_width * params.image_height
    P = params.photon_width * params.photon_height
    H = params.hash_width * params.hash_height
    L = params.max_path_length

    tuning = None
    if args.tuning:
        r.set_tuning(**json.loads(args.tuning))
    if args.autotune:
        tuning = r.autotune()
    r.run(args.warmup)
    r.sync()
    # every stage's launch time, from passes of their own (two eye cadences): timing events cost
    # the stream a few microseconds each, so the timed region times only the dominant stage
    r.profile(True)
    r.run(2 * L)
    stages = r.profile_read()
    r.profile(False)
    dom = max(stages, key=lambda k: stages[k][0])
    c0 = r.counters()
    # a captured pass carries no timing events (profiling launches stream by stream), so graph
    # replay runs its timed region untimed and takes the launch time from the stage passes
    r.profile(not args.graph, kinds=[dom])
    t0 = time.perf_counter()
    r.run(args.steps)
    r.sync()
    t1 = time.perf_counter()
    prof = r.profile_read()
    r.profile(False)
    c1 = r.counters()

    wall = t1 - t0
    passes = args.steps
    eye_passes = c1["eye_samples"] - c0["eye_samples"]
    emitted = c1["photons_emitted"] - c0["photons_emitted"]
    value = P * passes / wall
    # the dominant kernel (by total device time over the stage passes), timed over the timed region
    # (over the stage passes with graph replay)
    dom_ms, dom_n = stages[dom] if args.graph else prof[dom]
    units = {"photon": P, "gather": Q, "eye": Q, "clear": H}[dom]
    bytes_per_launch = BYTES_PER[dom] * units + SLOT_BYTES_PER[dom] * H
    if dom == "photon":
        # one ppm_pass call of dom_n passes: all but its first launch carry a gather
        bytes_per_launch += BYTES_PER["gather"] * Q * max(dom_n - 1, 0) / max(dom_n, 1)
    avg_ms = dom_ms / max(dom_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dom, args.workload))
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    # the rate a caller sees when it reads the image back after every pass: each ppm_pass(1) call
    # ends with its own gather, and the readback waits for it and copies the image to the host
    pcie = None
    if args.readback_each_pass:
        img = r.readback(args.readback_each_pass)
        t2 = time.perf_counter()
        for _ in range(args.steps):
            r.run(1)
            img = r.readback(args.readback_each_pass)
        t3 = time.perf_counter()
        pcie = {"value": P * args.steps / (t3 - t2), "unit": "photons/s", "readback": args.readback_each_pass,
                "bytes_per_pass": int(img.nbytes), "ms_per_step": (t3 - t2) * 1e3 / args.steps}

    cpu = None
    cpu_passes = w["cpu_passes"] if args.cpu_passes is None else args.cpu_passes
    if cpu_passes > 0:
        cpu = cpu_baseline(mesh, params, cpu_passes)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / passes,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, xorshift128 seeds of the reference)",
        "config": {"workload": w["desc"], "triangles": mesh.num_triangles, "eye_rays": Q, "photons_per_pass": P,
                   "hash_slots": H, "max_path_length": L, "bvh_build_s": round(build_s, 3)},
        "eye_rays_per_s": Q * eye_passes / wall,
        "passes_per_s": passes / wall,
        "emitted_photons_per_s": emitted / wall,
        "m_paths_per_s": (emitted + Q * eye_passes) / wall / (1024.0 * 1024.0),
        "kernel_ms_per_launch": {k: (v[0] / v[1] if v[1] else None) for k, v in stages.items()},
        "kernel_ms_per_launch_passes": 2 * L,
        "tuning": tuning if tuning is not None else r.tuning(),
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": avg_ms,
                     "launch_ms_from": "stage passes before the timed region (graph replay)" if args.graph
                     else "timed region"},
        "cpu_baseline": cpu,
    }
    if pcie is not None:
        line["pcie_inclusive"] = pcie
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
