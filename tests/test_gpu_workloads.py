"""The batched and served forms on workloads shaped like their benchmarks, against the oracle.

Batching workload: many seeded scenes (the room plus jittered icospheres) whose triangle counts are
log-uniform over 1k-200k and whose resolutions are drawn from 32^2-256^2 (photons = hash slots =
pixels), each with its own pass budget, advanced together by one launch sequence per pass with
graph replay.  The benchmark form (bench.py --workload batch-mix64) uses 64 scenes of 1k-500k
triangles at 128^2-1024^2 and 64-256 passes; the sizes here are scaled down so that the oracle can
check every job bit for bit.

Serving workload: a Poisson stream of render jobs over a pool of such scenes, each with a random
camera (position uniform in the room, field of view uniform over 40-70 degrees) and a target pass
count log-uniform over 16-64, admitted against a device-memory budget that holds only a few jobs at
a time, so that arrivals queue, are admitted as memory frees up and share batched passes with the
jobs already running.  bench.py --workload serve runs 20 jobs/s for 20 s with targets up to 512
over the 64-scene pool; here every job is compared with the oracle.
"""
import math
import threading
import time

import numpy as np
import pytest

import glslppm_amd as g
import oracle
from parity import STATE_BUFFERS, check_hash_properties, compare_all, format_report

pytestmark = pytest.mark.gpu


def uneven_scene(rs, t_max, res):
    """A room with icospheres, its triangle count drawn log-uniform over [1000, t_max]."""
    T = float(np.exp(rs.uniform(np.log(1000.0), np.log(t_max))))
    detail = 3 if T < 20000 else 4
    count = max(1, int(round((T - 36) / (20 * 4 ** detail))))
    seed = int(rs.randint(1 << 30))
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=seed, count=count, detail=detail)
    return m.set_camera((0, 0, 1.667), (0, 0, 0), res, res, 45.0).build()


class JobView:
    def __init__(self, batch, job):
        self.b, self.j = batch, job

    def readback(self, name):
        return self.b.readback(self.j, name)

    def counters(self):
        return self.b.counters(self.j)

    def grid(self):
        return self.b.grid(self.j)


@pytest.mark.parametrize("passes_per_graph", [4, 0])
def test_batch_of_uneven_scenes_equals_the_oracle_job_by_job(passes_per_graph):
    """passes_per_graph 4: graph replay, each eye stage inside its eye pass; 0: plain launches,
    each job's next eye stage begun ahead beside the photon passes before its eye pass."""
    rs = np.random.RandomState(0x67736C80)
    jobs = []
    for k in range(20):
        res = int(rs.choice([32, 64, 128, 256]))
        m = uneven_scene(rs, 200000.0, res)
        p = g.default_params(res, max_path_length=int(rs.choice([10, 10, 10, 6, 12])))
        p.xorshift[0] = (123456789 + 7919 * k) & 0xFFFFFFFF
        jobs.append((m, p, int(rs.randint(16, 41))))
    tris = [m.num_triangles for (m, _, _) in jobs]
    print("jobs %d, triangles %d-%d" % (len(jobs), min(tris), max(tris)), flush=True)
    b = g.Batch(passes_per_graph=passes_per_graph)
    b.set_tuning(trace_block=128, tile_wave_width=16, passes_per_graph=passes_per_graph)
    ids = [b.add(m, p, n) for (m, p, n) in jobs]
    most = max(n for (_, _, n) in jobs)
    b.run(most + 3)  # every budget runs out; the last passes leave finished jobs untouched
    b.sync()
    for k, ((m, p, n), j) in enumerate(zip(jobs, ids)):
        label = "job %d (%d triangles, %dx%d, L=%d, %d passes)" % (k, m.num_triangles, p.image_width,
                                                                   p.image_height, p.max_path_length, n)
        print("compare " + label, flush=True)
        assert b.progress(j) == (n, 0)
        view = JobView(b, j)
        check_hash_properties(view, p)
        cpu = oracle.OracleRenderer(m, p)
        cpu.run(n)
        rep = compare_all(view, cpu, STATE_BUFFERS)
        bad = format_report(rep)
        assert not bad, "%s differs from the oracle: %s" % (label, bad)
        cpu.close()
    b.close()


def test_poisson_stream_through_admission_equals_the_oracle():
    rs = np.random.RandomState(0x5E2E)
    pool = [uneven_scene(rs, 100000.0, 64) for _ in range(6)]
    # arrivals much faster than jobs finish, so that they queue for memory deterministically
    rate, seconds = 1000.0, 0.04
    n_jobs = int(rs.poisson(rate * seconds))
    arrivals = np.sort(rs.uniform(0.0, seconds, n_jobs))
    specs = []
    for k in range(n_jobs):
        si = int(rs.randint(len(pool)))
        res = int(rs.choice([32, 48, 64]))
        while True:
            o = rs.uniform(-0.45, 0.45, 3)
            if np.linalg.norm(o) > 0.15:
                break
        fov = float(rs.uniform(40.0, 70.0))
        target = int(round(math.exp(rs.uniform(math.log(16.0), math.log(64.0)))))
        p = g.default_params(64, image_width=res, image_height=res, max_path_length=10)
        p.xorshift[0] = (123456789 + 104729 * k) & 0xFFFFFFFF
        specs.append((si, (tuple(o), (0.0, 0.0, 0.0), res, res, fov), p, target))
    # memory for every scene and about three jobs' state at a time
    state = max(g.state_footprint(p) for (_, _, p, _) in specs)
    scenes = sum(g.scene_footprint(m, specs[0][2]) - g.state_footprint(specs[0][2]) for m in pool)
    budget = scenes + 3 * state
    srv = g.Server(hbm_budget=budget, passes_per_step=4)
    sids = [srv.add_scene(m) for m in pool]
    ids = [None] * n_jobs
    t0 = srv.now()

    def feeder():
        for k in range(n_jobs):
            delay = t0 + arrivals[k] - srv.now()
            if delay > 0:
                time.sleep(delay)
            si, cam, p, target = specs[k]
            ids[k] = srv.submit(sids[si], p, target, camera=g.make_camera(*cam))

    th = threading.Thread(target=feeder)
    th.start()
    while th.is_alive() or any(j is None or srv.info(j)["state"] != "done" for j in ids):
        st = srv.stats()
        assert st["device_bytes"] <= budget and st["failed"] == 0
        time.sleep(0.002)
    th.join()
    infos = [srv.info(j) for j in ids]
    first = np.array([i["t_first_pass"] - i["t_submit"] for i in infos])
    done = np.array([i["t_done"] - i["t_submit"] for i in infos])
    # most jobs resident at once (overlap of the [admitted, done) intervals), and jobs admitted only
    # after another job had finished and given its memory back
    events = sorted([(i["t_admit"], 1) for i in infos] + [(i["t_done"], -1) for i in infos], key=lambda e: (e[0], e[1]))
    peak, live = 0, 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    first_done = min(i["t_done"] for i in infos)
    waited = sum(i["t_admit"] >= first_done for i in infos)
    print("jobs %d, most resident at once %d, admitted after a job finished %d" % (n_jobs, peak, waited))
    print("time to first pass p50 %.4f p99 %.4f s, time to target p50 %.4f p99 %.4f s" % (
        np.percentile(first, 50), np.percentile(first, 99), np.percentile(done, 50), np.percentile(done, 99)),
        flush=True)
    assert peak >= 2 and waited > 0
    for k, info in enumerate(infos):
        assert info["passes_done"] == specs[k][3]
        assert info["t_submit"] <= info["t_admit"] <= info["t_first_pass"] <= info["t_done"]
    for k in range(n_jobs):
        si, cam, p, target = specs[k]
        m = pool[si]
        m.set_camera(*cam)
        cpu = oracle.OracleRenderer(m, p)
        cpu.run(target)
        for name in g.SERVER_BUFFERS:
            a, b = srv.result(ids[k], name), cpu.readback(name)
            bad = int(np.any(a.view(np.uint8).reshape(a.shape[0], -1) != b.view(np.uint8).reshape(b.shape[0], -1),
                             axis=1).sum())
            assert bad == 0, "job %d: %s differs from the oracle in %d elements" % (k, name, bad)
        assert srv.counters(ids[k]) == cpu.counters()
        cpu.close()
    srv.close()
