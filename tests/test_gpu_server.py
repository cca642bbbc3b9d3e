"""Serving loop (ppm_server_*): jobs arriving while others run give the results they give alone.

Each job's radiance, (tau, R), (E, N), RGBA8 and counters after its target passes are compared,
bit for bit, with a Renderer of the same scene, camera and parameters run alone, and with the
oracle run for the same passes.
"""
import threading
import time

import numpy as np
import pytest

import glslppm_amd as g
import oracle

pytestmark = pytest.mark.gpu


def scenes():
    a = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), 48, 48, 45.0).build()
    b = g.Mesh.generate(g.SCENE_SPHERES, seed=21, count=6, detail=2).set_camera(
        (0, 0, 1.667), (0, 0, 0), 64, 64, 45.0).build()
    return [a, b]


def job_list():
    """(scene index, camera args, params, target passes)."""
    out = []
    cams = [((0.2, -0.1, 1.5), (0, 0, 0), 45.0), ((-0.3, 0.2, 1.2), (0.1, 0, 0), 60.0),
            ((0.0, 0.3, 1.8), (0, -0.1, 0), 40.0)]
    specs = [(0, 48, 48, 5, 9), (1, 64, 64, 6, 17), (0, 40, 24, 4, 3), (1, 32, 32, 8, 26), (0, 48, 48, 5, 12),
             (1, 64, 40, 7, 7)]
    for k, (si, w, h, L, n) in enumerate(specs):
        o, at, fov = cams[k % len(cams)]
        p = g.default_params(64, image_width=w, image_height=h, photon_width=64, photon_height=32, hash_width=64,
                             hash_height=32, max_path_length=L)
        p.xorshift[0] = 123456789 + 31 * k
        out.append((si, (o, at, w, h, fov), p, n))
    return out


def check(server, jid, mesh, cam, params, passes):
    """A served job's results against the job alone in the plain loop and against the oracle."""
    o, at, w, h, fov = cam
    mesh.set_camera(o, at, w, h, fov)
    r = g.Renderer(mesh, params)
    r.run(passes)
    cpu = oracle.OracleRenderer(mesh, params)
    cpu.run(passes)
    for ref, label in ((r, "the job alone"), (cpu, "the oracle")):
        for n in g.SERVER_BUFFERS:
            a, b = server.result(jid, n), ref.readback(n)
            bad = int(np.any(a.view(np.uint8).reshape(a.shape[0], -1) != b.view(np.uint8).reshape(b.shape[0], -1),
                             axis=1).sum())
            assert bad == 0, "job %d: %s differs from %s in %d elements" % (jid, n, label, bad)
        assert server.counters(jid) == ref.counters()
    r.close()
    cpu.close()


def test_jobs_in_the_mix_equal_jobs_alone():
    meshes = scenes()
    srv = g.Server(passes_per_step=3)
    sids = [srv.add_scene(m) for m in meshes]
    jobs = job_list()
    ids = []
    # arrivals spread over the run, from a second thread, while earlier jobs are on the GPU
    def submit(lo, hi, delay):
        time.sleep(delay)
        for (si, cam, p, n) in jobs[lo:hi]:
            ids.append((srv.submit(sids[si], p, n, camera=g.make_camera(*cam)), si, cam, p, n))
    submit(0, 3, 0.0)
    t = threading.Thread(target=submit, args=(3, 6, 0.05))
    t.start()
    t.join()
    for (jid, si, cam, p, n) in ids:
        assert srv.wait(jid, timeout=120.0)
        info = srv.info(jid)
        assert info["state"] == "done" and info["passes_done"] == n
        assert 0.0 <= info["t_submit"] <= info["t_admit"] <= info["t_first_pass"] <= info["t_done"]
    for (jid, si, cam, p, n) in ids:
        print("check job %d (scene %d, %dx%d, L=%d, %d passes)" % (jid, si, cam[2], cam[3], p.max_path_length, n),
              flush=True)
        check(srv, jid, meshes[si], cam, p, n)
    st = srv.stats()
    assert st["done"] == len(ids) and st["failed"] == 0 and st["resident"] == 0 and st["queued"] == 0
    srv.close()


def test_admission_waits_for_memory_and_keeps_results():
    meshes = scenes()
    jobs = job_list()
    states = [g.state_footprint(p) for (_, _, p, _) in jobs]
    scenes_each = [g.scene_footprint(m, jobs[0][2]) - g.state_footprint(jobs[0][2]) for m in meshes]
    scene_bytes = sum(scenes_each)
    # room for both scenes and two jobs at a time, not three; a scene that no resident job renders
    # may give its memory back, which leaves room for a few more jobs of the other scene
    budget = scene_bytes + 2 * max(states) + min(states) // 2
    assert scene_bytes + 3 * min(states) > budget
    most = max(k for k in range(len(jobs) + 1) if min(scenes_each) + k * min(states) <= budget)
    assert most < len(jobs)
    srv = g.Server(hbm_budget=budget, passes_per_step=2)
    sids = [srv.add_scene(m) for m in meshes]
    ids = [(srv.submit(sids[si], p, n, camera=g.make_camera(*cam)), si, cam, p, n) for (si, cam, p, n) in jobs]
    peak = 0
    while not all(srv.info(j)["state"] == "done" for (j, *_rest) in ids):
        st = srv.stats()
        assert st["device_bytes"] <= budget
        peak = max(peak, st["resident"])
        time.sleep(0.002)
    print("budget %d, most resident %d, peak %d" % (budget, most, peak), flush=True)
    assert 1 <= peak <= most
    for (jid, si, cam, p, n) in ids:
        check(srv, jid, meshes[si], cam, p, n)
    assert srv.stats()["device_bytes"] <= budget
    srv.close()


def test_job_larger_than_the_budget_is_refused():
    meshes = scenes()
    srv = g.Server(hbm_budget=1 << 20)
    sid = srv.add_scene(meshes[0])
    with pytest.raises(g.PPMError):
        srv.submit(sid, g.default_params(64), 4)
    srv.close()
