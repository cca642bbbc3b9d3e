"""Batched pass loop (ppm_batch_*) and graph replay against the plain loop and the oracle, bit
for bit.

A job advanced inside a batch -- next to other scenes, resolutions, hash sizes, bounce caps and
pass budgets, added late or removed early -- must end with exactly the buffers of the same scene
and parameters run alone through ppm_create / ppm_pass for the same number of its passes, and
with the oracle's buffers for the same passes.
"""
import pytest

import glslppm_amd as g
import oracle
from parity import STATE_BUFFERS, ShrunkLeafBoxes, compare_all, format_report

pytestmark = pytest.mark.gpu


class JobView:
    """A batch job with the readback interface of a Renderer."""

    def __init__(self, batch, job):
        self.b, self.j = batch, job

    def readback(self, name):
        return self.b.readback(self.j, name)

    def counters(self):
        return self.b.counters(self.j)

    def grid(self):
        return self.b.grid(self.j)


def cornell(w, h=None):
    return g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), w, h or w, 45.0).build()


def spheres(seed, count, detail, w, h):
    m = g.Mesh.generate(g.SCENE_SPHERES, seed=seed, count=count, detail=detail)
    return m.set_camera((0.1, 0.05, 1.6), (0, 0, 0), w, h, 55.0).build()


def ibl(res):
    base = cornell(res).arrays()
    mats = [(rgb, eta, 0 if brdf == -1 else brdf, gl) for (rgb, eta, brdf, gl) in base["materials"]]
    m = g.Mesh.from_arrays(base["positions"], mats, base["material"], base["normals"])
    return m.set_camera((0, 0, 1.667), (0, 0, 0), res, res, 45.0).build()


def job_specs():
    """(label, mesh, params, budget) for a mix of shapes and cadences."""
    return [
        ("cornell-64", cornell(64), g.default_params(64, max_path_length=8), 20),
        ("spheres-48x40", spheres(7, 6, 2, 48, 40),
         g.default_params(64, image_width=48, image_height=40, photon_width=64, photon_height=32, hash_width=128,
                          hash_height=16, max_path_length=6), 13),
        ("cornell-32-unlimited", cornell(32), g.default_params(32, max_path_length=5, fixed_radius=0.02),
         g.PASSES_UNLIMITED),
        ("ibl-48", ibl(48), g.default_params(48, hash_width=64, hash_height=32, max_path_length=5), 7),
    ]


def check_equal_alone(view, mesh, params, passes, label, against_oracle=True):
    """The batch job against the same job alone in the plain loop and, unless told otherwise,
    against the oracle."""
    print("compare %s after %d passes" % (label, passes), flush=True)
    alone = g.Renderer(mesh, params)
    alone.run(passes)
    rep = compare_all(view, alone, STATE_BUFFERS)
    bad = format_report(rep)
    assert not bad, "%s: batch job differs from the job alone: %s" % (label, bad)
    assert view.counters()["frame_count"] == passes
    alone.close()
    if against_oracle:
        cpu = oracle.OracleRenderer(mesh, params)
        cpu.run(passes)
        rep = compare_all(view, cpu, STATE_BUFFERS)
        bad = format_report(rep)
        assert not bad, "%s: batch job differs from the oracle: %s" % (label, bad)
        cpu.close()


@pytest.mark.parametrize("passes_per_graph,eye_overlap", [(1, 0), (4, 0), (0, 0), (4, 1), (0, 1)])
def test_batch_jobs_equal_alone(passes_per_graph, eye_overlap):
    """eye_overlap 1: the eye stage on a second stream beside the photon stage, in graphs and in
    plain launches."""
    print("passes_per_graph=%d eye_overlap=%d" % (passes_per_graph, eye_overlap), flush=True)
    b = g.Batch(passes_per_graph=passes_per_graph)
    b.set_tuning(eye_overlap=eye_overlap)
    specs = job_specs()
    ids = [b.add(m, p, n) for (_, m, p, n) in specs]
    b.run(5)
    late_mesh = cornell(40, 24)
    late_params = g.default_params(40, image_width=40, image_height=24, hash_width=64, hash_height=16,
                                   max_path_length=4)
    late = b.add(late_mesh, late_params, 30)
    b.run(10)
    # the IBL job (budget 7) finished during the first 15 passes; check it and remove it
    assert b.progress(ids[3]) == (7, 0)
    check_equal_alone(JobView(b, ids[3]), specs[3][1], specs[3][2], 7, specs[3][0])
    b.remove(ids[3])
    b.run(10)
    b.sync()
    assert b.num_jobs == 4
    expect = {ids[0]: 20, ids[1]: 13, ids[2]: 25, late: 20}
    for (label, m, p, _), j in zip(specs[:3], ids[:3]):
        check_equal_alone(JobView(b, j), m, p, expect[j], label)
    check_equal_alone(JobView(b, late), late_mesh, late_params, 20, "late-40x24")
    b.close()


def test_batch_mixes_vertex_and_stored_leaf_boxes():
    """One batched launch with a scene whose leaf boxes are its triangles' vertex boxes (the walk
    takes them from the vertices) and the same scene with every leaf box shrunk (the walk reads
    them from the triangle records): each job equals the same job alone and the oracle."""
    p = g.default_params(64, image_width=48, image_height=40, photon_width=64, photon_height=32, hash_width=128,
                         hash_height=16, max_path_length=6)
    plain = spheres(7, 6, 2, 48, 40)
    shrunk = ShrunkLeafBoxes(spheres(7, 6, 2, 48, 40))
    b = g.Batch(passes_per_graph=0)
    ja = b.add(plain, p, 13)
    jb = b.add(shrunk, p, 13)
    b.run(13)
    b.sync()
    check_equal_alone(JobView(b, ja), plain, p, 13, "vertex-leaf-boxes")
    check_equal_alone(JobView(b, jb), shrunk, p, 13, "stored-leaf-boxes")
    b.close()


@pytest.mark.parametrize("passes_per_graph", [2, 0])
def test_batch_budget_extension_and_reuse_of_freed_memory(passes_per_graph):
    """passes_per_graph 0: plain launches, with each job's next eye stage begun ahead of its eye
    pass; the budget runs out and is extended around an eye pass (L=6)."""
    b = g.Batch(passes_per_graph=passes_per_graph)
    m, p = cornell(32), g.default_params(32, max_path_length=6)
    j0 = b.add(m, p, 3)
    b.run(5)
    assert b.progress(j0) == (3, 0)
    b.set_passes(j0, 4)
    b.run(6)
    check_equal_alone(JobView(b, j0), m, p, 7, "extended")
    held = b.device_bytes
    b.remove(j0)
    b.sync()
    j1 = b.add(m, p, 9)  # same sizes: served from the batch's cached blocks
    assert b.device_bytes == held
    b.run(9)
    check_equal_alone(JobView(b, j1), m, p, 9, "reused")
    b.close()


def test_plain_loop_graph_replay_equals_launches():
    m = spheres(3, 5, 2, 64, 64)
    p = g.default_params(64, max_path_length=7)
    pg = g.default_params(64, max_path_length=7, flags=g.FLAG_GRAPH)
    a = g.Renderer(m, p)
    b = g.Renderer(m, pg)
    a.run(9)
    b.run(4)
    b.run(5)
    rep = compare_all(b, a, STATE_BUFFERS)
    assert not format_report(rep), format_report(rep)


@pytest.mark.parametrize("splits,graph", [((1,) * 9, 0), ((2, 3, 4), 0), ((5, 4), 1), ((1, 1, 7), 1)])
def test_pass_calls_split_anywhere_equal_the_oracle(splits, graph):
    """Each pass's gather runs in the photon-stage launch of the pass after it, and a call's last
    pass is gathered at the end of the call: the same passes split over calls in any way, with
    every buffer read back between the calls, equal the oracle after each call (L = 3: eye passes
    first, inside and last in a call; one-wave photon blocks and four-wave ones, with and without
    graph replay)."""
    m = spheres(3, 5, 2, 50, 37)
    p = g.default_params(64, image_width=50, image_height=37, photon_width=40, photon_height=33, max_path_length=3)
    gpu = g.Renderer(m, p)
    cpu = oracle.OracleRenderer(m, p)
    gpu.set_tuning(eye_overlap=1, passes_per_graph=graph, trace_block=64 if graph else 256)
    for n in splits:
        gpu.run(n)
        cpu.run(n)
        rep = compare_all(gpu, cpu)
        assert not format_report(rep), (n, format_report(rep))
    gpu.close()
    cpu.close()


@pytest.mark.parametrize("shape", [dict(trace_block=64), dict(trace_block=512), dict(tile_wave_width=1),
                                   dict(tile_wave_width=4), dict(tile_wave_width=64),
                                   dict(trace_block=128, tile_wave_width=16, passes_per_graph=1),
                                   dict(walk_stack=0), dict(walk_stack=1), dict(walk_stack=2),
                                   dict(walk_stack=32, trace_block=512), dict(eye_overlap=1),
                                   dict(eye_overlap=1, passes_per_graph=1)])
def test_launch_shapes_do_not_change_results(shape):
    """Odd image sizes so that partial tiles and partial ray blocks occur under every shape; walk
    stacks of 0-2 entries send most walks on over the pre-order records part way."""
    print("shape %s" % shape, flush=True)
    m = spheres(5, 4, 2, 50, 37)
    p = g.default_params(64, image_width=50, image_height=37, photon_width=40, photon_height=33, max_path_length=6)
    a = g.Renderer(m, p)
    b = g.Renderer(m, p)
    b.set_tuning(**shape)
    a.run(13)
    b.run(13)
    rep = compare_all(b, a, STATE_BUFFERS)
    assert not format_report(rep), format_report(rep)


def test_eye_stage_ahead_across_calls_and_tuning_changes():
    """With eye_overlap the plain loop starts each eye pass's eye stage right after the eye pass
    before it: the stage stays pending across ppm_pass calls, readbacks and tuning changes (the
    overlap switched off, graphs switched on, the tiles reshaped) and every state matches one
    stream at every stop."""
    m = spheres(5, 4, 2, 50, 37)
    p = g.default_params(64, image_width=50, image_height=37, photon_width=40, photon_height=33, max_path_length=3)
    a = g.Renderer(m, p)
    b = g.Renderer(m, p)
    b.set_tuning(eye_overlap=1)
    done = 0
    for n, tuning in [(1, None), (1, None), (2, None), (4, dict(passes_per_graph=1)), (1, dict(eye_overlap=0)),
                      (3, dict(eye_overlap=1, tile_wave_width=8)), (2, dict(passes_per_graph=0)), (7, None)]:
        print("passes %d..%d tuning %s" % (done, done + n, tuning), flush=True)
        if tuning:
            b.set_tuning(**tuning)
        b.run(n)
        a.run(n)
        done += n
        rep = compare_all(b, a, STATE_BUFFERS)
        assert not format_report(rep), format_report(rep)


@pytest.mark.parametrize("max_path,passes_per_graph", [(1, 0), (1, 1), (2, 0), (2, 1)])
def test_eye_stage_ahead_at_the_shortest_cadences(max_path, passes_per_graph):
    """L=2: one photon pass between eye passes, so the stage that starts inside it is waited for by
    the very next pass; L=1: every pass is an eye pass and runs its own stage."""
    print("L=%d passes_per_graph=%d" % (max_path, passes_per_graph), flush=True)
    m = spheres(5, 4, 2, 50, 37)
    p = g.default_params(64, image_width=50, image_height=37, photon_width=40, photon_height=33,
                         max_path_length=max_path)
    a = g.Renderer(m, p)
    b = g.Renderer(m, p)
    b.set_tuning(eye_overlap=1, passes_per_graph=passes_per_graph)
    for n in (1, 2, 3, 5):
        a.run(n)
        b.run(n)
        rep = compare_all(b, a, STATE_BUFFERS)
        assert not format_report(rep), format_report(rep)


def test_profiling_times_the_chosen_kernel_kinds_and_keeps_results():
    """ppm_profile_enable: 1 times every kind, PPM_PROFILE_KINDS | bits only the chosen ones; the
    launch counts follow the pass cadence (L=3: eye passes 0, 3 and 6 in 7 passes, and their 3 eye
    stages; the stage of eye pass 9 would start inside pass 7); the gathers of passes 0-5 run in the
    photon-stage launches of passes 1-6, so k_gather runs once, for pass 6, at the end of the call."""
    m = spheres(5, 4, 2, 50, 37)
    p = g.default_params(64, image_width=50, image_height=37, photon_width=40, photon_height=33, max_path_length=3)
    a = g.Renderer(m, p)
    b = g.Renderer(m, p)
    b.set_tuning(eye_overlap=1)
    b.profile(True)
    b.run(7)
    every = b.profile_read()
    assert [every[k][1] for k in g.KERNELS] == [3, 7, 7, 1], every
    assert all(every[k][0] > 0.0 for k in g.KERNELS)
    b.profile(True, kinds=["photon"])
    b.run(5)
    only = b.profile_read()
    assert only["photon"][1] == 5 and only["photon"][0] > 0.0
    assert all(only[k] == (0.0, 0) for k in ("eye", "clear", "gather")), only
    b.profile(False)
    a.run(12)
    rep = compare_all(b, a, STATE_BUFFERS)
    assert not format_report(rep), format_report(rep)
    for bad in (2, 0x100 | 0x10, -1):
        with pytest.raises(g.PPMError):
            b._check(g.library().ppm_profile_enable(b._h, bad))


def test_autotune_keeps_results_and_picks_a_valid_shape():
    m = spheres(9, 6, 2, 64, 48)
    p = g.default_params(64, image_height=48, max_path_length=5)
    r = g.Renderer(m, p)
    t = r.autotune(passes_per_trial=3)
    assert t["trace_block"] in (64, 128, 256, 512) and t["tile_wave_width"] in (1, 2, 4, 8, 16, 32, 64)
    assert t["walk_stack"] in (8, 16, 32) and t["eye_overlap"] in (0, 1)
    assert t["ms_per_pass"] > 0
    done = r.counters()["frame_count"]
    # 13 trials (the start, 3 + 4 + 1 + 2 + 1 alternatives, the winner), each 1 + 3 passes
    assert done == 13 * 4
    a = g.Renderer(m, p)
    a.run(done)
    rep = compare_all(r, a, STATE_BUFFERS)
    assert not format_report(rep), format_report(rep)

    bt = g.Batch()
    specs = job_specs()[:3]
    ids = [bt.add(mm, pp, g.PASSES_UNLIMITED) for (_, mm, pp, _) in specs]
    tb = bt.autotune(passes_per_trial=2)
    assert tb["ms_per_pass"] > 0
    bt.run(3)
    n = bt.progress(ids[0])[0]
    for (label, mm, pp, _), j in zip(specs, ids):
        check_equal_alone(JobView(bt, j), mm, pp, n, label)
    bt.close()


def test_tuning_out_of_range_is_refused():
    m = spheres(5, 4, 2, 32, 32)
    r = g.Renderer(m, g.default_params(32, max_path_length=4))
    for bad in (dict(trace_block=96), dict(trace_block=1024), dict(tile_wave_width=3), dict(passes_per_graph=2),
                dict(walk_stack=-1), dict(walk_stack=33), dict(eye_overlap=2), dict(ms_per_pass=1.0),
                dict(no_such_field=1)):
        with pytest.raises(g.PPMError):
            r.set_tuning(**bad)
    assert r.tuning()["trace_block"] == 64
