"""Pass-level pins of the oracle.

- hash.fs / scatter / correction (glslppm.cpp:401-470): per slot, the count equals the number of
  valid photons hashed there and the representative equals the brute-force minimum photon index
  with the depth-clip exclusion of the last photon row's right half; the slot of every valid
  photon equals hash() of its cell.
- progressive.fs gather (progressive.fs:65-127): a brute-force O(Q P) search over all photons
  with the same slot rules reproduces the photon count M and flux of every hit point bit for bit.
- progressive update (progressive.fs:129-137): g = 1 when N + M = 0, and the closed-form
  recurrence of R and N under a constant M.
- re.fs: closed form of the radiance, tone map and sRGB curve.
- energy: the flux of a freshly emitted photon is Le * A * pi (photontrace.fs:409-411).
- the oracle's result does not depend on the number of OpenMP threads.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import glslppm_amd as g
import oracle

f32 = np.float32


def cornell(res, **kw):
    m = g.Mesh.generate(g.SCENE_CORNELL).set_camera((0, 0, 1.667), (0, 0, 0), res, res, 45.0).build()
    return m, g.default_params(res, **kw)


def test_slot_counts_and_representatives_brute_force():
    m, p = cornell(32, max_path_length=6)
    o = oracle.OracleRenderer(m, p)
    for npass in (1, 4, 3):
        o.run(npass)
        flux = o.readback("photon_flux")
        slot = o.readback("photon_slot")
        cnt = o.readback("slot_count")
        rep = o.readback("slot_rep")
        P, H, Pw = o.p, o.h, p.photon_width
        valid = ~(flux[:, 0] < 0)
        assert np.all(slot[~valid] == -1)
        assert cnt.sum() == np.count_nonzero(valid & (slot >= 0))
        bc = np.bincount(slot[slot >= 0], minlength=H)
        assert np.array_equal(bc, cnt)
        limit = P - Pw // 2
        for s in range(H):
            cand = np.where((slot == s) & (np.arange(P) < limit))[0]
            assert rep[s] == (cand.min() if cand.size else -1)
        # the slot of each valid photon is hash() of its cell
        gr = o.grid()
        pos = o.readback("photon_position")[:, :3]
        for i in np.where(valid)[0][:300]:
            cell = np.floor((pos[i] - np.array(gr["bbmin"], f32)).astype(f32) * f32(gr["grid_scale"]))
            h = oracle.cell_hash(cell, gr["grid_scale"], gr["hash_scale"], H)
            assert slot[i] == (int(h) if h < H else -1)


def _cell_slot(cell, gr, H, Hw, Hh):
    h = oracle.cell_hash(cell, gr["grid_scale"], gr["hash_scale"], H)
    # the hash texture is sampled with GL_CLAMP + GL_NEAREST: a row past the last clamps
    return int(h) if h < H else (Hh - 1) * Hw


def test_gather_brute_force_over_all_photons():
    m, p = cornell(16, max_path_length=5)
    p.photon_width = p.photon_height = 24
    p.hash_width, p.hash_height = 16, 8
    o = oracle.OracleRenderer(m, p)
    o.run(2)  # pass 1 is not an eye pass, so the hit points stay
    fr0 = o.readback("flux_radius")
    en0 = o.readback("emission_count")
    o.run(1)
    fr1, en1 = o.readback("flux_radius"), o.readback("emission_count")
    qpos, qnrm = o.readback("hit_position")[:, :3], o.readback("hit_normal")[:, :3]
    qrefl, qtri = o.readback("hit_reflectance")[:, :3], o.readback("hit_triangle")[:, 0]
    ppos, pflux = o.readback("photon_position")[:, :3], o.readback("photon_flux")[:, :3]
    pdir, pslot = o.readback("photon_direction")[:, :3], o.readback("photon_slot")
    gr = o.grid()
    P, H, Q = o.p, o.h, o.q
    valid = ~(pflux[:, 0] < 0)
    limit = P - p.photon_width // 2
    bb = np.array(gr["bbmin"], f32)
    gs = f32(gr["grid_scale"])
    checked = 0
    for q in range(Q):
        if qtri[q] < 0:
            assert np.array_equal(fr1[q], fr0[q]) and en1[q, 3] == en0[q, 3]
            continue
        ii = np.floor(np.abs(qpos[q] - bb) * gs - f32(0.5)).astype(f32)
        R = fr0[q, 3]
        flux = np.zeros(3, f32)
        mcount = f32(0)
        for iz in (0, 1):
            for iy in (0, 1):
                for ix in (0, 1):
                    cell = (ii + np.array([ix, iy, iz], f32)).astype(f32)
                    s = _cell_slot(cell, gr, H, p.hash_width, p.hash_height)
                    members = np.where(valid & (pslot == s))[0]  # O(P) scan per cell
                    if members.size == 0:
                        continue
                    reps = members[members < limit]
                    r = reps.min() if reps.size else 0  # a cleared slot reads photon texel (0, 0)
                    lo = (cell / gs + bb).astype(f32)
                    hi = ((cell + f32(1)) / gs + bb).astype(f32)
                    x = ppos[r]
                    if np.all(x > lo) and np.all(x < hi):
                        diff = (x - qpos[q]).astype(f32)
                        dist = np.sqrt(np.float32(diff[0] * diff[0] + diff[1] * diff[1]) + diff[2] * diff[2])
                        facing = f32(f32(qnrm[q, 0] * -pdir[r, 0]) + f32(qnrm[q, 1] * -pdir[r, 1])) + f32(qnrm[q, 2] * -pdir[r, 2])
                        if dist < R and facing > 0:
                            c = f32(members.size)
                            flux = (flux + (pflux[r] * c).astype(f32)).astype(f32)
                            mcount = f32(mcount + c)
        assert en1[q, 3] == f32(en0[q, 3] + f32(mcount * f32(0.7))), q
        tau, r1, n1 = oracle.update(fr0[q, :3], fr0[q, 3], en0[q, 3], flux, mcount, qrefl[q], 0.7)
        assert np.array_equal(fr1[q], np.array([tau[0], tau[1], tau[2], r1], f32)), q
        checked += mcount > 0
    assert checked > 10


def test_update_without_photons_keeps_radius():
    # N + M = 0 gives 0/0 in g; the NaN-rejecting min turns it into 1
    tau, r, n = oracle.update([0.5, 0.25, 0.125], 0.02, 0.0, [0.0, 0.0, 0.0], 0.0, [0.7, 0.7, 0.7], 0.7)
    assert r == f32(0.02) and n == 0.0 and list(tau) == [f32(0.5), f32(0.25), f32(0.125)]


@pytest.mark.parametrize("c", [1.0, 3.0, 17.0])
def test_update_closed_form_recurrence(c):
    alpha = 0.7
    tau, r, n = np.zeros(3, f32), f32(0.05), f32(0.0)
    phi = np.array([2.0, 1.0, 0.5], f32)
    refl = np.array([0.9, 0.6, 0.3], f32)
    R, N, T = 0.05, 0.0, np.zeros(3)
    for k in range(40):
        tau, r, n = oracle.update(tau, r, n, phi, c, refl, alpha)
        gk = (N + alpha * c) / (N + c)
        R = R * math.sqrt(gk)
        N = N + alpha * c
        T = (T + phi.astype(np.float64) * refl * 0.31830995240629591621063460818591) * gk
        # closed form of the product: N_k = alpha c k, g_i = alpha (i + 1) / (alpha i + 1)
        assert abs(N - alpha * c * (k + 1)) < 1e-9 * (k + 1) * c
        assert abs(float(n) - N) <= 4e-7 * N * (k + 1)
        assert abs(float(r) - R) <= 4e-7 * R * (k + 1)
        assert np.all(np.abs(tau - T) <= 4e-7 * np.abs(T) * (k + 1) + 1e-12)
    prod = 0.05 * math.sqrt(math.prod(alpha * (i + 1) / (alpha * i + 1) for i in range(40)))
    assert abs(float(r) - prod) <= 2e-5 * prod


def test_radiance_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(500):
        fr = np.array([*rng.uniform(0, 5e-4, 3), rng.uniform(0.002, 0.03)], f32)
        em = np.array([*rng.uniform(0, 2e-5, 3), 1.0], f32)
        nph = f32(rng.uniform(1e5, 1e8))
        lin, mp = oracle.radiance(fr, em, nph)
        den = float(fr[3]) ** 2 * 3.141592 * float(nph)
        L = np.array([float(fr[k]) / den + float(em[k]) for k in range(3)])
        assert np.allclose(lin[:3], L, rtol=1e-6, atol=1e-12)
        for k in range(3):
            t = 1.0 - math.exp(-float(lin[k]) * 60000.0)
            s = 12.92 * t if t < 0.0031308 else 1.055 * t ** 0.416666666667 - 0.055
            assert abs(float(mp[k]) - s) <= 3e-6
        assert mp[3] == f32(1.0)


def test_emitted_photon_energy():
    m, p = cornell(32, max_path_length=8)
    a = m.arrays()
    o = oracle.OracleRenderer(m, p)
    o.run(1)
    c = o.counters()
    assert c["photons_emitted"] == o.p and c["num_photons_f"] == f32(o.p)
    flux = o.readback("photon_flux")[:, :3]
    st = o.readback("photon_state")
    le = f32(a["light_area"] * f32(3.141592))
    hit = st[:, 0] >= 0
    # every first segment of the Cornell room ends on a flat triangle: shading and geometric
    # normals agree to rounding, so the g-factor is 1 within a few ulp
    assert np.allclose(np.abs(flux[hit]), le, rtol=2e-6)
    assert np.all(st[:, 2] == 1)


_DET = r"""
import sys, hashlib
sys.path.insert(0, sys.argv[1])
import glslppm_amd as g, oracle
m = g.Mesh.generate(g.SCENE_SPHERES, seed=4, count=3, detail=2).set_camera((0,0,1.667),(0,0,0),32,32,45.0).build()
o = oracle.OracleRenderer(m, g.default_params(32, max_path_length=6))
o.run(13)
h = hashlib.sha256()
for b in ("flux_radius", "emission_count", "photon_flux", "photon_rng", "slot_rep", "slot_count", "hit_triangle"):
    h.update(o.readback(b).tobytes())
print(oracle.num_threads(), h.hexdigest())
"""


def test_oracle_independent_of_thread_count():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for n in ("1", "4"):
        env = dict(os.environ, OMP_NUM_THREADS=n)
        r = subprocess.run([sys.executable, "-c", _DET, root], env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout.split()[-2:])
    assert outs[0][0] == "1" and outs[1][0] == "4"
    assert outs[0][1] == outs[1][1]
