"""The benchmark configurations at their full sizes on the GPU, against the
oracle (the reference's C restated, pinned bit for bit against the
compiled reference) on the same seeded inputs.

* one grain x 20 000 g-vectors (12 000 lattice + noise, 8 000 random):
  score_and_refine n exact, mean drlv2 within 1e-12, UBI within 1e-10;
* 5000 grains, lognormal lists from a 2e6-peak table, 10 iterations:
  every iteration's n exact, every iteration's mean within 1e-12, UBIs
  within 1e-10;
* the adjoint at 5000 grains x 2e6 peaks: the loss against the oracle's
  objective, the geometry gradient against central differences;
* 20 000 grains x 8e6 peaks (per-peak dty, 15% spurious): labels and drlv2
  of a seeded sample of peaks after all 20 000 grains, in the parity mode
  bit for bit (fast mode: labels equal), and 20 captured refine
  iterations checked on a sample of grains.
"""
import os

import numpy as np
import pytest
import torch

from imaged11_amd import adjoint, batched, drivers, geometry, refine, simulate
from imaged11_amd import cImageD11 as gpu

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)


def T(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.int64)


def test_c1_one_grain_20000_gvectors(dev, orc):
    print("C1: 1 grain x 20000 g-vectors, tol 0.05", flush=True)
    ubi, gv, _ = simulate.lattice_gvectors(seed=1, n_lattice=12000, n_random=8000)
    rng = np.random.default_rng(10)
    start = ubi @ (np.eye(3) + 1e-3 * rng.normal(size=(3, 3)))
    u_o = np.ascontiguousarray(start.copy())
    for call in range(3):  # a chain of calls, as refinegrains.refine makes
        n_o, m_o = orc.score_and_refine(u_o, gv, 0.05)
        u_h = np.ascontiguousarray(start.copy()) if call == 0 else u_h
        n_h, m_h = gpu.score_and_refine(u_h, gv, 0.05)
        assert n_h == n_o
        assert abs(m_h - m_o) <= 1e-12 * abs(m_o)
        assert np.abs(u_h - u_o).max() <= 1e-10 * np.abs(u_o).max()
    # device tensors, one call
    u_d = T(start, dev)
    n_d, m_d = gpu.score_and_refine(u_d, T(gv, dev), 0.05)
    u_1 = np.ascontiguousarray(start.copy())
    n_1, m_1 = orc.score_and_refine(u_1, gv, 0.05)
    assert n_d == n_1 and abs(m_d - m_1) <= 1e-12 * abs(m_1)
    assert np.abs(u_d.cpu().numpy() - u_1).max() <= 1e-10 * np.abs(u_1).max()
    assert gpu.score(ubi, gv, 0.05) == orc.score(ubi, gv, 0.05)


@pytest.fixture(scope="module")
def c3(orc):
    print("C3/C4 workload: 5000 grains, 2e6 peaks, lognormal lists", flush=True)
    K = 5000
    wl = simulate.scoring_workload(K=K, N=2_000_000, seed=3)
    idx, off = simulate.refine_lists(wl["peaks"]["grain"], K, seed=3)
    return wl, idx, off


def test_c3_batched_refine_5000_grains(dev, orc, c3):
    wl, idx, off = c3
    K = len(off) - 1
    pk, pars = wl["peaks"], wl["pars"]
    gv = batched.sf2gv(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev), pars,
                       parity=True)
    rng = np.random.default_rng(4)
    start = wl["ubi"] @ (np.eye(3) + rng.normal(scale=1e-3, size=(K, 3, 3)))
    n_iter = 10
    print("C3: 5000 grains x 10 iterations, %d list entries" % int(off[-1]), flush=True)
    gvh = gv.cpu().numpy()
    u_o, n_o, m_o = orc.refine_lists(start, gvh, idx, off, 0.05, n_iter, nthreads=NTHREADS)
    idx_d, off_d = T(idx, dev, torch.int32), T(off, dev, torch.int64)
    u_g, n_g, m_g = batched.refine_lists(T(start, dev), gv, idx_d, off_d, 0.05, n_iter)
    assert np.array_equal(n_g.cpu().numpy(), n_o)
    rel = np.abs(m_g.cpu().numpy() - m_o) / np.maximum(np.abs(m_o), 1e-300)
    assert rel.max() <= 1e-12
    err = np.abs(u_g.cpu().numpy() - u_o).max(axis=(1, 2)) / np.abs(u_o).max(axis=(1, 2))
    assert err.max() <= 1e-10
    # the captured hipGraph of the same 10 iterations, replayed
    work = T(start, dev)
    g = refine.RefineGraph(work, gv, idx_d, off_d, 0.05, n_iter)
    g.replay()
    torch.cuda.synchronize()
    assert np.array_equal(g.n.cpu().numpy(), n_o)
    err = np.abs(work.cpu().numpy() - u_o).max(axis=(1, 2)) / np.abs(u_o).max(axis=(1, 2))
    assert err.max() <= 1e-10
    g.close()


def test_c4_adjoint_5000_grains(dev, orc, c3):
    wl, idx, off = c3
    K = len(off) - 1
    pk, pars = wl["peaks"], wl["pars"]
    print("C4: adjoint, 5000 grains x 2e6 peaks", flush=True)
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                pars["omegasign"], T(idx, dev, torch.int32),
                                T(off, dev, torch.int64), parity=True)
    params = {k: torch.tensor(float(pars[k]), dtype=torch.float64, device=dev)
              for k in adjoint.FREE[:8]}
    rng = np.random.default_rng(4)
    ubis = T(wl["ubi"] @ (np.eye(3) + rng.normal(scale=1e-3, size=(K, 3, 3))), dev)
    trans = torch.zeros((K, 3), dtype=torch.float64, device=dev)
    geo = adjoint.pack_geometry(params, pars, parity=True)
    prob.assign_hkl(geo, ubis, trans)
    loss, gu, gt, gg, sums = prob.evaluate(geo, ubis, trans, want_sums=True)
    g = geo.detach().cpu().numpy()
    so, co = orc.sincos_omega(pk["omega"], pars["omegasign"])
    want = orc.joint_loss(pk["sc"], pk["fc"], so, co, g[:4], g[4:13], g[13:16], g[16], g[17:],
                          ubis.cpu().numpy(), trans.cpu().numpy(), idx, off,
                          prob.hkl.cpu().numpy())
    assert abs(float(loss) - want) <= 1e-10 * abs(want)
    # central differences for the 8 geometry parameters, h = 1e-6 relative
    # with floors: 1e-6 rad for the tilts, 1e-4 degree for the wedge.  The
    # loss (~2e5 over 2e6 terms) carries rounding of ~1e-9 relative, so a
    # step of 1e-9 rad would leave the difference of the two losses within
    # a few ulps of the loss; these steps keep it 1e5 times above that.
    p = {k: v.clone().requires_grad_(True) for k, v in params.items()}
    L = adjoint.joint_loss(prob, ubis, trans, p, pars)
    L.backward()
    floors = {"tilt_x": 1.0, "tilt_y": 1.0, "tilt_z": 1.0, "wedge": 100.0}

    def at(q):
        return float(prob.evaluate(adjoint.pack_geometry(q, pars, parity=True), ubis,
                                   trans)[0])
    for k in adjoint.FREE[:8]:
        v = float(params[k])
        h = 1e-6 * max(abs(v), floors.get(k, 0.0))
        up = {kk: vv.clone() for kk, vv in params.items()}
        dn = {kk: vv.clone() for kk, vv in params.items()}
        up[k] = up[k] + h
        dn[k] = dn[k] - h
        fd = (at(up) - at(dn)) / (2 * h)
        gk = float(p[k].grad)
        assert abs(gk - fd) <= 1e-5 * max(abs(fd), abs(gk)) + 1e-9 * abs(float(L)), (k, gk, fd)


@pytest.fixture(scope="module")
def c5(orc):
    print("C5 workload: 20000 grains x 8e6 peaks", flush=True)
    return simulate.scan_workload(K=20000, N=8_000_000, seed=5, device="cuda")


@pytest.mark.parametrize("parity", [True, False], ids=["parity", "fast"])
def test_c5_scan_20000_grains(dev, orc, c5, parity):
    wl = c5
    pk, pars = wl["peaks"], wl["pars"]
    K, N = len(wl["ubi"]), len(pk["sc"])
    print("C5: assign 20000 x 8e6, parity=%s" % parity, flush=True)
    scan = drivers.Scan(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev), pars,
                        dty=T(pk["dty"], dev), parity=parity)
    ubis, trans = T(wl["ubi"], dev), T(wl["trans"], dev)
    labels, drlv2, counts = drivers.assignlabels(scan, ubis, trans, 0.05)
    # a peak's label and drlv2 depend on that peak alone: a seeded sample of
    # peaks run through every grain by the oracle must end in the same state
    rng = np.random.default_rng(55)
    sel = np.sort(rng.choice(N, size=50_000, replace=False))
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    xl = np.zeros((N, 3))
    orc.compute_xlylzl(pk["sc"], pk["fc"], p, r, dist, xl)
    xl[:, 1] -= pk["dty"]
    xs = np.ascontiguousarray(xl[sel])
    so, co = orc.sincos_omega(pk["omega"][sel], sign)
    mat = orc.wedge_chi_matrix(wedge, chi)
    d, lab = np.ones(len(sel)), np.full(len(sel), -1, np.int32)
    orc.gv_assign_many(xs, so, co, mat, wvln, wl["ubi"], wl["trans"], np.arange(K), 0.05, d,
                       lab, nthreads=NTHREADS)
    assert np.array_equal(labels.cpu().numpy()[sel], lab)
    got_d = drlv2.cpu().numpy()[sel]
    if parity:
        assert np.array_equal(bits(got_d), bits(d))
    else:
        np.testing.assert_allclose(got_d, d, rtol=1e-9, atol=1e-18)
    if not parity:
        return
    # 20 captured refine iterations on every grain's assigned peaks, checked
    # against the oracle's score_and_refine chain on a sample of grains
    print("C5: 20 captured refine iterations", flush=True)
    gvl, idx, off = drivers.grain_gvectors(scan, trans, labels, K)
    n_iter = 20
    work = ubis.clone()
    g = refine.RefineGraph(work, gvl, None, off, 0.05, n_iter)
    g.replay()
    torch.cuda.synchronize()
    n_g = g.n.cpu().numpy()
    u_g = work.cpu().numpy()
    gvl_h, off_h = gvl.cpu().numpy(), off.cpu().numpy()
    for k in rng.choice(K, size=64, replace=False):
        u = wl["ubi"][k].copy()
        rows = np.ascontiguousarray(gvl_h[off_h[k]:off_h[k + 1]])
        for it in range(n_iter):
            nn, _ = orc.score_and_refine(u, rows, 0.05)
            assert int(n_g[it, k]) == nn
        assert np.abs(u_g[k] - u).max() <= 1e-10 * np.abs(u).max()
    g.close()
