"""Batched refinement on the GPU: the fused persistent launch, the
hipGraph of per-iteration launches and the autotuner, against the oracle's
per-grain score_and_refine loop (the calls refinegrains.refine makes)."""
import numpy as np
import pytest
import torch

from imaged11_amd import batched, refine, simulate

pytestmark = pytest.mark.gpu

UBI_TOL = 1e-10
MEAN_RTOL = 1e-12
CHAIN_MEAN_RTOL = 1e-9


def T(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


@pytest.fixture(scope="module")
def problem(orc):
    from imaged11_amd import geometry
    wl = simulate.scoring_workload(K=300, N=120000, seed=31)
    pk, pars = wl["peaks"], wl["pars"]
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    xl = np.zeros((len(pk["sc"]), 3))
    orc.compute_xlylzl(pk["sc"], pk["fc"], p, r, dist, xl)
    gv = np.zeros_like(xl)
    orc.compute_gv(xl, pk["omega"], sign, wvln, wedge, chi, np.zeros(3), gv)
    K = 300
    idx, off = simulate.refine_lists(pk["grain"], K, seed=32, median=300, lo=20, hi=2000)
    rng = np.random.default_rng(33)
    start = wl["ubi"] @ (np.eye(3) + rng.normal(scale=2e-3, size=(K, 3, 3)))
    return dict(gv=gv, idx=idx, off=off, start=start)


def _dev(problem, dev):
    return (T(problem["gv"], dev), T(problem["idx"], dev, torch.int32),
            T(problem["off"], dev, torch.int64))


def test_graph_equals_fused_and_oracle(dev, orc, problem):
    gv, idx, off = _dev(problem, dev)
    n_iter = 10
    u_o, n_o, m_o = orc.refine_lists(problem["start"], problem["gv"], problem["idx"],
                                     problem["off"], 0.05, n_iter)
    for lanes in (16, 64):
        uf, nf, mf = batched.refine_lists(T(problem["start"], dev), gv, idx, off, 0.05, n_iter,
                                          lanes_per_grain=lanes)
        work = T(problem["start"], dev)
        g = refine.RefineGraph(work, gv, idx, off, 0.05, n_iter, lanes_per_grain=lanes)
        for rep in range(3):
            # eager work and a sync between replays; every replay checked
            work.copy_(T(problem["start"], dev))
            junk = torch.randn(1000, device=dev).sum()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(g.n, nf) and torch.equal(g.mean, mf) and torch.equal(work, uf)
            assert float(junk) == float(junk)
        g.close()
        assert np.array_equal(nf.cpu().numpy(), n_o)
        m = mf.cpu().numpy()
        assert np.all(np.abs(m[0] - m_o[0]) <= MEAN_RTOL * np.abs(m_o[0]))
        assert np.all(np.abs(m - m_o) <= CHAIN_MEAN_RTOL * np.abs(m_o))
        np.testing.assert_allclose(uf.cpu().numpy(), u_o, rtol=0, atol=UBI_TOL)


def test_autotune_and_refine(dev, orc, problem, tmp_path, monkeypatch):
    monkeypatch.setenv(refine.CACHE_ENV, str(tmp_path / "tune.json"))
    gv, idx, off = _dev(problem, dev)
    start = T(problem["start"], dev)
    best = refine.autotune(start, gv, idx, off, 0.05, 4, reps=2)
    assert best["mode"] in refine.MODES and best["lanes_per_grain"] in refine.LANES
    assert refine.autotune(start, gv, idx, off, 0.05, 4)["key"] == best["key"]  # cached
    u, n, m = refine.refine(start, gv, idx, off, 0.05, 4, config=best)
    u_o, n_o, m_o = orc.refine_lists(problem["start"], problem["gv"], problem["idx"],
                                     problem["off"], 0.05, 4)
    assert np.array_equal(n.cpu().numpy(), n_o)
    np.testing.assert_allclose(u.cpu().numpy(), u_o, rtol=0, atol=UBI_TOL)
    assert torch.equal(start, T(problem["start"], dev))  # input untouched


def test_refine_lists_bad_entries_ignored(dev, orc):
    """Out-of-range grain numbers, offsets and peak indices are skipped."""
    rng = np.random.default_rng(40)
    ubi, gv, _ = simulate.lattice_gvectors(seed=41, n_lattice=600, n_random=400)
    K = 4
    ubis = np.repeat(ubi[None], K, axis=0)
    idx = rng.integers(0, len(gv), 400).astype(np.int32)
    idx[[3, 50, 77]] = [-5, len(gv), 2 ** 31 - 1]
    off = np.array([0, 100, 200, 300, 400], np.int64)
    order = np.array([3, 2, 1, 0], np.int32)
    good = (idx >= 0) & (idx < len(gv))
    u, n, m = batched.refine_lists(T(ubis, dev), T(gv, dev), T(idx, dev, torch.int32),
                                   T(off, dev, torch.int64), 0.05, 1,
                                   order=T(order, dev, torch.int32))
    for g in range(K):
        sel = idx[off[g]:off[g + 1]][good[off[g]:off[g + 1]]]
        uu = ubi.copy()
        nn, mm = orc.score_and_refine(uu, gv[sel], 0.05)
        assert int(n[0, g]) == nn
        np.testing.assert_allclose(u[g].cpu().numpy(), uu, rtol=0, atol=UBI_TOL)
    # a bad grain number in order and an offset past the end: no crash,
    # the affected grain is simply not refined
    order_bad = np.array([3, 2, 99, 0], np.int32)
    off_bad = np.array([0, 100, 200, 300, 10 ** 9], np.int64)
    u2, n2, _ = batched.refine_lists(T(ubis, dev), T(gv, dev), T(idx, dev, torch.int32),
                                     T(off_bad, dev, torch.int64), 0.05, 1,
                                     order=T(order_bad, dev, torch.int32))
    torch.cuda.synchronize()
    assert torch.equal(u2[1], T(ubi, dev))
