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


@pytest.mark.parametrize("lanes", [16, 64])
def test_gathered_rows_bit_identical(dev, problem, lanes):
    """The refinement on rows packed once (gather_rows, idx None) sees the
    same values in the same order as the indexed one: bit-identical
    results, bad indices included (a NaN row is never kept)."""
    gv, idx, off = _dev(problem, dev)
    idx = idx.clone()
    idx[[5, 700, 4000]] = torch.tensor([-1, gv.shape[0], 2 ** 31 - 1], dtype=torch.int32,
                                       device=dev)
    start = T(problem["start"], dev)
    ui, ni, mi = batched.refine_lists(start, gv, idx, off, 0.05, 3, lanes_per_grain=lanes,
                                      gather=False)
    ug, ng_, mg = batched.refine_lists(start, gv, idx, off, 0.05, 3, lanes_per_grain=lanes,
                                       gather=True)
    rows = batched.gather_rows(gv, idx)
    ur, nr, mr = batched.refine_lists(start, rows, None, off, 0.05, 3, lanes_per_grain=lanes)
    torch.cuda.synchronize()
    assert torch.isnan(rows[[5, 700, 4000]]).all()
    assert torch.equal(rows[6], gv[idx[6].long()])
    for u, n, m in ((ug, ng_, mg), (ur, nr, mr)):
        assert torch.equal(u, ui) and torch.equal(n, ni) and torch.equal(m, mi)


def test_graph_replay_gathers_current_rows(dev, problem):
    """A replay packs the rows again, so it sees gv as it is at the replay."""
    gv, idx, off = _dev(problem, dev)
    gv = gv.clone()
    start = T(problem["start"], dev)
    work = start.clone()
    g = refine.RefineGraph(work, gv, idx, off, 0.05, 3)
    assert g.rows is not None
    g.replay()
    torch.cuda.synchronize()
    u1, _, _ = batched.refine_lists(start, gv, idx, off, 0.05, 3)
    assert torch.equal(work, u1)
    gv.mul_(1.0 + 1e-4)
    work.copy_(start)
    g.replay()
    torch.cuda.synchronize()
    u2, n2, m2 = batched.refine_lists(start, gv, idx, off, 0.05, 3)
    assert torch.equal(work, u2) and torch.equal(g.n, n2) and torch.equal(g.mean, m2)
    assert not torch.equal(u1, u2)
    g.close()


def test_rows_must_cover_the_lists(dev):
    """Without an index the rows must cover every list entry."""
    import ctypes
    from imaged11_amd import _lib
    from imaged11_amd._lib import dptr, stream_ptr
    lib = _lib.load()
    ubis = torch.eye(3, dtype=torch.float64, device=dev)[None].repeat(2, 1, 1)
    rows = torch.zeros((10, 3), dtype=torch.float64, device=dev)
    off = torch.tensor([0, 6, 11], dtype=torch.int64, device=dev)
    order = torch.tensor([0, 1], dtype=torch.int32, device=dev)
    n = torch.zeros((1, 2), dtype=torch.int32, device=dev)
    m = torch.zeros((1, 2), dtype=torch.float64, device=dev)
    q = torch.zeros(1, dtype=torch.int32, device=dev)
    rc = lib.gs_refine_lists(dptr(ubis), 2, dptr(rows), 10, None, ctypes.c_longlong(11), dptr(off),
                             dptr(order), 0.05, 1, dptr(n), dptr(m), 0, 0, dptr(q), stream_ptr())
    assert rc != 0
    assert b"cover" in lib.gs_last_error()
    lib.gs_clear_error()


def test_empty_lists_rows_mode(dev):
    """No list entries at all, gathered (n_iter >= 2) or given as rows: the
    UBIs are kept, n = 0 and mean = 0 for every call, as score_and_refine
    on no peaks returns."""
    ubis = torch.eye(3, dtype=torch.float64, device=dev)[None].repeat(5, 1, 1) * 4.0
    gv = torch.zeros((10, 3), dtype=torch.float64, device=dev)
    idx = torch.zeros(0, dtype=torch.int32, device=dev)
    off = torch.zeros(6, dtype=torch.int64, device=dev)
    for g_idx, g_rows in ((idx, gv), (None, gv[:0])):
        u, n, m = batched.refine_lists(ubis, g_rows, g_idx, off, 0.05, 3)
        torch.cuda.synchronize()
        assert torch.equal(u, ubis)
        assert int(n.abs().sum()) == 0 and float(m.abs().sum()) == 0.0
    assert batched.gather_rows(gv, idx).shape == (0, 3)


@pytest.mark.parametrize("lanes", [16, 32])
def test_long_lists_get_a_wave(dev, orc, problem, lanes):
    """long_min: the grains with the longest lists refined by a whole wave,
    the rest at `lanes` lanes per grain; against the oracle's per-grain
    loop like every other launch shape, graph and fused alike, and the
    count of long grains is that of the lengths."""
    gv, idx, off = _dev(problem, dev)
    lengths = np.diff(problem["off"])
    long_min = int(np.quantile(lengths, 0.8))
    n_iter = 4
    u_o, n_o, m_o = orc.refine_lists(problem["start"], problem["gv"], problem["idx"],
                                     problem["off"], 0.05, n_iter)
    uf, nf, mf = batched.refine_lists(T(problem["start"], dev), gv, idx, off, 0.05, n_iter,
                                      lanes_per_grain=lanes, long_min=long_min)
    work = T(problem["start"], dev)
    g = refine.RefineGraph(work, gv, idx, off, 0.05, n_iter, lanes_per_grain=lanes,
                           long_min=long_min)
    g.replay()
    torch.cuda.synchronize()
    assert int(g.queue[n_iter]) == int((lengths >= long_min).sum()) > 0
    assert torch.equal(g.n, nf) and torch.equal(g.mean, mf) and torch.equal(work, uf)
    g.close()
    assert np.array_equal(nf.cpu().numpy(), n_o)
    m = mf.cpu().numpy()
    assert np.all(np.abs(m[0] - m_o[0]) <= MEAN_RTOL * np.abs(m_o[0]))
    assert np.all(np.abs(m - m_o) <= CHAIN_MEAN_RTOL * np.abs(m_o))
    np.testing.assert_allclose(uf.cpu().numpy(), u_o, rtol=0, atol=UBI_TOL)


@pytest.mark.parametrize("lanes,long_q,mid_q", [(8, 0.9, 0.5), (16, 0.8, 0.4), (16, None, 0.6),
                                                (32, 0.9, 0.5)])
def test_three_lane_tiers(dev, orc, problem, lanes, long_q, mid_q):
    """Lanes in proportion to list length: lists of at least long_min
    entries on 64 lanes, at least mid_min on 32, the rest on `lanes`
    (mid_min only, or mid tier wider than `lanes`, included).  Against the
    oracle's per-grain loop, graph equal to fused bit for bit, and the two
    tier counts those of the lengths."""
    gv, idx, off = _dev(problem, dev)
    lengths = np.diff(problem["off"])
    long_min = int(np.quantile(lengths, long_q)) if long_q else 0
    mid_min = int(np.quantile(lengths, mid_q))
    n_iter = 4
    u_o, n_o, m_o = orc.refine_lists(problem["start"], problem["gv"], problem["idx"],
                                     problem["off"], 0.05, n_iter)
    uf, nf, mf = batched.refine_lists(T(problem["start"], dev), gv, idx, off, 0.05, n_iter,
                                      lanes_per_grain=lanes, long_min=long_min, mid_min=mid_min)
    work = T(problem["start"], dev)
    g = refine.RefineGraph(work, gv, idx, off, 0.05, n_iter, lanes_per_grain=lanes,
                           long_min=long_min, mid_min=mid_min)
    g.replay()
    torch.cuda.synchronize()
    n64 = int((lengths >= long_min).sum()) if long_min else 0
    n32 = int(((lengths >= mid_min) | ((lengths >= long_min) if long_min else False)).sum())
    assert int(g.queue[n_iter]) == n64 and int(g.queue[n_iter + 1]) == n32 > n64
    assert torch.equal(g.n, nf) and torch.equal(g.mean, mf) and torch.equal(work, uf)
    g.close()
    assert np.array_equal(nf.cpu().numpy(), n_o)
    m = mf.cpu().numpy()
    assert np.all(np.abs(m[0] - m_o[0]) <= MEAN_RTOL * np.abs(m_o[0]))
    assert np.all(np.abs(m - m_o) <= CHAIN_MEAN_RTOL * np.abs(m_o))
    np.testing.assert_allclose(uf.cpu().numpy(), u_o, rtol=0, atol=UBI_TOL)


@pytest.fixture(scope="module")
def dyadic():
    """Data whose score_and_refine sums are exact in any order: g-vectors
    on a 1/16 grid (|g| <= 2), integer UBIs, tol 0.2, so h, its rounding,
    drlv2 (multiples of 1/256) and every term of n, sum drlv2, R and H are
    small dyadic numbers whose sums over a list never round.  The totals a
    launch forms are then the reference's whatever the lanes and their
    order, and one iteration's UBIs must equal the oracle's bit for bit:
    that pins the 3x3 solve (two inverses, each with nine IEEE divisions),
    including the one that shares the divisions out over a DPP row."""
    rng = np.random.default_rng(41)
    K = 240
    ubis = rng.integers(-2, 3, size=(K, 3, 3)).astype(np.float64) + 5.0 * np.eye(3)
    lengths = np.clip(np.exp(rng.normal(np.log(300), 0.8, size=K)), 20, 2000).astype(np.int64)
    blocks = []
    for k in range(K):
        # 70% near the grain's lattice (UBI^-1 hkl put on the 1/16 grid),
        # the rest anywhere on the grid
        n_lat = int(0.7 * lengths[k])
        hkl = rng.integers(-4, 5, size=(n_lat, 3))
        lat = np.rint(16.0 * hkl @ np.linalg.inv(ubis[k]).T) / 16.0
        blocks += [lat, rng.integers(-16, 17, size=(lengths[k] - n_lat, 3)) / 16.0]
    gv = np.concatenate(blocks)
    perm = rng.permutation(len(gv))  # lists scattered over the table
    gv_table = np.empty_like(gv)
    gv_table[perm] = gv
    idx = perm.astype(np.int32)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    gv = gv_table
    return dict(gv=gv, ubis=ubis, idx=idx, off=off, lengths=lengths)


@pytest.mark.parametrize("lanes,long_q,mid_q", [(8, None, None), (16, None, None), (32, None, None),
                                                (64, None, None), (16, 0.9, 0.6), (8, 0.9, 0.6)])
def test_exact_sums_bit_identical_to_oracle(dev, orc, dyadic, lanes, long_q, mid_q):
    d = dyadic
    lengths = d["lengths"]
    long_min = int(np.quantile(lengths, long_q)) if long_q else 0
    mid_min = int(np.quantile(lengths, mid_q)) if mid_q else 0
    u_o, n_o, m_o = orc.refine_lists(d["ubis"], d["gv"], d["idx"], d["off"], 0.2, 1)
    assert np.median(n_o) >= 50 and np.mean(m_o > 0) > 0.9  # kept peaks off the lattice too
    assert not np.array_equal(u_o, d["ubis"])
    gv, idx, off = T(d["gv"], dev), T(d["idx"], dev, torch.int32), T(d["off"], dev, torch.int64)
    uf, nf, mf = batched.refine_lists(T(d["ubis"], dev), gv, idx, off, 0.2, 1,
                                      lanes_per_grain=lanes, long_min=long_min, mid_min=mid_min)
    assert np.array_equal(nf.cpu().numpy(), n_o)
    assert np.array_equal(mf.cpu().numpy().view(np.int64), m_o.view(np.int64))
    assert np.array_equal(uf.cpu().numpy().view(np.int64), u_o.view(np.int64))
    work = T(d["ubis"], dev)
    g = refine.RefineGraph(work, gv, idx, off, 0.2, 1, lanes_per_grain=lanes, long_min=long_min,
                           mid_min=mid_min)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(work, uf)
    g.close()
