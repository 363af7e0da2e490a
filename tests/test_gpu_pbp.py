"""Point-by-point refinement on the GPU against the oracle's voxel loop.

The first selection (gs_pbp_assign) is compared entry for entry: every
task's peak list and hkl.  refine_map's results: NaN patterns, peak and
distinct-reflection counts exactly; UBIs to 1e-9 and strain tensors to
1e-10 (the GPU's g-vectors follow compute_gv's arithmetic, the reference
the numpy chain: ~1e-15 relative; the fits solve normal equations where
the reference calls lstsq)."""
import numpy as np
import pytest
import torch

from imaged11_amd import simulate
from imaged11_amd import point_by_point as PP
from oracle import pbp_oracle as P

pytestmark = pytest.mark.gpu

ARGS = ("refine_points", "all_pbpmap_ubis", "ri_col", "rj_col", "sx_grid", "sy_grid", "mask",
        "sc", "fc", "eta", "sum_intensity", "sinomega", "cosomega", "omega", "dty", "dtyi",
        "xpos", "ystep", "y0", "B0")


@pytest.fixture(scope="module")
def wl():
    return simulate.pbp_workload()


def test_compute_gve_matches_oracle(dev, wl):
    g = PP.compute_gve(wl["sc"], wl["fc"], wl["omega"], wl["xpos"], **wl["geometry"])
    want = P.compute_gve(wl["sc"], wl["fc"], wl["omega"], wl["xpos"], **wl["geometry"])
    np.testing.assert_allclose(g, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_assign_tasks_matches_oracle(dev, wl):
    """Every voxel of the map with all of its UBIs, the kernel's lists
    against get_voxel_idx + the first scoring of refine_map."""
    T_ = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                     device=dev)
    gv = P.compute_gve(wl["sc"], wl["fc"], wl["omega"], wl["xpos"], **wl["geometry"]).T.copy()
    NY = wl["mask"].shape[0]
    pts, ptask, ubis = [], [0], []
    M = wl["all_pbpmap_ubis"].shape[2]
    for i in range(NY):
        for j in range(NY):
            here = [m for m in range(M) if wl["ri_col"][m] == i and wl["rj_col"][m] == j and
                    not np.isnan(wl["all_pbpmap_ubis"][0, 0, m])]
            if not here:
                continue
            pts.append((wl["sx_grid"][i, j], wl["sy_grid"][i, j]))
            ubis.extend(wl["all_pbpmap_ubis"][:, :, m] for m in here)
            ptask.append(len(ubis))
    pts, ptask, ubis = np.array(pts), np.array(ptask, np.int32), np.array(ubis)
    off, _, peak, hkl = PP.assign_tasks(T_(wl["sinomega"]), T_(wl["cosomega"]), T_(wl["dty"]),
                                        T_(gv), T_(pts), T_(ptask, torch.int32), T_(ubis),
                                        wl["y0"], wl["ystep"], 0.1)
    off, peak, hkl = off.cpu().numpy(), peak.cpu().numpy(), hkl.cpu().numpy().T
    total = 0
    for p in range(len(pts)):
        li, _ = P.get_voxel_idx(wl["y0"], pts[p, 0], pts[p, 1], wl["sinomega"], wl["cosomega"],
                                wl["dty"], wl["ystep"])
        for t in range(ptask[p], ptask[p + 1]):
            hi, m = P._assign(ubis[t], gv[li].T, 0.1)
            sl = slice(off[t], off[t + 1])
            assert np.array_equal(peak[sl], li[m])
            assert np.array_equal(hkl[:, sl], hi[:, m])
            total += m.sum()
    assert total == len(peak) > 5000


@pytest.mark.parametrize("tol,merge_tol,minpks", [(0.1, 0.05, 6), (0.05, 0.03, 50)])
def test_refine_map_matches_oracle(dev, wl, tol, merge_tol, minpks):
    kw = dict(tol=tol, merge_tol=merge_tol, min_grain_npks=minpks)
    a = PP.refine_map(*[wl[k] for k in ARGS], **wl["geometry"], **kw)
    b = P.refine_map(*[wl[k] for k in ARGS], **wl["geometry"], **kw)
    for x, y in zip(a, b):
        assert np.array_equal(np.isnan(x), np.isnan(y))
    np.testing.assert_allclose(a[0], b[0], atol=1e-9, equal_nan=True)
    np.testing.assert_allclose(a[1], b[1], atol=1e-10, equal_nan=True)
    assert np.array_equal(a[2], b[2], equal_nan=True)
    assert np.array_equal(a[3], b[3], equal_nan=True)
    assert (~np.isnan(a[1][0, 0])).sum() >= 40


def test_segment_sum_is_bincount_order(dev):
    """gs_segment_sum adds each segment row by row from 0.0, the order of
    np.bincount: bit for bit equal, with empty segments."""
    rng = np.random.default_rng(7)
    lens = rng.integers(0, 40, 500)
    off = np.zeros(501, np.int64)
    off[1:] = np.cumsum(lens)
    v = rng.normal(size=(off[-1], 5)) * 10.0 ** rng.integers(-8, 8, size=(off[-1], 1))
    got = PP.segment_sum(torch.as_tensor(v, device=dev), torch.as_tensor(off, device=dev))
    lab = np.repeat(np.arange(500), lens)
    want = np.stack([np.bincount(lab, weights=v[:, c], minlength=500) for c in range(5)], 1)
    assert np.array_equal(got.cpu().numpy().view(np.int64), want.view(np.int64))


def test_refine_map_edge_cases(dev, wl):
    """No point to refine, nothing assigned, and too few peaks to fit: all
    NaN, as the reference leaves them."""
    args = {k: wl[k] for k in ARGS}
    # only masked-out points
    outside = [(i, j) for i in range(wl["mask"].shape[0]) for j in range(wl["mask"].shape[1])
               if not wl["mask"][i, j]]
    a = dict(args, refine_points=np.array(outside[:5]))
    for r in PP.refine_map(*[a[k] for k in ARGS], **wl["geometry"]):
        assert np.isnan(r).all()
    # a tolerance nothing passes
    for r in PP.refine_map(*[args[k] for k in ARGS], **wl["geometry"], tol=1e-300):
        assert np.isnan(r).all()
    # a peak minimum no UBI reaches: nothing fitted, nothing written
    got = PP.refine_map(*[args[k] for k in ARGS], **wl["geometry"], min_grain_npks=10 ** 9)
    want = P.refine_map(*[args[k] for k in ARGS], **wl["geometry"], min_grain_npks=10 ** 9)
    for x, y in zip(got, want):
        assert np.isnan(x).all() and np.isnan(y).all()


def _T(a, dev, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)


def _lstsq_tasks(rng, nc, nr, T=300):
    """Random per-task systems: full rank, rank-deficient (rows in a
    subspace), fewer rows than columns, and scaled rows."""
    rows, off = [], [0]
    for t in range(T):
        n = int(rng.integers(1, 40))
        a = rng.normal(size=(n, nc)) * 10.0 ** rng.uniform(-2, 2)
        kind = t % 4
        if kind == 1:  # rank nc - 1 or less: rows in a random subspace
            k = int(rng.integers(1, nc))
            a = rng.normal(size=(n, k)) @ rng.normal(size=(k, nc))
        elif kind == 2:  # exactly repeated directions
            a = np.repeat(a[:1], n, 0) * rng.uniform(0.5, 2, size=(n, 1))
        b = rng.normal(size=(n, nr))
        rows.append(np.hstack([a, b]))
        off.append(off[-1] + n)
    return np.vstack(rows), np.array(off, np.int64)


def test_pbp_strain_fit_is_lstsq(dev):
    """gs_pbp_strain_fit against np.linalg.lstsq(a, b, rcond=eps*max(n, 6))
    per task: rank-deficient and under-determined tasks get the minimum-norm
    solution; tasks not marked stay NaN; a non-finite row gives NaN."""
    rng = np.random.default_rng(8)
    d, off = _lstsq_tasks(rng, 6, 1)
    T = len(off) - 1
    worth = np.ones(T, bool)
    worth[::7] = False
    d[off[5]] = np.nan  # task 5: a NaN row
    got = PP.strain_fit(_T(d, dev), _T(off, dev, torch.int64), _T(worth, dev, torch.bool)).cpu().numpy()
    eps = np.finfo(float).eps
    for t in range(T):
        if not worth[t] or t == 5:
            assert np.isnan(got[t]).all()
            continue
        a, b = d[off[t]:off[t + 1], :6], d[off[t]:off[t + 1], 6]
        want = np.linalg.lstsq(a, b, rcond=eps * max(len(a), 6))[0]
        scale = np.abs(want).max() + 1e-300
        assert np.abs(got[t] - want).max() <= 1e-9 * scale, t


def test_pbp_ubi_fit_is_lstsq_and_quality_test(dev):
    """gs_pbp_ubi_fit against np.linalg.lstsq per task (solution and rank)
    and the reference's cond / det / matrix_rank test; ub0 = U.B0 of the
    fitted cell (the oracle's ubi_and_ucell_to_u) where the fit is kept."""
    rng = np.random.default_rng(9)
    d, off = _lstsq_tasks(rng, 3, 3)
    # half of the tasks: a lattice-like fit (rows w g, w hkl with hkl = UBI g)
    for t in range(0, len(off) - 1, 2):
        n = off[t + 1] - off[t]
        ub = np.linalg.inv(np.diag([4.0, 5.0, 6.0]) + rng.normal(scale=0.1, size=(3, 3)))
        h = rng.integers(-4, 5, size=(n, 3)).astype(float)
        w = rng.uniform(0.3, 1.0, size=(n, 1))
        d[off[t]:off[t + 1]] = np.hstack([w * (h @ ub.T), w * h])
    T = len(off) - 1
    fit = np.ones(T, bool)
    fit[1::9] = False
    ubis = rng.normal(size=(T, 3, 3))
    B0 = np.eye(3) / 4.05
    ubifit, rank, worth, ubi_out, ub0 = (x.cpu().numpy() for x in PP.ubi_fit(
        _T(d, dev), _T(off, dev, torch.int64), _T(fit, dev, torch.bool), _T(ubis, dev), B0))
    eps = np.finfo(float).eps
    n_worth = 0
    for t in range(T):
        if not fit[t]:
            assert not worth[t] and np.isnan(ubifit[t]).all()
            assert np.array_equal(ubi_out[t], ubis[t])
            continue
        a, b = d[off[t]:off[t + 1], :3], d[off[t]:off[t + 1], 3:]
        x, _, r, _ = np.linalg.lstsq(a, b, rcond=eps * max(len(a), 3))
        assert rank[t] == r, t
        u = x.T
        np.testing.assert_allclose(ubifit[t], u, rtol=0, atol=1e-9 * (np.abs(u).max() + 1e-300))
        ok = (r == 3 and np.linalg.cond(u) < 1e14 and np.linalg.det(u) > 0 and
              np.linalg.matrix_rank(u) == 3)
        assert bool(worth[t]) == ok, t
        if ok:
            n_worth += 1
            np.testing.assert_array_equal(ubi_out[t], ubifit[t])
            want = P.ubi_and_ucell_to_u(u, P.ubi_to_unitcell(u)) @ B0
            np.testing.assert_allclose(ub0[t], want, rtol=0, atol=1e-9 * np.abs(want).max())
        else:
            assert np.array_equal(ubi_out[t], ubis[t]) and np.isnan(ub0[t]).all()
    assert n_worth > 100


def test_segment_max_is_max_with_zero(dev):
    """gs_segment_max: the maximum of each segment's rows and 0.0 (the
    zeros torch's scatter_reduce reduces into), NaN propagated, empty
    segments 0.0."""
    rng = np.random.default_rng(10)
    lens = rng.integers(0, 30, 400)
    off = np.zeros(401, np.int64)
    off[1:] = np.cumsum(lens)
    v = rng.normal(size=(off[-1], 3))
    v[off[7]:off[8]] = -np.abs(v[off[7]:off[8]])  # all negative: 0.0
    if lens[9]:
        v[off[9], 1] = np.nan
    got = PP.segment_max(_T(v, dev), _T(off, dev, torch.int64)).cpu().numpy()
    t = np.repeat(np.arange(400), lens)
    want = torch.zeros((400, 3), dtype=torch.float64).scatter_reduce(
        0, torch.as_tensor(np.repeat(t[:, None], 3, 1)), torch.as_tensor(v), "amax").numpy()
    assert np.array_equal(got, want, equal_nan=True)


def test_merge_sums_equal_the_unfused_columns(dev):
    """gs_pbp_merge_sums against gathering, weighting and permuting the
    columns then summing each group with gs_segment_sum: bit for bit."""
    rng = np.random.default_rng(11)
    n, E = 5000, 40000
    peak = _T(rng.integers(0, n, E), dev, torch.int64)
    gid = _T(rng.integers(0, 3000, E), dev, torch.int64)
    perm = torch.sort(gid, stable=True).indices
    counts = torch.bincount(gid, minlength=3000)
    goff = torch.zeros(3001, dtype=torch.int64, device=dev)
    goff[1:] = torch.cumsum(counts, 0)
    si = _T(rng.uniform(1, 100, n), dev)
    cols = [_T(rng.normal(size=n) * 10.0 ** rng.integers(-3, 3), dev) for _ in range(6)]
    got = PP.merge_sums(perm, peak, si, cols, goff)
    wI = si[peak]
    want = PP.segment_sum(torch.stack([wI] + [c[peak] * wI for c in cols], 1)[perm], goff)
    assert torch.equal(got, want)


def test_rows_dot_kernel_bit_exact(dev):
    """gs_pbp_rows_dot against the same row products in numpy, bit for bit:
    h = ubis[task].g, its half-to-even rounding and the tolerance test, with
    rows exactly on half-integers and on the tolerance."""
    import torch
    from imaged11_amd import point_by_point as PP
    rng = np.random.default_rng(17)
    T, R = 50, 20000
    ubis = rng.normal(size=(T, 3, 3)) * 3
    task = rng.integers(0, T, R)
    g = rng.normal(size=(R, 3))
    g[:10] = 0.0
    ubis[0] = np.eye(3)
    task[:10] = 0
    g[:10, 0] = [0.5, 1.5, 2.5, -0.5, 0.05, 0.049999999999999996, 0.0500000001, 1.0, 3.25, -2.5]
    u = ubis[task]
    want = np.stack([(u[:, i, 0] * g[:, 0] + u[:, i, 1] * g[:, 1]) + u[:, i, 2] * g[:, 2]
                     for i in range(3)], 1)
    hw = np.round(want)
    e = want - hw
    okw = ((e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]) < 0.05 * 0.05
    D = lambda a: torch.as_tensor(a, device=dev)
    h = PP._rows_dot(D(ubis), D(task), D(g)).cpu().numpy()
    assert np.array_equal(h.view(np.int64), want.view(np.int64))
    hk, ok = PP._score(D(ubis), D(task), D(g), 0.05)
    assert np.array_equal(hk.cpu().numpy(), hw.astype(np.int64))
    assert np.array_equal(ok.cpu().numpy(), okw)
    assert okw[:10].tolist() == [False] * 4 + [False, True, False, True, False, False]
