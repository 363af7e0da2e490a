"""Point-by-point refinement (``ImageD11/sinograms/point_by_point.py``)
with the per-voxel scoring on the GPU.

``refine_map`` (:1432-1670) refines every UBI of a multi-valued map: for
each voxel, the peaks whose ray passes through it (``get_voxel_idx``) are
scored against each of its UBIs, the assigned ones merged by (h, k, l, eta
sign, dtyi) with intensity weights, the merged peaks re-selected for the
voxel and re-scored with ``merge_tol``, a weighted least-squares UBI fitted,
the peaks re-scored with it, the distinct reflections counted, and a strain
tensor fitted to the directional strains.  The reference runs it with numba
``prange`` over voxels, and each voxel first tests every peak of the scan.

Here that O(voxels x peaks) selection and first scoring of every (voxel,
UBI) task is one kernel in two passes (``gs_pbp_assign``: count, then fill
the assigned (peak, hkl) lists), and the g-vectors are computed on the GPU
(``compute_gve``: ``gs_compute_xlylzl_xpos_dev`` + ``gs_compute_gv_dev``).
Everything after the first selection works on the assigned rows only and
is vectorised across all tasks at once: the merge is one ``np.bincount``
per column over (task, key) groups (the reference's per-voxel bincount,
the same sums in the same order), the least-squares fits are per-task 3x3
and 6x6 normal equations accumulated with ``np.bincount`` and solved in one
batched call (the reference's ``np.linalg.lstsq``: equal for full-rank
problems up to rounding; the rank it reports is computed from the same
singular values), and the outlier statistics are per-task sums.
"""
import numpy as np
import torch

from . import _lib, geometry
from ._lib import darr, dptr, stream_ptr

# ---------------------------------------------------------------------------
# helpers with the reference's names and semantics (vectorised numpy)


def get_voxel_idx(y0, xi0, yi0, sinomega, cosomega, dty, ystep):
    """Peaks whose ray passes through the voxel at (xi0, yi0): (indices,
    ydist) (:359-368)."""
    ydist = np.abs(y0 - xi0 * sinomega - yi0 * cosomega - dty)
    return np.where(ydist <= ystep)[0], ydist


def count_unique_peaks(hkl, etasign, dtyi):
    """Rank of each peak's (h, k, l, etasign, dtyi) among the sorted
    distinct keys (:284-307)."""
    return np.unique(np.vstack((hkl, etasign, dtyi)), axis=1, return_inverse=True)[1].ravel()


def count_unique_peaks_no_dtyi(hkl, etasign):
    """As count_unique_peaks without dtyi (:311-334)."""
    return np.unique(np.vstack((hkl, etasign)), axis=1, return_inverse=True)[1].ravel()


def merge(hkl, etasign, dtyi, sum_intensity, sc, fc, omega, dty, xpos_refined, eta):
    """Intensity-weighted means of the peaks sharing (h, k, l, etasign,
    dtyi) (:338-356)."""
    labels = count_unique_peaks(hkl, etasign, dtyi)
    sI = np.bincount(labels, weights=sum_intensity)
    f = lambda x: np.bincount(labels, weights=x * sum_intensity) / sI
    return sI, f(sc), f(fc), f(omega), f(dty), f(xpos_refined), f(eta)


def gve_norm(gve):
    """Column norms of a (3, n) array (:423-429)."""
    return np.sqrt(np.sum(gve * gve, axis=0))


def divide_where(arr1, arr2, out, wherearr):
    """arr1 / arr2 where wherearr != 0, else out (:433-439)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(wherearr != 0, np.divide(arr1, arr2), out)


def ubi_to_unitcell(ubi):
    """(..., 3, 3) -> (..., 6) cells from the metric UBI.UBI^T (:443-451)."""
    u = np.asarray(ubi, float)
    G = np.matmul(u, np.swapaxes(u, -1, -2))
    a, b, c = (np.sqrt(G[..., i, i]) for i in range(3))
    al = np.degrees(np.arccos(G[..., 1, 2] / b / c))
    be = np.degrees(np.arccos(G[..., 0, 2] / a / c))
    ga = np.degrees(np.arccos(G[..., 0, 1] / a / b))
    return np.stack([a, b, c, al, be, ga], -1)


def _b_busing_levy(ucell):
    """(..., 6) -> (..., 3, 3) B of ubi_and_ucell_to_u (:458-487)."""
    a, b, c = ucell[..., 0], ucell[..., 1], ucell[..., 2]
    ca, cb, cg = (np.cos(np.radians(ucell[..., i])) for i in (3, 4, 5))
    g = np.stack([np.stack([a * a, a * b * cg, a * c * cb], -1),
                  np.stack([a * b * cg, b * b, b * c * ca], -1),
                  np.stack([a * c * cb, b * c * ca, c * c], -1)], -2)
    gi = np.linalg.inv(g)
    astar, bstar, cstar = (np.sqrt(gi[..., i, i]) for i in range(3))
    betas = np.degrees(np.arccos(gi[..., 0, 2] / astar / cstar))
    gammas = np.degrees(np.arccos(gi[..., 0, 1] / astar / bstar))
    B = np.zeros(ucell.shape[:-1] + (3, 3))
    B[..., 0, 0] = astar
    B[..., 0, 1] = bstar * np.cos(np.radians(gammas))
    B[..., 0, 2] = cstar * np.cos(np.radians(betas))
    B[..., 1, 1] = bstar * np.sin(np.radians(gammas))
    B[..., 1, 2] = -cstar * np.sin(np.radians(betas)) * ca
    B[..., 2, 2] = 1.0 / c
    return B


def ubi_and_ucell_to_u(ubi, ucell):
    """U = (B(ucell).UBI)^T (:455-490), for one or a stack of UBIs."""
    B = _b_busing_levy(np.asarray(ucell, float))
    return np.swapaxes(np.matmul(B, np.asarray(ubi, float)), -1, -2)


def hkluniq(ubi, gx, gy, gz, eta, m, tol, hmax):
    """(assigned peaks, distinct (h, k, l, eta > 0) with |h|,|k|,|l| < hmax)
    among the peaks of mask m (:52-81)."""
    g = np.stack([gx, gy, gz])[:, np.asarray(m, bool)]
    e = np.asarray(eta)[np.asarray(m, bool)]
    H = np.dot(ubi, g)
    ih = np.round(H)
    ok = ((H - ih) ** 2).sum(axis=0) < tol * tol
    inside = ok & np.all(np.abs(ih) < hmax, axis=0)
    keys = np.vstack((ih[:, inside], (e[inside] > 0)))
    return int(ok.sum()), int(np.unique(keys, axis=1).shape[1]) if inside.any() else 0


# ---------------------------------------------------------------------------
# device pieces


def _dev():
    _lib.load()
    _lib.require_gpu()
    return torch.device("cuda")


def _t(a, dev, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)


def compute_gve(sc, fc, omega, xpos, distance, y_center, y_size, tilt_y, z_center, z_size,
                tilt_z, tilt_x, o11, o12, o21, o22, t_x, t_y, t_z, wedge, chi, wavelength):
    """g-vectors (3, n) of peaks seen from a per-peak origin x position
    ``xpos`` (:372-404) on the GPU: compute_xlylzl with x - xpos, then
    compute_gv with the translation (t_x, t_y, t_z) at omega (no sign)."""
    dev = _dev()
    g = _gve_device(_t(sc, dev), _t(fc, dev), _t(omega, dev), _t(xpos, dev), dict(
        distance=distance, y_center=y_center, y_size=y_size, tilt_y=tilt_y, z_center=z_center,
        z_size=z_size, tilt_z=tilt_z, tilt_x=tilt_x, o11=o11, o12=o12, o21=o21, o22=o22,
        t_x=t_x, t_y=t_y, t_z=t_z, wedge=wedge, chi=chi, wavelength=wavelength))
    return g.cpu().numpy().T.copy()


def _gve_device(sc, fc, omega, xpos, geo):
    lib = _lib.load()
    lib.gs_clear_error()
    n = sc.numel()
    dev = sc.device
    xl = torch.empty((n, 3), dtype=torch.float64, device=dev)
    gv = torch.empty((n, 3), dtype=torch.float64, device=dev)
    if n == 0:
        return gv
    pars = dict(geo, omegasign=1.0)
    p, r, dist = geometry.pack_xlylzl(pars)
    _lib.check(lib.gs_compute_xlylzl_xpos_dev(dptr(sc), dptr(fc), darr(p), darr(r), darr(dist),
                                              dptr(xpos), dptr(xl), n, stream_ptr()),
               "xlylzl_xpos")
    t = darr([geo["t_x"], geo["t_y"], geo["t_z"]])
    _lib.check(lib.gs_compute_gv_dev(dptr(xl), dptr(omega), 1.0, float(geo["wavelength"]),
                                     float(geo["wedge"]), float(geo["chi"]), t, dptr(gv), n,
                                     stream_ptr()), "compute_gv")
    return gv


def assign_tasks(so, co, dty, gv, pts, ptask, ubis, y0, ystep, tol):
    """gs_pbp_assign, both passes: for every task t (tasks grouped by voxel:
    ptask (P+1,) offsets over the voxels pts (P,2)), the peaks local to its
    voxel and assigned to its UBI.  Device tensors in; returns (task_off
    (T+1,) int64 numpy, peak (E,) int64 numpy, hkl (3, E) int32 numpy),
    peaks ascending within each task."""
    lib = _lib.load()
    lib.gs_clear_error()
    dev = gv.device
    n = gv.shape[0]
    P = pts.shape[0]
    T = ubis.shape[0]
    counts = torch.empty(T, dtype=torch.int32, device=dev)
    _lib.check(lib.gs_pbp_assign(dptr(so), dptr(co), dptr(dty), dptr(gv), n, dptr(pts),
                                 dptr(ptask), P, dptr(ubis), T, float(y0), float(ystep),
                                 float(tol), dptr(counts), None, None, None, None, stream_ptr()),
               "pbp_assign count")
    cnt = counts.cpu().numpy().astype(np.int64)
    off = np.zeros(T + 1, np.int64)
    off[1:] = np.cumsum(cnt)
    E = int(off[-1])
    peak = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
    hkl = torch.empty((max(E, 1), 3), dtype=torch.int32, device=dev)
    cursor = torch.empty(T, dtype=torch.int32, device=dev)
    off_d = _t(off, dev, torch.int64)
    _lib.check(lib.gs_pbp_assign(dptr(so), dptr(co), dptr(dty), dptr(gv), n, dptr(pts),
                                 dptr(ptask), P, dptr(ubis), T, float(y0), float(ystep),
                                 float(tol), dptr(counts), dptr(off_d), dptr(cursor), dptr(peak),
                                 dptr(hkl), stream_ptr()), "pbp_assign fill")
    pk = peak[:E].cpu().numpy().astype(np.int64)
    hk = hkl[:E].cpu().numpy()
    task = np.repeat(np.arange(T), cnt)
    order = np.lexsort((pk, task))
    return off, pk[order], np.ascontiguousarray(hk[order].T)


# ---------------------------------------------------------------------------
# batched small linear algebra on per-task groups of rows


def _segment_sum(task, values, T):
    """Per-task sums of each column of values (rows, k) in row order."""
    return np.stack([np.bincount(task, weights=values[:, k], minlength=T)
                     for k in range(values.shape[1])], -1)


def _lstsq_rank(ata, n_rows):
    """Numerical rank of the (n x m) design matrix with Gram matrix ata, as
    np.linalg.lstsq counts it: singular values above eps * max(n, m) times
    the largest."""
    m = ata.shape[-1]
    s = np.sqrt(np.clip(np.linalg.eigvalsh(ata), 0, None))
    rcond = np.finfo(np.float64).eps * np.maximum(n_rows, m)
    return np.sum(s > (rcond * s.max(-1))[..., None], axis=-1)


def _solve_normal(ata, atb):
    """Least-squares solutions from the normal equations; tasks whose Gram
    matrix is singular get NaN (the caller handles them)."""
    out = np.full(atb.shape, np.nan)
    ok = np.abs(np.linalg.det(ata)) > 0
    if ok.any():
        out[ok] = np.linalg.solve(ata[ok], atb[ok])
    return out


def _sym_outer(x, y=None):
    """Per-row products x_i y_j (rows, 3, 3) flattened to (rows, 9)."""
    y = x if y is None else y
    return (x[:, :, None] * y[:, None, :]).reshape(len(x), -1)


def refine_map(refine_points, all_pbpmap_ubis, ri_col, rj_col, sx_grid, sy_grid, mask,
               sc, fc, eta, sum_intensity, sinomega, cosomega, omega, dty, dtyi, xpos,
               ystep, y0, B0, distance, y_center, y_size, tilt_y, z_center, z_size, tilt_z,
               tilt_x, o11, o12, o21, o22, t_x, t_y, t_z, wedge, chi, wavelength,
               tol=0.1, merge_tol=0.05, min_grain_npks=6, gnoise_std=1e-4):
    """Refine every UBI of a point-by-point map (:1432-1670).  Same
    arguments and results as the reference: (final_ubis (3,3,M),
    final_eps (3,3,M), final_npks (M,), final_nuniq (M,)), NaN where a UBI
    was not refined."""
    dev = _dev()
    geo = dict(distance=distance, y_center=y_center, y_size=y_size, tilt_y=tilt_y,
               z_center=z_center, z_size=z_size, tilt_z=tilt_z, tilt_x=tilt_x, o11=o11, o12=o12,
               o21=o21, o22=o22, t_x=t_x, t_y=t_y, t_z=t_z, wedge=wedge, chi=chi,
               wavelength=wavelength)
    M = all_pbpmap_ubis.shape[2]
    final_ubis = np.full_like(all_pbpmap_ubis, np.nan, dtype=np.float64)
    final_eps = np.full_like(all_pbpmap_ubis, np.nan, dtype=np.float64)
    final_npks = np.full(M, np.nan)
    final_nuniq = np.full(M, np.nan)
    sc, fc, eta, sum_intensity = (np.asarray(a, np.float64) for a in (sc, fc, eta, sum_intensity))
    sinomega, cosomega, omega, dty, xpos = (np.asarray(a, np.float64)
                                            for a in (sinomega, cosomega, omega, dty, xpos))
    dtyi = np.asarray(dtyi)
    ri_col, rj_col = np.asarray(ri_col), np.asarray(rj_col)
    # ---- tasks: (voxel, UBI) pairs in the reference's visiting order
    by_point = {}
    for m in range(M):
        by_point.setdefault((int(ri_col[m]), int(rj_col[m])), []).append(m)
    pts, ptask, tmap = [], [0], []
    for ri, rj in refine_points:
        ri, rj = int(ri), int(rj)
        if not mask[ri, rj]:
            continue
        here = by_point.get((ri, rj), [])
        valid = [m for m in here if not np.isnan(all_pbpmap_ubis[0, 0, m])]
        if not valid:
            continue
        pts.append((sx_grid[ri, rj], sy_grid[ri, rj]))
        tmap.extend(valid)
        ptask.append(len(tmap))
    T = len(tmap)
    if T == 0:
        return final_ubis, final_eps, final_npks, final_nuniq
    tmap = np.array(tmap, np.int64)
    pts = np.array(pts, np.float64)
    ptask = np.array(ptask, np.int32)
    tpoint = np.repeat(np.arange(len(pts)), np.diff(ptask))
    ubis = np.ascontiguousarray(np.moveaxis(all_pbpmap_ubis[:, :, tmap], -1, 0))
    # ---- GPU: g-vectors of every peak, selection + first scoring
    gv_all = _gve_device(_t(sc, dev), _t(fc, dev), _t(omega, dev), _t(xpos, dev), geo)
    off, peak, hkl = assign_tasks(_t(sinomega, dev), _t(cosomega, dev), _t(dty, dev), gv_all,
                                  _t(pts, dev), _t(ptask, dev, torch.int32), _t(ubis, dev), y0,
                                  ystep, tol)
    E = len(peak)
    task = np.repeat(np.arange(T), np.diff(off))
    # ---- merge by (task, h, k, l, etasign, dtyi); groups in sorted key order
    etas = np.sign(eta[peak])
    keys = (dtyi[peak], etas, hkl[2], hkl[1], hkl[0], task)
    order = np.lexsort(keys)
    ks = np.stack([k[order] for k in keys[::-1]])
    new = np.ones(E, bool)
    new[1:] = np.any(ks[:, 1:] != ks[:, :-1], axis=0)
    gid_sorted = np.cumsum(new) - 1
    gid = np.empty(E, np.int64)
    gid[order] = gid_sorted
    G = int(gid_sorted[-1]) + 1 if E else 0
    gtask = np.zeros(G, np.int64)
    gtask[gid] = task
    ghkl = np.zeros((3, G), np.int64)
    ghkl[:, gid] = hkl
    wI = sum_intensity[peak]
    sI = np.bincount(gid, weights=wI, minlength=G)
    mean = lambda x: np.bincount(gid, weights=x[peak] * wI, minlength=G) / sI
    m_sc, m_fc, m_om, m_dty, m_xpos, m_eta = (mean(x) for x in (sc, fc, omega, dty, xpos, eta))
    # ---- re-selection of the merged peaks for the voxel, merged g-vectors
    gpt = pts[tpoint[gtask]]
    yd = np.abs(y0 - gpt[:, 0] * np.sin(np.radians(m_om)) - gpt[:, 1] *
                np.cos(np.radians(m_om)) - m_dty)
    keep = np.nonzero(yd <= ystep)[0]
    rtask = gtask[keep]
    rydist = yd[keep]
    reta = m_eta[keep]
    g = _gve_device(_t(m_sc[keep], dev), _t(m_fc[keep], dev), _t(m_om[keep], dev),
                    _t(m_xpos[keep], dev), geo).cpu().numpy()

    def score(u):
        h = np.einsum("rij,rj->ri", u, g)
        hi = np.round(h).astype(np.int32)
        e = h - hi
        return hi, (e ** 2).sum(axis=1) < merge_tol ** 2
    hi, sel = score(ubis[rtask])
    npks = np.bincount(rtask, weights=sel, minlength=T).astype(np.int64)
    fit = npks > min_grain_npks
    # ---- weighted least-squares UBI for the tasks with enough peaks
    rows = sel & fit[rtask]
    w = 1.0 / (rydist + 1)
    wg = w[:, None] * g
    wh = w[:, None] * hi.astype(np.float64)
    rt = rtask[rows]
    ata = _segment_sum(rt, _sym_outer(wg[rows]), T).reshape(T, 3, 3)
    atb = _segment_sum(rt, _sym_outer(wg[rows], wh[rows]), T).reshape(T, 3, 3)
    ubifit = np.full((T, 3, 3), np.nan)
    ubifit[fit] = np.swapaxes(_solve_normal(ata[fit], atb[fit]), -1, -2)
    rank = np.zeros(T, np.int64)
    rank[fit] = _lstsq_rank(ata[fit], npks[fit])
    worth = fit & (rank == 3) & np.all(np.isfinite(ubifit), axis=(1, 2))
    if worth.any():
        u = ubifit[worth]
        good = (np.linalg.cond(u) < 1e14) & (np.linalg.det(u) > 0) & \
            (np.linalg.matrix_rank(u) == 3)
        worth[np.nonzero(worth)[0][~good]] = False
    ubi_out = np.where(worth[:, None, None], ubifit, ubis)
    # ---- re-score with the fitted UBIs, count distinct reflections
    hi2, sel2 = score(ubi_out[rtask])
    npk2 = np.bincount(rtask, weights=sel2, minlength=T).astype(np.int64)
    r2 = sel2 & fit[rtask]
    ukeys = np.stack([rtask[r2], hi2[r2, 0], hi2[r2, 1], hi2[r2, 2],
                      np.sign(reta[r2]).astype(np.int64)])
    nuniq = np.bincount(np.unique(ukeys, axis=1)[0], minlength=T) if r2.any() else \
        np.zeros(T, np.int64)
    for t in np.nonzero(fit)[0]:
        m = tmap[t]
        final_ubis[:, :, m] = ubi_out[t]
        final_npks[m] = npk2[t]
        final_nuniq[m] = nuniq[t]
    # ---- strain tensors of the well-fitted tasks
    if worth.any():
        eps = _strain_fit(worth, ubifit, rtask, sel2, g, hi2, rydist, B0, gnoise_std, T)
        for t in np.nonzero(worth)[0]:
            final_eps[:, :, tmap[t]] = eps[t]
    return final_ubis, final_eps, final_npks, final_nuniq


def _strain_fit(worth, ubifit, rtask, sel2, g, hi2, rydist, B0, gnoise_std, T):
    """The directional-strain fit of refine_map (:1624-1668) for every task
    in ``worth`` at once."""
    rows = sel2 & worth[rtask]
    t = rtask[rows]
    gg = g[rows]
    hk = hi2[rows].astype(np.float64)
    ucell = ubi_to_unitcell(ubifit[worth])
    U = np.full((T, 3, 3), np.nan)
    U[worth] = ubi_and_ucell_to_u(ubifit[worth], ucell)
    UB0 = np.matmul(U, B0)
    gve0 = np.einsum("rij,rj->ri", UB0[t], hk)
    gTg0 = np.sum(gg * gve0, axis=1)
    gTg = np.sum(gg * gg, axis=1)
    strain = (gTg0 / gTg) - 1
    kap = gg / np.sqrt(np.sum(gg * gg, axis=1))[:, None]
    kx, ky, kz = kap.T
    Mrow = np.stack((kx * kx, ky * ky, kz * kz, 2 * kx * ky, 2 * kx * kz, 2 * ky * kz), 1)
    w = 1. / (rydist[rows] + 1)
    a = np.sum(gve0 * (gnoise_std ** 2) * gve0, axis=1)
    noise = np.sqrt(divide_where(a, gTg ** 2, out=np.ones_like(gTg), wherearr=gTg))
    w = w * (1. / noise)
    n = np.bincount(t, minlength=T).astype(np.float64)
    mean = np.bincount(t, weights=strain, minlength=T) / np.maximum(n, 1)
    std = np.sqrt(np.bincount(t, weights=(strain - mean[t]) ** 2, minlength=T) / np.maximum(n, 1))
    w = np.where(strain > mean[t] + std[t] * 3.5, 0.0, w)
    w = np.where(strain < mean[t] - std[t] * 3.5, 0.0, w)
    wmax = np.zeros(T)
    np.maximum.at(wmax, t, w)
    w = w / wmax[t]
    aM = w[:, None] * Mrow
    b = w * strain
    ata = _segment_sum(t, (aM[:, :, None] * aM[:, None, :]).reshape(len(t), 36), T)
    atb = _segment_sum(t, aM * b[:, None], T)
    ata = ata.reshape(T, 6, 6)
    e = np.full((T, 6), np.nan)
    idx = np.nonzero(worth)[0]
    sol = _solve_normal(ata[idx], atb[idx][:, :, None])[:, :, 0]
    e[idx] = sol
    # rank-deficient tasks: the minimum-norm solution, as lstsq gives it
    for k, tt in enumerate(idx):
        if not np.all(np.isfinite(sol[k])):
            r = t == tt
            e[tt] = np.linalg.lstsq(aM[r], b[r], rcond=np.finfo(float).eps * max(r.sum(), 6))[0]
    sxx, syy, szz, sxy, sxz, syz = e.T
    return np.stack([np.stack([sxx, sxy, sxz], -1), np.stack([sxy, syy, syz], -1),
                     np.stack([sxz, syz, szz], -1)], -2)
