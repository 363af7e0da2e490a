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
is vectorised across all tasks at once on the device: the merge is one
ordered segmented sum per column over (task, key) groups (the reference's
per-voxel ``np.bincount``, the same sums in the same order), the
least-squares fits run one task per thread (``gs_pbp_ubi_fit``,
``gs_pbp_strain_fit``: the task's weighted rows streamed into a QR factor
by Givens rotations, the factor decomposed by one-sided Jacobi rotations,
then ``np.linalg.lstsq``'s minimum-norm solution, rank and rcond rule, and
the reference's cond / det / matrix_rank test of the fitted UBI), and the
outlier statistics are per-task sums.
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
    voxel and assigned to its UBI.  Device tensors in; returns device
    tensors (task_off (T+1,) int64, task (E,) int64, peak (E,) int64, hkl
    (E,3) int64), ordered by task, then peak index."""
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
    off = torch.zeros(T + 1, dtype=torch.int64, device=dev)
    off[1:] = torch.cumsum(counts.long(), 0)
    E = int(off[-1])
    peak = torch.empty(max(E, 1), dtype=torch.int32, device=dev)
    hkl = torch.empty((max(E, 1), 3), dtype=torch.int32, device=dev)
    cursor = torch.empty(T, dtype=torch.int32, device=dev)
    _lib.check(lib.gs_pbp_assign(dptr(so), dptr(co), dptr(dty), dptr(gv), n, dptr(pts),
                                 dptr(ptask), P, dptr(ubis), T, float(y0), float(ystep),
                                 float(tol), dptr(counts), dptr(off), dptr(cursor), dptr(peak),
                                 dptr(hkl), stream_ptr()), "pbp_assign fill")
    # ascending peak index inside each task: sort (task, peak) keys
    task = torch.repeat_interleave(torch.arange(T, device=dev), counts.long(), output_size=E)
    _, order = torch.sort(task * max(n, 1) + peak[:E].long())
    return off, task[order], peak[:E].long()[order], hkl[:E].long()[order]


def segment_sum(v, off):
    """Column sums of the row segments off[s]..off[s+1] of v (rows, C),
    row by row in order (gs_segment_sum: np.bincount's order)."""
    lib = _lib.load()
    lib.gs_clear_error()
    v = v.contiguous()
    S = off.numel() - 1
    C = v.shape[1]
    if v.shape[0] == 0:  # no rows: every segment is empty
        return torch.zeros((S, C), dtype=torch.float64, device=v.device)
    out = torch.empty((S, C), dtype=torch.float64, device=v.device)
    _lib.check(lib.gs_segment_sum(dptr(v), C, dptr(off), S, dptr(out), stream_ptr()),
               "segment_sum")
    return out


def merge_sums(perm, peak, si, cols, goff):
    """The merge's weighted sums per group (gs_pbp_merge_sums): (G, 7) =
    sums of w and of each of the six columns (sc, fc, omega, dty, xpos,
    eta) times w over the rows goff[g]..goff[g+1] of the grouped order (row
    r is peak peak[perm[r]], w its intensity), in row order from 0.0."""
    lib = _lib.load()
    lib.gs_clear_error()
    G = goff.numel() - 1
    out = torch.empty((G, 7), dtype=torch.float64, device=si.device)
    if G == 0:
        return out
    perm, peak, si = perm.contiguous(), peak.contiguous(), si.contiguous()
    cols = [c.contiguous() for c in cols]
    _lib.check(lib.gs_pbp_merge_sums(dptr(perm), dptr(peak), si.shape[0], dptr(si),
                                     *[dptr(c) for c in cols], dptr(goff), G, dptr(out),
                                     stream_ptr()), "pbp_merge_sums")
    return out


def segment_max(v, off):
    """Column maxima of the row segments off[s]..off[s+1] of v (rows, C)
    and 0.0 (gs_segment_max; an empty segment gives 0.0)."""
    lib = _lib.load()
    lib.gs_clear_error()
    v = v.contiguous()
    S = off.numel() - 1
    C = v.shape[1]
    if v.shape[0] == 0:
        return torch.zeros((S, C), dtype=torch.float64, device=v.device)
    out = torch.empty((S, C), dtype=torch.float64, device=v.device)
    _lib.check(lib.gs_segment_max(dptr(v), C, dptr(off), S, dptr(out), stream_ptr()),
               "segment_max")
    return out


def _offsets(seg, S):
    """CSR offsets of rows already ordered by segment id seg (0..S-1)."""
    off = torch.zeros(S + 1, dtype=torch.int64, device=seg.device)
    off[1:] = torch.cumsum(torch.bincount(seg, minlength=S), 0)
    return off


def _pack(cols):
    """One int64 key per row ordering rows as the tuples of cols do
    (lexicographic, first column most significant), or None when the value
    ranges need more than 62 bits."""
    lo = torch.stack([c.min() for c in cols])
    hi = torch.stack([c.max() for c in cols])
    lo_h, hi_h = lo.cpu().numpy(), hi.cpu().numpy()
    bits = [max(1, int(h - l).bit_length()) for l, h in zip(lo_h, hi_h)]
    if sum(bits) > 62:
        return None
    key = torch.zeros_like(cols[0])
    for c, l, b in zip(cols, lo_h, bits):
        key = (key << b) | (c - int(l))
    return key


def _groups(cols):
    """Group ids (rows with equal tuples share one, ids in sorted tuple
    order) and the number of groups."""
    key = _pack(cols)
    if key is None:
        _, inv = torch.unique(torch.stack(cols, 1), dim=0, sorted=True, return_inverse=True)
    else:
        _, inv = torch.unique(key, sorted=True, return_inverse=True)
    G = int(inv.max()) + 1 if inv.numel() else 0
    return inv, G


def _rows_dot(ubis, task, g):
    """h = U.g per row with U = ubis[task[r]] (gs_pbp_rows_dot: each
    component ((u0 g0 + u1 g1) + u2 g2), every operation rounded on its
    own).  ubis (T,3,3) float64, task (R,) int64, g (R,3) float64, all on
    the device; returns (R,3)."""
    ubis, task, g = ubis.contiguous(), task.to(torch.int64).contiguous(), g.contiguous()
    lib = _lib.load()
    R = g.shape[0]
    h = torch.empty((R, 3), dtype=torch.float64, device=g.device)
    _lib.check(lib.gs_pbp_rows_dot(dptr(ubis), dptr(task), R, dptr(g), 0.0, dptr(h), None,
                                   None, stream_ptr()), "pbp_rows_dot")
    return h


def _score(ubis, task, g, tol):
    """(hkl (R,3) int64, ok (R,) bool) of every row: h = ubis[task[r]].g
    rounded half to even, ok = the squared distance to it < tol^2
    (gs_pbp_rows_dot)."""
    ubis, task, g = ubis.contiguous(), task.to(torch.int64).contiguous(), g.contiguous()
    lib = _lib.load()
    R = g.shape[0]
    hkl = torch.empty((R, 3), dtype=torch.int64, device=g.device)
    ok = torch.empty(R, dtype=torch.bool, device=g.device)
    _lib.check(lib.gs_pbp_rows_dot(dptr(ubis), dptr(task), R, dptr(g), float(tol * tol), None,
                                   dptr(hkl), dptr(ok), stream_ptr()), "pbp_rows_dot")
    return hkl, ok


def ubi_fit(design, off, fit, ubis, B0):
    """weighted_lstsq_ubi_fit (:408-419) and the fit-quality test
    (:1595-1597) of every task with fit[t], on the device
    (gs_pbp_ubi_fit): design (rows, 6) = (w g, w hkl) grouped by task with
    CSR offsets off (T+1).  Returns (ubifit, rank, worth, ubi_out, ub0):
    lstsq's solution and rank, the quality test, the fitted UBI where it
    passes (else the input) and U.B0 of the fitted cell."""
    lib = _lib.load()
    T = ubis.shape[0]
    dev = ubis.device
    fit_i = fit.to(torch.int32).contiguous()
    design = design.contiguous()
    ubis = ubis.contiguous()
    ubifit = torch.empty((T, 3, 3), dtype=torch.float64, device=dev)
    ubi_out = torch.empty((T, 3, 3), dtype=torch.float64, device=dev)
    ub0 = torch.empty((T, 3, 3), dtype=torch.float64, device=dev)
    rank = torch.empty(T, dtype=torch.int32, device=dev)
    worth = torch.empty(T, dtype=torch.int32, device=dev)
    b0 = darr(np.asarray(B0, np.float64).reshape(9))
    _lib.check(lib.gs_pbp_ubi_fit(dptr(design), dptr(off), dptr(fit_i), dptr(ubis), b0, T,
                                  dptr(ubifit), dptr(rank), dptr(worth), dptr(ubi_out), dptr(ub0),
                                  stream_ptr()), "pbp_ubi_fit")
    return ubifit, rank, worth.bool(), ubi_out, ub0


def strain_fit(design, off, worth):
    """The strain lstsq of every task with worth[t] (:1652-1656) on the
    device (gs_pbp_strain_fit): design (rows, 7) = (w M, w strain) grouped
    by task.  Returns (T, 6) = (sxx, syy, szz, sxy, sxz, syz), NaN where
    not fitted."""
    lib = _lib.load()
    T = worth.shape[0]
    w_i = worth.to(torch.int32).contiguous()
    design = design.contiguous()
    eps = torch.empty((T, 6), dtype=torch.float64, device=worth.device)
    _lib.check(lib.gs_pbp_strain_fit(dptr(design), dptr(off), dptr(w_i), T, dptr(eps),
                                     stream_ptr()), "pbp_strain_fit")
    return eps


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
    final_ubis = np.full(all_pbpmap_ubis.shape, np.nan)
    final_eps = np.full(all_pbpmap_ubis.shape, np.nan)
    final_npks = np.full(M, np.nan)
    final_nuniq = np.full(M, np.nan)
    ri_col, rj_col = np.asarray(ri_col), np.asarray(rj_col)
    # ---- tasks: (voxel, UBI) pairs of the voxels to refine, grouped by
    # voxel in the order of refine_points (a repeated point is refined once:
    # the reference would write the same results twice)
    rp = np.asarray(refine_points, np.int64).reshape(-1, 2)
    rank = np.full(mask.shape, -1, np.int64)
    rank[rp[::-1, 0], rp[::-1, 1]] = np.arange(len(rp))[::-1]
    sel = np.asarray(mask, bool) & (rank >= 0)
    inb = (ri_col >= 0) & (ri_col < mask.shape[0]) & (rj_col >= 0) & (rj_col < mask.shape[1])
    ri_c, rj_c = np.where(inb, ri_col, 0), np.where(inb, rj_col, 0)
    ok = inb & sel[ri_c, rj_c] & ~np.isnan(all_pbpmap_ubis[0, 0, :])
    m_idx = np.nonzero(ok)[0]
    prank = rank[ri_c[m_idx], rj_c[m_idx]]
    o = np.lexsort((m_idx, prank))
    tmap, tp = m_idx[o], prank[o]
    T = len(tmap)
    if T == 0:
        return final_ubis, final_eps, final_npks, final_nuniq
    upr, counts = np.unique(tp, return_counts=True)
    ptask = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    pts = np.stack([sx_grid[rp[upr, 0], rp[upr, 1]], sy_grid[rp[upr, 0], rp[upr, 1]]], 1)
    tpoint = np.repeat(np.arange(len(upr)), counts)
    ubis_h = np.ascontiguousarray(np.moveaxis(all_pbpmap_ubis[:, :, tmap], -1, 0), np.float64)
    D = lambda a, dt=torch.float64: _t(a, dev, dt)
    sc_d, fc_d, om_d, xpos_d = D(sc), D(fc), D(omega), D(xpos)
    eta_d, si_d, dty_d = D(eta), D(sum_intensity), D(dty)
    dtyi_d = D(dtyi, torch.int64)
    pts_d, ubis_d = D(pts), D(ubis_h)
    tx_d, ty_d = pts_d[D(tpoint, torch.int64), 0], pts_d[D(tpoint, torch.int64), 1]
    # ---- selection and first scoring of every task over every peak
    gv_all = _gve_device(sc_d, fc_d, om_d, xpos_d, geo)
    _, task, peak, hkl = assign_tasks(D(sinomega), D(cosomega), dty_d, gv_all, pts_d,
                                      D(ptask, torch.int32), ubis_d, y0, ystep, tol)
    if peak.numel() == 0:  # no task has any peak: nothing is refined
        return final_ubis, final_eps, final_npks, final_nuniq
    # ---- merge by (task, h, k, l, eta sign, dtyi): groups in sorted key
    # order, sums in row (= ascending peak) order, as np.bincount per voxel
    gid, G = _groups([task, hkl[:, 0], hkl[:, 1], hkl[:, 2],
                      torch.sign(eta_d[peak]).long(), dtyi_d[peak]])
    perm = torch.sort(gid, stable=True).indices
    goff = _offsets(gid, G)
    sums = merge_sums(perm, peak, si_d, (sc_d, fc_d, om_d, dty_d, xpos_d, eta_d), goff)
    sI = sums[:, 0]
    m_sc, m_fc, m_om, m_dty, m_xpos, m_eta = (sums[:, k] / sI for k in range(1, 7))
    gtask = task[perm][goff[:-1]]
    # ---- re-selection of the merged peaks for the voxel, merged g-vectors
    rad = m_om * (np.pi / 180.0)
    yd = torch.abs(y0 - tx_d[gtask] * torch.sin(rad) - ty_d[gtask] * torch.cos(rad) - m_dty)
    keep = torch.nonzero(yd <= ystep).flatten()
    rtask, rydist, reta = gtask[keep], yd[keep], m_eta[keep]
    g = _gve_device(m_sc[keep].contiguous(), m_fc[keep].contiguous(), m_om[keep].contiguous(),
                    m_xpos[keep].contiguous(), geo)
    hi, sel = _score(ubis_d, rtask, g, merge_tol)
    npks = torch.bincount(rtask[sel], minlength=T)
    fit = npks > min_grain_npks
    # ---- weighted least-squares UBI of every task with enough peaks
    rows = torch.nonzero(sel & fit[rtask]).flatten()
    rt = rtask[rows]
    w = (1.0 / (rydist[rows] + 1))[:, None]
    design = torch.cat([w * g[rows], w * hi[rows].double()], 1)
    _, _, worth_d, ubi_out_d, ub0_d = ubi_fit(design, _offsets(rt, T), fit, ubis_d, B0)
    fit_h = fit.cpu().numpy()
    worth = worth_d.cpu().numpy()
    ubi_out = ubi_out_d.cpu().numpy()
    # ---- re-score with the fitted UBIs, count distinct reflections
    hi2, sel2 = _score(ubi_out_d, rtask, g, merge_tol)
    npk2 = torch.bincount(rtask[sel2], minlength=T).cpu().numpy()
    r2 = torch.nonzero(sel2 & fit[rtask]).flatten()
    nuniq = np.zeros(T, np.int64)
    if r2.numel():
        uid, U_ = _groups([rtask[r2], hi2[r2, 0], hi2[r2, 1], hi2[r2, 2],
                           torch.sign(reta[r2]).long()])
        first = torch.zeros(U_, dtype=torch.int64, device=dev).scatter_(0, uid, rtask[r2])
        nuniq = torch.bincount(first, minlength=T).cpu().numpy()
    fidx = np.nonzero(fit_h)[0]
    final_ubis[:, :, tmap[fidx]] = np.moveaxis(ubi_out[fidx], 0, -1)
    final_npks[tmap[fidx]] = npk2[fidx]
    final_nuniq[tmap[fidx]] = nuniq[fidx]
    # ---- strain tensors of the well-fitted tasks
    if worth.any():
        eps = _strain_fit(worth_d, ub0_d, rtask, sel2, g, hi2, rydist, gnoise_std, T)
        widx = np.nonzero(worth)[0]
        final_eps[:, :, tmap[widx]] = np.moveaxis(eps[widx], 0, -1)
    return final_ubis, final_eps, final_npks, final_nuniq


def _strain_fit(worth_d, ub0_d, rtask, sel2, g, hi2, rydist, gnoise_std, T):
    """The directional-strain fit of refine_map (:1624-1668) for every task
    with worth_d at once: per-row terms on the device, per-task mean and
    standard deviation by gs_segment_sum, the weight maximum by
    gs_segment_max, the lstsq solves by
    gs_pbp_strain_fit.  Returns (T, 3, 3) strain tensors (NaN where not
    fitted)."""
    dev = g.device
    rows = torch.nonzero(sel2 & worth_d[rtask]).flatten()
    t = rtask[rows]
    gg = g[rows]
    hk = hi2[rows].double()
    gve0 = _rows_dot(ub0_d, t, hk)
    gTg0 = (gg[:, 0] * gve0[:, 0] + gg[:, 1] * gve0[:, 1]) + gg[:, 2] * gve0[:, 2]
    gTg = (gg[:, 0] * gg[:, 0] + gg[:, 1] * gg[:, 1]) + gg[:, 2] * gg[:, 2]
    strain = gTg0 / gTg - 1
    kap = gg / torch.sqrt(gTg)[:, None]
    kx, ky, kz = kap[:, 0], kap[:, 1], kap[:, 2]
    Mrow = torch.stack((kx * kx, ky * ky, kz * kz, 2 * kx * ky, 2 * kx * kz, 2 * ky * kz), 1)
    w = 1. / (rydist[rows] + 1)
    a = (gve0[:, 0] * (gnoise_std ** 2) * gve0[:, 0] + gve0[:, 1] * (gnoise_std ** 2) *
         gve0[:, 1]) + gve0[:, 2] * (gnoise_std ** 2) * gve0[:, 2]
    noise = torch.sqrt(torch.where(gTg != 0, a / gTg ** 2, torch.ones_like(gTg)))
    w = w * (1. / noise)
    off = _offsets(t, T)
    n = (off[1:] - off[:-1]).double()
    mean = segment_sum(strain[:, None], off)[:, 0] / torch.clamp(n, min=1)
    std = torch.sqrt(segment_sum(((strain - mean[t]) ** 2)[:, None], off)[:, 0] /
                     torch.clamp(n, min=1))
    w = torch.where(strain > mean[t] + std[t] * 3.5, torch.zeros_like(w), w)
    w = torch.where(strain < mean[t] - std[t] * 3.5, torch.zeros_like(w), w)
    wmax = segment_max(w[:, None], off)[:, 0]
    w = w / wmax[t]
    design = torch.cat([w[:, None] * Mrow, (w * strain)[:, None]], 1)
    e = strain_fit(design, off, worth_d).cpu().numpy()
    sxx, syy, szz, sxy, sxz, syz = e.T
    return np.stack([np.stack([sxx, sxy, sxz], -1), np.stack([sxy, syy, syz], -1),
                     np.stack([sxz, syz, szz], -1)], -2)
