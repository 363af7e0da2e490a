"""ORACLE — test infrastructure only.

Sequential CPU restatement of point-by-point refinement
(``ImageD11/sinograms/point_by_point.py``): ``refine_map`` (:1432-1670)
voxel by voxel and UBI by UBI, with the helpers it calls —
``get_voxel_idx`` (:359-368), ``count_unique_peaks[_no_dtyi]`` (:284-335),
``merge`` (:338-356), ``compute_gve`` (:372-404),
``weighted_lstsq_ubi_fit`` (:408-419), ``gve_norm``, ``divide_where``,
``ubi_to_unitcell`` (:443-451), ``ubi_and_ucell_to_u`` (:455-490),
``hkluniq`` (:52-81) — in plain numpy (the reference compiles them with
numba; the arithmetic is the same numpy arithmetic).

Only ``tests/`` and the CPU leg of ``bench.py`` import this module.

Pinning: the helpers as the reference's own tests pin them
(``test/test_point_by_point.py``: ``count_unique_peaks`` against
``np.unique(..., axis=1, return_inverse=True)``, ``gve_norm`` against
``np.linalg.norm``, ``divide_where`` against ``np.divide(where=)``, the
geometry against ``transform``), ``compute_gve`` against the compiled
reference's ``compute_xlylzl_xpos_variable`` + ``compute_gv``, ``merge``
against a dictionary grouping, ``weighted_lstsq_ubi_fit`` and the strain
fit against exact (noise-free) lattices with a known strain, and
``ubi_and_ucell_to_u`` against U.B = UBI^-1.  ``refine_map`` as a whole
(which voxels and peaks it selects, the merge keys, the order of its
stages) is checked only through those helpers and through a synthetic map
whose imposed UBIs and strain it recovers to the noise level: beyond that
its parity is unpinned (the reference's numba code cannot run here and
ships no output for a whole map).
"""
import numpy as np

from oracle import refinegrains_oracle as RGO


def get_voxel_idx(y0, xi0, yi0, sinomega, cosomega, dty, ystep):
    """:359-368: peaks whose ray passes the voxel; (indices, ydist)."""
    ydist = np.abs(y0 - xi0 * sinomega - yi0 * cosomega - dty)
    return np.where(ydist <= ystep)[0], ydist


def _rank_of_keys(keys):
    """Rank of each row's key among the sorted distinct keys."""
    order = sorted(range(len(keys)), key=lambda i: keys[i])
    out = np.zeros(len(keys), np.int64)
    seen = {}
    for i in order:
        if keys[i] not in seen:
            seen[keys[i]] = len(seen)
        out[i] = seen[keys[i]]
    return out


def count_unique_peaks(hkl, etasign, dtyi):
    """:284-307."""
    return _rank_of_keys([(hkl[0, i], hkl[1, i], hkl[2, i], etasign[i], dtyi[i])
                          for i in range(hkl.shape[1])])


def count_unique_peaks_no_dtyi(hkl, etasign):
    """:311-334."""
    return _rank_of_keys([(hkl[0, i], hkl[1, i], hkl[2, i], etasign[i])
                          for i in range(hkl.shape[1])])


def merge(hkl, etasign, dtyi, sum_intensity, sc, fc, omega, dty, xpos_refined, eta):
    """:338-356: intensity-weighted means per (h, k, l, etasign, dtyi)."""
    labels = count_unique_peaks(hkl, etasign, dtyi)
    wI = sum_intensity
    sI = np.bincount(labels, weights=wI)
    f = lambda x: np.bincount(labels, weights=x * wI) / sI
    return sI, f(sc), f(fc), f(omega), f(dty), f(xpos_refined), f(eta)


def compute_gve(sc, fc, omega, xpos, distance, y_center, y_size, tilt_y, z_center, z_size,
                tilt_z, tilt_x, o11, o12, o21, o22, t_x, t_y, t_z, wedge, chi, wavelength):
    """:372-404: g-vectors (3, n) with the per-peak distance - xpos."""
    xyz = RGO.compute_xyz_lab([sc, fc], y_center=y_center, y_size=y_size, tilt_y=tilt_y,
                              z_center=z_center, z_size=z_size, tilt_z=tilt_z, tilt_x=tilt_x,
                              distance=distance - xpos, o11=o11, o12=o12, o21=o21, o22=o22)
    origins = RGO.compute_grain_origins(omega, wedge, chi, t_x, t_y, t_z)
    s1 = xyz - origins
    eta = np.degrees(np.arctan2(-s1[1, :], s1[2, :]))
    s1_perp_x = np.sqrt(s1[1, :] * s1[1, :] + s1[2, :] * s1[2, :])
    tth = np.degrees(np.arctan2(s1_perp_x, s1[0, :]))
    return RGO.compute_g_vectors(tth, eta, omega, wavelength, wedge=wedge, chi=chi)


def weighted_lstsq_ubi_fit(ydist, gve, hkl):
    """:408-419."""
    w = (1. / (ydist + 1)).reshape(gve.shape[1], 1)
    a = w * gve.T
    b = w * hkl.T
    m, n = a.shape[-2:]
    rcond = np.finfo(b.dtype).eps * max(n, m)
    ubifitT, residuals, rank, sing_vals = np.linalg.lstsq(a, b, rcond=rcond)
    return w, ubifitT.T, residuals, rank, sing_vals


def gve_norm(gve):
    """:423-429."""
    return np.array([np.sqrt(np.sum(np.power(gve[:, i], 2))) for i in range(gve.shape[1])])


def divide_where(arr1, arr2, out, wherearr):
    """:433-439."""
    with np.errstate(divide="ignore", invalid="ignore"):
        div = np.divide(arr1, arr2)
    return np.where(wherearr != 0, div, out)


def ubi_to_unitcell(ubi):
    """:443-451."""
    G = np.dot(ubi, ubi.T)
    a, b, c = np.sqrt(np.diag(G))
    al = np.degrees(np.arccos(G[1, 2] / b / c))
    be = np.degrees(np.arccos(G[0, 2] / a / c))
    ga = np.degrees(np.arccos(G[0, 1] / a / b))
    return np.array([a, b, c, al, be, ga])


def ubi_and_ucell_to_u(ubi, ucell):
    """:455-490: U = (B.UBI)^T with the Busing-Levy B of ucell."""
    a, b, c = ucell[:3]
    ca, cb, cg = np.cos(np.radians(ucell[3:]))
    g = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca],
                  [a * c * cb, b * c * ca, c * c]])
    gi = np.linalg.inv(g)
    astar, bstar, cstar = np.sqrt(np.diag(gi))
    betas = np.degrees(np.arccos(gi[0, 2] / astar / cstar))
    gammas = np.degrees(np.arccos(gi[0, 1] / astar / bstar))
    B = np.zeros((3, 3))
    B[0, 0] = astar
    B[0, 1] = bstar * np.cos(np.radians(gammas))
    B[0, 2] = cstar * np.cos(np.radians(betas))
    B[1, 1] = bstar * np.sin(np.radians(gammas))
    B[1, 2] = -cstar * np.sin(np.radians(betas)) * ca
    B[2, 2] = 1.0 / c
    return np.dot(B, ubi).T


def hkluniq(ubi, gx, gy, gz, eta, m, tol, hmax):
    """:52-81: (assigned count, distinct (h, k, l, eta > 0) count)."""
    seen = set()
    tcount = 0
    for i in range(len(gx)):
        if not m[i]:
            continue
        H = ubi[0][0] * gx[i] + ubi[0][1] * gy[i] + ubi[0][2] * gz[i]
        K = ubi[1][0] * gx[i] + ubi[1][1] * gy[i] + ubi[1][2] * gz[i]
        L = ubi[2][0] * gx[i] + ubi[2][1] * gy[i] + ubi[2][2] * gz[i]
        ih, ik, il = np.round(H), np.round(K), np.round(L)
        if (H - ih) ** 2 + (K - ik) ** 2 + (L - il) ** 2 < tol * tol:
            if abs(ih) < hmax and abs(ik) < hmax and abs(il) < hmax:
                seen.add((int(ih), int(ik), int(il), 1 if eta[i] > 0 else 0))
            tcount += 1
    return tcount, len(seen)


def _assign(ubi, gve, tol):
    hklf = ubi.dot(gve)
    hkli = np.round(hklf).astype(np.int32)
    err = hklf - hkli
    return hkli, (err ** 2).sum(axis=0) < tol ** 2


def refine_map(refine_points, all_pbpmap_ubis, ri_col, rj_col, sx_grid, sy_grid, mask,
               sc, fc, eta, sum_intensity, sinomega, cosomega, omega, dty, dtyi, xpos,
               ystep, y0, B0, distance, y_center, y_size, tilt_y, z_center, z_size, tilt_z,
               tilt_x, o11, o12, o21, o22, t_x, t_y, t_z, wedge, chi, wavelength,
               tol=0.1, merge_tol=0.05, min_grain_npks=6, gnoise_std=1e-4):
    """:1432-1670, sequentially.  Returns (final_ubis (3,3,M), final_eps
    (3,3,M), final_npks (M,), final_nuniq (M,))."""
    geo = dict(distance=distance, y_center=y_center, y_size=y_size, tilt_y=tilt_y,
               z_center=z_center, z_size=z_size, tilt_z=tilt_z, tilt_x=tilt_x, o11=o11, o12=o12,
               o21=o21, o22=o22, t_x=t_x, t_y=t_y, t_z=t_z, wedge=wedge, chi=chi,
               wavelength=wavelength)
    final_ubis = np.full_like(all_pbpmap_ubis, np.nan)
    final_eps = np.full_like(all_pbpmap_ubis, np.nan)
    final_npks = np.full(all_pbpmap_ubis.shape[2], np.nan)
    final_nuniq = np.full(all_pbpmap_ubis.shape[2], np.nan)
    all_idx = np.arange(len(ri_col))
    gve_all = compute_gve(sc, fc, omega, xpos, **geo)
    for ri, rj in refine_points:
        if not mask[ri, rj]:
            continue
        pbpmap_mask = (ri_col == ri) & (rj_col == rj)
        pbpmap_idx = all_idx[pbpmap_mask]
        ubis_here = all_pbpmap_ubis[:, :, pbpmap_mask]
        if np.all(np.isnan(ubis_here[0, 0, :])):
            continue
        xi0, yi0 = sx_grid[ri, rj], sy_grid[ri, rj]
        local_idx, _ = get_voxel_idx(y0, xi0, yi0, sinomega, cosomega, dty, ystep)
        gve_voxel = gve_all[:, local_idx]
        for ubi_idx in range(ubis_here.shape[2]):
            ubi = ubis_here[:, :, ubi_idx]
            if np.isnan(ubi[0, 0]):
                continue
            hkli, gmask = _assign(ubi, gve_voxel, tol)
            hkl = hkli[:, gmask]
            etasign = np.sign(eta[local_idx][gmask])
            sel = local_idx[gmask]
            (m_si, m_sc, m_fc, m_om, m_dty, m_xpos, m_eta) = merge(
                hkl, etasign, dtyi[sel], sum_intensity[sel], sc[sel], fc[sel], omega[sel],
                dty[sel], xpos[sel], eta[sel])
            m_sin = np.sin(np.radians(m_om))
            m_cos = np.cos(np.radians(m_om))
            lig, ydist = get_voxel_idx(y0, xi0, yi0, m_sin, m_cos, m_dty, ystep)
            gve_merged = compute_gve(m_sc[lig], m_fc[lig], m_om[lig], m_xpos[lig], **geo)
            ydist_grain = ydist[lig]
            m_eta_grain = m_eta[lig]
            hkli, gmask = _assign(ubi, gve_merged, merge_tol)
            grain_npks = np.sum(gmask)
            gve_grain = gve_merged[:, gmask]
            hkl = hkli[:, gmask]
            grain_ydist = ydist_grain[gmask]
            if grain_npks > min_grain_npks:
                w, ubifit, residuals, rank, sing_vals = weighted_lstsq_ubi_fit(grain_ydist,
                                                                               gve_grain, hkl)
                worth = (ubifit is not None) and (rank == 3) and \
                    (np.linalg.cond(ubifit) < 1e14) and (np.linalg.det(ubifit) > 0) and \
                    (np.linalg.matrix_rank(ubifit) == 3)
                ubi_out = ubifit if worth else ubi.copy()
                hkli, gmask = _assign(ubi_out, gve_merged, merge_tol)
                grain_npks = np.sum(gmask)
                labels = count_unique_peaks_no_dtyi(hkli[:, gmask], np.sign(m_eta_grain[gmask]))
                final_ubis[:, :, pbpmap_idx[ubi_idx]] = ubi_out
                final_npks[pbpmap_idx[ubi_idx]] = grain_npks
                final_nuniq[pbpmap_idx[ubi_idx]] = np.unique(labels).shape[0]
                if worth:
                    g = gve_merged[:, gmask]
                    hkl = hkli[:, gmask]
                    yd = ydist_grain[gmask]
                    final_eps[:, :, pbpmap_idx[ubi_idx]] = strain_fit(ubifit, g, hkl, yd, B0,
                                                                      gnoise_std)
    return final_ubis, final_eps, final_npks, final_nuniq


def strain_fit(ubifit, g, hkl, ydist, B0, gnoise_std):
    """The strain-tensor fit of refine_map (:1624-1668)."""
    ucell = ubi_to_unitcell(ubifit)
    U = ubi_and_ucell_to_u(ubifit, ucell)
    gve0 = U.dot(B0).dot(hkl.astype(np.float64))
    gTg0 = np.sum(g * gve0, axis=0)
    gTg = np.sum(g * g, axis=0)
    directional_strain = (gTg0 / gTg) - 1
    kappa = g / gve_norm(g)
    kx, ky, kz = kappa
    M = np.column_stack((kx * kx, ky * ky, kz * kz, 2 * kx * ky, 2 * kx * kz, 2 * ky * kz))
    w = (1. / (ydist + 1)).reshape(g.shape[1], 1)
    a = np.sum(gve0 * (gnoise_std ** 2) * gve0, axis=0)
    strain_noise_std = np.sqrt(divide_where(a, gTg ** 2, out=np.ones_like(gTg), wherearr=gTg))
    w = w * (1. / strain_noise_std.reshape(w.shape))
    mean, std = np.mean(directional_strain), np.std(directional_strain)
    w[directional_strain > mean + std * 3.5] = 0
    w[directional_strain < mean - std * 3.5] = 0
    w = w / np.max(w)
    a = w * M
    b = w.flatten() * directional_strain
    m, n = a.shape[-2:]
    rcond = np.finfo(b.dtype).eps * max(n, m)
    sxx, syy, szz, sxy, sxz, syz = np.linalg.lstsq(a, b, rcond=rcond)[0].T
    return np.array([[sxx, sxy, sxz], [sxy, syy, syz], [sxz, syz, szz]])
