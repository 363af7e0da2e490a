"""Seeded synthetic inputs for the grain-scoring path.

The reference publishes no benchmark inputs for these functions (its only
timed simulation needs the external PolyXSim tool), so the workloads are
generated here from fixed seeds:

* ``lattice_gvectors`` — one grain's g-vectors plus random ones (the
  one-grain ``score_and_refine`` case: a=4.05 A cubic, 12 000 lattice
  g-vectors with N(0, 1e-3) noise + 8 000 uniform in a 1.6 A^-1 ball);
* ``simulate_peaks`` — a forward model that places the diffraction spots of
  many grains on a tilted 2048x2048 detector over an omega scan, with pixel
  jitter and spurious peaks, returning the raw (sc, fc, omega) peak table
  that compute_xlylzl -> compute_gv -> score consume;
* ``refine_lists`` — uneven (lognormal) per-grain peak lists for the batched
  refinement.

The forward model inverts compute_gv (``src/cdiffraction.c:112-168``): for a
sample-frame g it solves the Bragg condition for omega, forms the scattered
beam direction, and intersects it with the detector plane of
compute_xlylzl (``src/cdiffraction.c:190-224``).  Everything is numpy on the
host; it only produces inputs.
"""
import math

import numpy as np

from . import geometry


def random_rotations(n, rng):
    """Uniformly distributed rotation matrices (unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def ub_from_cell(U, a, b, c):
    """UB for an orthogonal cell: g = U . diag(1/a, 1/b, 1/c) . hkl."""
    return U * np.array([1.0 / a, 1.0 / b, 1.0 / c])[None, :]


def lattice_gvectors(seed=1, a=4.05, n_lattice=12000, n_random=8000, gmax=1.6,
                     hmax=8, noise=1e-3):
    """One grain's UBI and a g-vector table (n_lattice + n_random, 3).

    Lattice rows are UB.hkl for integer |h|,|k|,|l| <= hmax with
    |g| <= gmax, cycled to n_lattice rows, each with N(0, noise) added per
    component; random rows are uniform in the ball |g| <= gmax.  The rows
    are shuffled.  Returns (ubi, gv, is_lattice).
    """
    rng = np.random.default_rng(seed)
    U = random_rotations(1, rng)[0]
    ub = ub_from_cell(U, a, a, a)
    r = np.arange(-hmax, hmax + 1)
    hkl = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T
    hkl = hkl[np.any(hkl != 0, axis=1)]
    g = hkl @ ub.T
    g = g[np.linalg.norm(g, axis=1) <= gmax]
    reps = int(math.ceil(n_lattice / len(g)))
    g = np.tile(g, (reps, 1))[:n_lattice]
    g = g + rng.normal(scale=noise, size=g.shape)
    v = rng.normal(size=(n_random, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    v *= gmax * rng.random(n_random)[:, None] ** (1.0 / 3.0)
    gv = np.concatenate([g, v])
    lat = np.concatenate([np.ones(len(g), bool), np.zeros(n_random, bool)])
    perm = rng.permutation(len(gv))
    ubi = np.linalg.inv(ub)
    return np.ascontiguousarray(ubi), np.ascontiguousarray(gv[perm]), lat[perm]


def detector_parameters(rng, npx=2048, pixel=50.0, distance=150000.0, wavelength=0.2,
                        tilt=0.01, wedge=0.0, chi=0.0):
    """An ImageD11 parameter dict for a 2048^2 detector at 50 um, 150 mm,
    tilts U[-tilt, tilt] rad, wavelength 0.2 A (lengths in microns)."""
    return {
        "y_center": npx / 2 + rng.uniform(-5, 5), "z_center": npx / 2 + rng.uniform(-5, 5),
        "y_size": pixel, "z_size": pixel, "distance": distance, "wavelength": wavelength,
        "omegasign": 1.0,
        "tilt_x": rng.uniform(-tilt, tilt), "tilt_y": rng.uniform(-tilt, tilt),
        "tilt_z": rng.uniform(-tilt, tilt),
        "o11": 1.0, "o12": 0.0, "o21": 0.0, "o22": -1.0, "wedge": wedge, "chi": chi,
    }


def random_grains(K, rng, cubic=(3.0, 6.0), ortho=(3.0, 9.0)):
    """K UB matrices: the first half cubic a~U[cubic], the rest orthorhombic
    a,b,c~U[ortho].  Returns (ub[K,3,3], cells[K,3])."""
    U = random_rotations(K, rng)
    cells = np.empty((K, 3))
    nc = K // 2
    cells[:nc] = rng.uniform(*cubic, size=nc)[:, None]
    cells[nc:] = rng.uniform(*ortho, size=(K - nc, 3))
    ub = U / cells[:, None, :]
    return ub, cells


def _grain_spots(ub, t, pars, gmax, omega_range, omega_step, dty_rng=None, dty=0.0):
    """Spots of one grain: (sc, fc, omega, dty) for every hkl within gmax
    whose Bragg solution lands on the detector inside the omega range.
    With dty > 0 each spot is recorded at a sample y-translation drawn
    U[-dty, dty] that shifts the scattering origin by (0, dty_k, 0)."""
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    mat = geometry.wedge_chi_matrix(wedge, chi).reshape(3, 3)
    inv_len = np.linalg.norm(np.linalg.inv(ub), axis=1)  # cell edge lengths
    hm = np.floor(gmax * inv_len).astype(int)
    axes = [np.arange(-m, m + 1) for m in hm]
    hkl = np.array(np.meshgrid(*axes, indexing="ij")).reshape(3, -1).T
    hkl = hkl[np.any(hkl != 0, axis=1)]
    g = hkl @ ub.T
    gg = np.einsum("ij,ij->i", g, g)
    keep = gg <= gmax * gmax
    g, gg = g[keep], gg[keep]
    # k_x(theta) = A cos(theta) + B sin(theta) + m02 g2 must equal -wvln |g|^2 / 2
    A = mat[0, 0] * g[:, 0] + mat[0, 1] * g[:, 1]
    B = mat[0, 1] * g[:, 0] - mat[0, 0] * g[:, 1]
    C = -0.5 * wvln * gg - mat[0, 2] * g[:, 2]
    R = np.hypot(A, B)
    ok = (R > 0) & (np.abs(C) < R)
    g, A, B, C, R = g[ok], A[ok], B[ok], C[ok], R[ok]
    phi = np.arctan2(B, A)
    acos = np.arccos(C / R)
    sols = []
    for sgn in (1.0, -1.0):
        theta = phi + sgn * acos
        om = np.degrees(theta) * sign
        om = (om - omega_range[0]) % 360.0 + omega_range[0]
        inr = (om >= omega_range[0]) & (om < omega_range[1])
        if not np.any(inr):
            continue
        gi = g[inr]
        th = np.radians(om[inr] * sign)
        co, so = np.cos(th), np.sin(th)
        w = np.stack([co * gi[:, 0] - so * gi[:, 1], so * gi[:, 0] + co * gi[:, 1], gi[:, 2]], 1)
        k = w @ mat.T
        nhat = wvln * k
        nhat[:, 0] += 1.0
        u = np.stack([co * t[0] - so * t[1], so * t[0] + co * t[1], np.full_like(co, t[2])], 1)
        o = u @ mat.T
        shift = dty_rng.uniform(-dty, dty, len(o)) if dty > 0 else np.zeros(len(o))
        o[:, 1] += shift
        R3 = r.reshape(3, 3)
        c1, c2 = R3[:, 1], R3[:, 2]
        # o + s nhat = dist + v1 c1 + v2 c2  ->  [nhat, -c1, -c2] (s, v1, v2) = dist - o
        M = np.empty((len(nhat), 3, 3))
        M[:, :, 0] = nhat
        M[:, :, 1] = -c1[None, :]
        M[:, :, 2] = -c2[None, :]
        rhs = dist[None, :] - o
        sol = np.linalg.solve(M, rhs[:, :, None])[:, :, 0]
        good = sol[:, 0] > 0
        fc = sol[:, 1] / p[3] + p[1]
        sc = sol[:, 2] / p[2] + p[0]
        sols.append(np.stack([sc[good], fc[good], om[inr][good], shift[good]], 1))
    if not sols:
        return np.zeros((0, 4))
    return np.concatenate(sols)


def simulate_peaks(ub, translations, pars, n_peaks, seed=2, spurious=0.10, gmax=1.6,
                   omega_range=(0.0, 180.0), omega_step=0.25, jitter_px=0.5, npx=2048, dty=0.0):
    """Forward-simulate a raw peak table of exactly n_peaks rows.

    Spots of every grain that land on the npx x npx detector are collected,
    omega is binned to the centre of its omega_step frame (kept exact when
    omega_step <= 0), (sc, fc) get
    N(0, jitter_px) jitter, the grain spots are subsampled to
    (1 - spurious) * n_peaks and uniform spurious peaks fill the rest.
    Returns dict(sc, fc, omega, dty, grain) with grain = -1 for spurious
    rows; rows are shuffled.  dty > 0: per-peak sample y-translations
    U[-dty, dty] (same length unit as the detector) shift each spot's
    scattering origin; the analysis removes them from yl.
    """
    rng = np.random.default_rng(seed)
    K = ub.shape[0]
    rows, owner = [], []
    for g in range(K):
        s = _grain_spots(ub[g], translations[g], pars, gmax, omega_range, omega_step,
                         dty_rng=rng, dty=dty)
        on = (s[:, 0] >= 0) & (s[:, 0] < npx) & (s[:, 1] >= 0) & (s[:, 1] < npx)
        s = s[on]
        rows.append(s)
        owner.append(np.full(len(s), g, np.int32))
    rows = np.concatenate(rows)
    owner = np.concatenate(owner)
    n_real = int(round(n_peaks * (1.0 - spurious)))
    if len(rows) < n_real:
        raise ValueError("simulation produced %d spots, %d needed" % (len(rows), n_real))
    pick = rng.choice(len(rows), n_real, replace=False)
    rows, owner = rows[pick], owner[pick]
    rows[:, 0] += rng.normal(scale=jitter_px, size=n_real)
    rows[:, 1] += rng.normal(scale=jitter_px, size=n_real)
    n_sp = n_peaks - n_real
    if omega_step > 0:
        nbin = np.floor((rows[:, 2] - omega_range[0]) / omega_step)
        rows[:, 2] = omega_range[0] + (nbin + 0.5) * omega_step
        nfr = int(round((omega_range[1] - omega_range[0]) / omega_step))
        sp_omega = omega_range[0] + (rng.integers(0, nfr, n_sp) + 0.5) * omega_step
    else:  # exact omega, no frames
        sp_omega = rng.uniform(omega_range[0], omega_range[1], n_sp)
    sp = np.stack([rng.uniform(0, npx, n_sp), rng.uniform(0, npx, n_sp), sp_omega,
                   rng.uniform(-dty, dty, n_sp) if dty > 0 else np.zeros(n_sp)], 1)
    allrows = np.concatenate([rows, sp])
    allown = np.concatenate([owner, np.full(n_sp, -1, np.int32)])
    perm = rng.permutation(n_peaks)
    allrows, allown = allrows[perm], allown[perm]
    return {
        "sc": np.ascontiguousarray(allrows[:, 0]),
        "fc": np.ascontiguousarray(allrows[:, 1]),
        "omega": np.ascontiguousarray(allrows[:, 2]),
        "dty": np.ascontiguousarray(allrows[:, 3]),
        "grain": allown,
    }


def scoring_workload(K=1000, N=1_000_000, seed=2, spurious=0.10, translations=False, dty=0.0):
    """The many-grain scoring workload: K grains (half cubic a~U[3,6],
    half orthorhombic a,b,c~U[3,9]), N raw peaks simulated on the detector,
    10% spurious.  Returns dict(pars, ubi[K,3,3], ub, trans[K,3], peaks).
    dty > 0 adds per-peak sample y-translations U[-dty, dty] (scanning)."""
    rng = np.random.default_rng(seed)
    pars = detector_parameters(rng)
    ub, cells = random_grains(K, rng)
    if translations:
        trans = rng.uniform(-200.0, 200.0, size=(K, 3))
        trans[:, 2] *= 0.25
    else:
        trans = np.zeros((K, 3))
    peaks = simulate_peaks(ub, trans, pars, N, seed=seed + 1000, spurious=spurious, dty=dty)
    ubi = np.linalg.inv(ub)
    return {"pars": pars, "ubi": np.ascontiguousarray(ubi), "ub": ub, "trans": trans,
            "cells": cells, "peaks": peaks}


def refine_lists(owner, K, seed=3, median=300, sigma=0.8, lo=20, hi=2000):
    """Per-grain peak lists with lognormal lengths (median, sigma) clipped
    to [lo, hi].  Each list holds the grain's own peaks first (up to its
    length), then random other rows of the table.  Returns CSR (idx, off)."""
    rng = np.random.default_rng(seed)
    n = len(owner)
    lengths = np.clip(np.round(rng.lognormal(math.log(median), sigma, K)), lo, hi).astype(np.int64)
    order = np.argsort(owner, kind="stable")
    starts = np.searchsorted(owner[order], np.arange(K))
    ends = np.searchsorted(owner[order], np.arange(K), side="right")
    idx = np.empty(int(lengths.sum()), np.int32)
    off = np.zeros(K + 1, np.int64)
    off[1:] = np.cumsum(lengths)
    for g in range(K):
        own = order[starts[g]:ends[g]]
        m = int(lengths[g])
        take = own[:m]
        if len(take) < m:
            extra = rng.integers(0, n, m - len(take))
            take = np.concatenate([take, extra])
        idx[off[g]:off[g + 1]] = rng.permutation(take)
    return idx, off


def fast_peaks(ub, translations, pars, n_peaks, seed=5, spurious=0.15, gmax=1.6,
               omega_range=(0.0, 180.0), omega_step=0.25, jitter_px=0.5, npx=2048, dty=0.0,
               candidates=2000, chunk=1000, device="cpu"):
    """Large peak tables quickly: the forward model of `simulate_peaks`,
    vectorised over grains with torch (on `device`), drawing `candidates`
    random hkl per grain from its bounding box instead of enumerating the
    whole reciprocal-lattice ball.  Same outputs: dict(sc, fc, omega, dty,
    grain) numpy arrays of exactly n_peaks rows."""
    import torch
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    f64 = dict(dtype=torch.float64, device=dev)
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    mat = torch.tensor(geometry.wedge_chi_matrix(wedge, chi).reshape(3, 3), **f64)
    R3 = torch.tensor(r.reshape(3, 3), **f64)
    c1, c2 = R3[:, 1], R3[:, 2]
    w = torch.linalg.cross(c1, c2)
    distv = torch.tensor(dist, **f64)
    ub_t = torch.as_tensor(ub, **f64)
    tr_t = torch.as_tensor(np.asarray(translations, float), **f64)
    K = ub_t.shape[0]
    hm = torch.floor(gmax * torch.linalg.norm(torch.linalg.inv(ub_t), dim=2))  # (K, 3)
    out_rows, out_own = [], []
    for g0 in range(0, K, chunk):
        g1 = min(K, g0 + chunk)
        G = g1 - g0
        u = torch.rand((G, candidates, 3), generator=gen, **f64)
        hkl = torch.floor(u * (2 * hm[g0:g1, None, :] + 1)) - hm[g0:g1, None, :]
        key = ((hkl + 512) * torch.tensor([1 << 20, 1 << 10, 1], **f64)).sum(-1)
        gvec = torch.einsum("gij,gcj->gci", ub_t[g0:g1], hkl)
        gg = (gvec * gvec).sum(-1)
        own = torch.arange(g0, g1, device=dev)[:, None].expand(G, candidates)
        keep = (gg <= gmax * gmax) & (hkl.abs().sum(-1) > 0)
        gvec, gg, own, key = gvec[keep], gg[keep], own[keep], key[keep]
        # one copy of each (grain, hkl)
        uk = torch.unique(own.double() * 2.0 ** 31 + key, return_inverse=True)[1]
        first = torch.full((int(uk.max()) + 1 if uk.numel() else 0,), len(uk), device=dev,
                           dtype=torch.long)
        first.scatter_reduce_(0, uk, torch.arange(len(uk), device=dev), reduce="amin")
        gvec, gg, own = gvec[first], gg[first], own[first]
        A = mat[0, 0] * gvec[:, 0] + mat[0, 1] * gvec[:, 1]
        B = mat[0, 1] * gvec[:, 0] - mat[0, 0] * gvec[:, 1]
        C = -0.5 * wvln * gg - mat[0, 2] * gvec[:, 2]
        Rn = torch.hypot(A, B)
        ok = (Rn > 0) & (C.abs() < Rn)
        gvec, A, B, C, Rn, own = gvec[ok], A[ok], B[ok], C[ok], Rn[ok], own[ok]
        phi = torch.atan2(B, A)
        ac = torch.arccos(C / Rn)
        for sgn in (1.0, -1.0):
            om = torch.rad2deg(phi + sgn * ac) * sign
            om = torch.remainder(om - omega_range[0], 360.0) + omega_range[0]
            inr = (om >= omega_range[0]) & (om < omega_range[1])
            gi, omi, owi = gvec[inr], om[inr], own[inr]
            th = torch.deg2rad(omi * sign)
            co, so = torch.cos(th), torch.sin(th)
            wv = torch.stack([co * gi[:, 0] - so * gi[:, 1], so * gi[:, 0] + co * gi[:, 1],
                              gi[:, 2]], 1)
            nh = wvln * (wv @ mat.T)
            nh[:, 0] += 1.0
            t = tr_t[owi]
            uu = torch.stack([co * t[:, 0] - so * t[:, 1], so * t[:, 0] + co * t[:, 1], t[:, 2]],
                             1)
            o = uu @ mat.T
            if dty > 0:
                shift = (torch.rand(len(o), generator=gen, **f64) * 2 - 1) * dty
            else:
                shift = torch.zeros(len(o), **f64)
            o[:, 1] += shift
            b = distv[None, :] - o
            s = (b @ w) / (nh @ w)  # ray parameter to the detector plane
            P = o + s[:, None] * nh - distv[None, :]
            v1 = P @ c1
            v2 = P @ c2
            fc = v1 / p[3] + p[1]
            sc = v2 / p[2] + p[0]
            on = (s > 0) & (sc >= 0) & (sc < npx) & (fc >= 0) & (fc < npx)
            out_rows.append(torch.stack([sc[on], fc[on], omi[on], shift[on]], 1))
            out_own.append(owi[on])
    rows = torch.cat(out_rows)
    owner = torch.cat(out_own).to(torch.int32)
    n_real = int(round(n_peaks * (1.0 - spurious)))
    if len(rows) < n_real:
        raise ValueError("simulation produced %d spots, %d needed" % (len(rows), n_real))
    pick = torch.randperm(len(rows), generator=gen, device=dev)[:n_real]
    rows, owner = rows[pick], owner[pick]
    rows[:, 0] += torch.randn(n_real, generator=gen, **f64) * jitter_px
    rows[:, 1] += torch.randn(n_real, generator=gen, **f64) * jitter_px
    nbin = torch.floor((rows[:, 2] - omega_range[0]) / omega_step)
    rows[:, 2] = omega_range[0] + (nbin + 0.5) * omega_step
    n_sp = n_peaks - n_real
    nfr = int(round((omega_range[1] - omega_range[0]) / omega_step))
    sp = torch.stack([torch.rand(n_sp, generator=gen, **f64) * npx,
                      torch.rand(n_sp, generator=gen, **f64) * npx,
                      omega_range[0] + (torch.randint(0, nfr, (n_sp,), generator=gen,
                                                      device=dev).double() + 0.5) * omega_step,
                      (torch.rand(n_sp, generator=gen, **f64) * 2 - 1) * dty], 1)
    allrows = torch.cat([rows, sp])
    allown = torch.cat([owner, torch.full((n_sp,), -1, dtype=torch.int32, device=dev)])
    perm = torch.randperm(n_peaks, generator=gen, device=dev)
    allrows, allown = allrows[perm].cpu().numpy(), allown[perm].cpu().numpy()
    return {"sc": np.ascontiguousarray(allrows[:, 0]), "fc": np.ascontiguousarray(allrows[:, 1]),
            "omega": np.ascontiguousarray(allrows[:, 2]), "dty": np.ascontiguousarray(allrows[:, 3]),
            "grain": allown}


def scan_workload(K=20000, N=8_000_000, seed=5, spurious=0.15, dty=500.0, device="cpu"):
    """The large scanning workload: K grains (cells as scoring_workload)
    with translations, N peaks with per-peak dty in U[-dty, dty] and 15%
    spurious peaks, generated with `fast_peaks`."""
    rng = np.random.default_rng(seed)
    pars = detector_parameters(rng)
    ub, cells = random_grains(K, rng)
    trans = rng.uniform(-200.0, 200.0, size=(K, 3))
    trans[:, 2] *= 0.25
    peaks = fast_peaks(ub, trans, pars, N, seed=seed + 1000, spurious=spurious, dty=dty,
                       device=device)
    return {"pars": pars, "ubi": np.ascontiguousarray(np.linalg.inv(ub)), "ub": ub,
            "trans": trans, "cells": cells, "peaks": peaks}


def indexing_workload(K=200, cell=(4.05, 4.05, 4.05, 90.0, 90.0, 90.0), symmetry="F",
                      dsmax=1.2, keep=0.5, noise=2e-4, spurious=0.1, seed=7):
    """g-vectors for the ring-pair indexer: K randomly oriented grains of
    one phase, each contributing a random `keep` fraction of its
    reflections with d* < dsmax (g = U.B.hkl + N(0, noise) per component),
    plus `spurious` uniform peaks in the d* shell.  The unit cell is built
    from `cell`/`symmetry` with the reference's B convention.  Returns
    dict(gv (N,3), ds (N,), grain (N,) with -1 for spurious, U (K,3,3),
    cell, symmetry); rows are shuffled."""
    from . import unitcell as uc_mod

    class _NoExt:  # gethkls needs no compiled calls
        pass

    rng = np.random.default_rng(seed)
    uc = uc_mod.unitcell(list(cell), symmetry, ext=_NoExt())
    hkl = np.array([h for _, h in uc.gethkls(dsmax)], float)
    Us = random_rotations(K, rng)
    rows, owner = [], []
    for k in range(K):
        sel = rng.random(len(hkl)) < keep
        g = (Us[k] @ uc.B @ hkl[sel].T).T
        rows.append(g + rng.normal(scale=noise, size=g.shape))
        owner.append(np.full(len(g), k, np.int32))
    g = np.concatenate(rows)
    own = np.concatenate(owner)
    n_sp = int(round(len(g) * spurious / (1.0 - spurious)))
    d = rng.normal(size=(n_sp, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(0.2, dsmax, n_sp)
    g = np.concatenate([g, d * r[:, None]])
    own = np.concatenate([own, np.full(n_sp, -1, np.int32)])
    perm = rng.permutation(len(g))
    g, own = np.ascontiguousarray(g[perm]), own[perm]
    return {"gv": g, "ds": np.sqrt((g * g).sum(axis=1)), "grain": own, "U": Us,
            "cell": tuple(cell), "symmetry": symmetry}


def refinegrains_workload(K=20, N=20000, seed=7, wedge=0.0, chi=0.0, translation=100.0,
                          spurious=0.05, omega_step=0.25, jitter_px=0.3, gmax=1.2):
    """Grains with translations and a binned-omega peak table, the input of
    the refinegrains driver: dict(pars, ubi[K,3,3], trans[K,3], peaks).
    ``pars`` holds every parameter the driver reads (t_x..t_z zero)."""
    rng = np.random.default_rng(seed)
    pars = detector_parameters(rng, wedge=wedge, chi=chi)
    pars.update({"t_x": 0.0, "t_y": 0.0, "t_z": 0.0})
    ub, cells = random_grains(K, rng)
    trans = rng.uniform(-translation, translation, size=(K, 3))
    trans[:, 2] *= 0.25
    peaks = simulate_peaks(ub, trans, pars, N, seed=seed + 1000, spurious=spurious, gmax=gmax,
                           omega_step=omega_step, jitter_px=jitter_px)
    return {"pars": pars, "ubi": np.ascontiguousarray(np.linalg.inv(ub)), "trans": trans,
            "cells": cells, "peaks": peaks}


def _grain_scan_spots(ub, sx, sy, pars, gmax, ystep, y0, omega_range=(0.0, 180.0)):
    """Spots of the voxels (sx, sy) (arrays) of one grain in a scanning
    experiment: at rotation omega a voxel sits at lab (lx, ly) =
    (sx cos - sy sin, sx sin + sy cos); the stage translation dty (on a grid
    of step ystep) that brings it into the beam is the grid value nearest
    y0 - ly, and the scattering origin is (lx, ly + dty - y0, 0).  Returns
    rows (sc, fc, omega, dty, dtyi, xpos=lx, reflection, solution), the
    last two identifying the spot (shared by every voxel of the grain)."""
    p, r, dist = geometry.pack_xlylzl(pars)
    wvln = pars["wavelength"]
    inv_len = np.linalg.norm(np.linalg.inv(ub), axis=1)
    hm = np.floor(gmax * inv_len).astype(int)
    axes = [np.arange(-m, m + 1) for m in hm]
    hkl = np.array(np.meshgrid(*axes, indexing="ij")).reshape(3, -1).T
    hkl = hkl[np.any(hkl != 0, axis=1)]
    g = hkl @ ub.T
    gg = np.einsum("ij,ij->i", g, g)
    keep = gg <= gmax * gmax
    g, gg = g[keep], gg[keep]
    A, B, C = g[:, 0], -g[:, 1], -0.5 * wvln * gg
    R = np.hypot(A, B)
    ok = (R > 0) & (np.abs(C) < R)
    refl = np.nonzero(ok)[0]
    g, A, B, C, R = g[ok], A[ok], B[ok], C[ok], R[ok]
    phi = np.arctan2(B, A)
    acos = np.arccos(C / R)
    R3 = r.reshape(3, 3)
    sx, sy = np.asarray(sx, float)[:, None], np.asarray(sy, float)[:, None]
    rows = []
    for sgn in (1.0, -1.0):
        om = np.degrees(phi + sgn * acos)
        om = (om - omega_range[0]) % 360.0 + omega_range[0]
        inr = (om >= omega_range[0]) & (om < omega_range[1])
        gi, om, ri = g[inr], om[inr], refl[inr]
        th = np.radians(om)
        co, so = np.cos(th), np.sin(th)
        k = np.stack([co * gi[:, 0] - so * gi[:, 1], so * gi[:, 0] + co * gi[:, 1], gi[:, 2]], 1)
        nhat = wvln * k
        nhat[:, 0] += 1.0
        # origin + s nhat = dist + v1 c1 + v2 c2 for every (voxel, spot):
        # one 3x3 matrix per spot, one right-hand side per voxel
        Mi = np.empty((len(nhat), 3, 3))
        Mi[:, :, 0] = nhat
        Mi[:, :, 1] = -R3[:, 1][None, :]
        Mi[:, :, 2] = -R3[:, 2][None, :]
        Mi = np.linalg.inv(Mi)
        lx, ly = sx * co - sy * so, sx * so + sy * co            # (V, S)
        dtyi = np.round((y0 - ly) / ystep)
        dty = dtyi * ystep
        rhs = np.stack([dist[0] - lx, dist[1] - (ly + dty - y0), np.full_like(lx, dist[2])], -1)
        sol = np.einsum("sij,vsj->vsi", Mi, rhs)
        fc = sol[..., 1] / p[3] + p[1]
        sc = sol[..., 2] / p[2] + p[0]
        good = (sol[..., 0] > 0) & (sc >= 0) & (sc < 2048) & (fc >= 0) & (fc < 2048)
        full = lambda a: np.broadcast_to(a, sc.shape)
        rows.append(np.stack([sc, fc, full(om), dty, dtyi, lx, full(ri.astype(float)),
                              np.full(sc.shape, sgn)], -1)[good])
    return np.concatenate(rows) if rows else np.zeros((0, 8))


def pbp_workload(NY=9, ystep=4.0, seed=31, a=4.05, gmax=0.9, strain=1e-3, extra=0.25,
                 grains=2):
    """A point-by-point refinement input: an NY x NY sample grid (step
    ystep) with a disc of voxels shared by cubic grains (two: the left and
    right halves; more: the Voronoi cells of random seeds), every grain but
    the first strained by a symmetric tensor of size ``strain``, and the
    spots of every voxel with the stage position dty that puts it in the
    beam.  Voxels of one grain on the same ray (same reflection, omega and
    dty) make one detector peak, as on a real detector: their spots are
    summed (sum_intensity ~ voxel count) at the mean position, xpos the mean
    origin x.  The per-voxel UBIs to refine are the grain's exact UBI
    rotated by ~1e-3 rad; a fraction ``extra`` of voxels also carry a
    neighbouring grain's UBI, and one carries a NaN entry.  Returns a dict
    with every argument refine_map takes (peak columns, map columns, grids,
    mask, B0 and the detector parameters) plus the truth."""
    rng = np.random.default_rng(seed)
    pars = detector_parameters(rng)
    pars.update(distance=150000.0, wedge=0.0, chi=0.0, t_x=0.0, t_y=0.0, t_z=0.0)
    U = random_rotations(grains, rng)
    B0 = np.eye(3) / a
    epss, ubs = [], []
    for k in range(grains):
        e = rng.normal(scale=strain, size=(3, 3)) if k > 0 else np.zeros((3, 3))
        e = 0.5 * (e + e.T)
        epss.append(e)
        ubs.append(np.linalg.inv(np.eye(3) + e).T @ U[k] @ B0)
    c = NY // 2
    ii, jj = np.meshgrid(np.arange(NY), np.arange(NY), indexing="ij")
    sx_grid = (ii - c) * ystep
    sy_grid = (jj - c) * ystep
    mask = (ii - c) ** 2 + (jj - c) ** 2 <= c * c
    if grains == 2:
        owner = np.where(jj < c, 0, 1)
    else:
        seeds = rng.uniform(-c, c, size=(grains, 2))
        d2 = (ii[..., None] - c - seeds[:, 0]) ** 2 + (jj[..., None] - c - seeds[:, 1]) ** 2
        owner = np.argmin(d2, axis=-1)
    rows = []
    for k in range(grains):
        sel = mask & (owner == k)
        if not sel.any():
            continue
        sp = _grain_scan_spots(ubs[k], sx_grid[sel], sy_grid[sel], pars, gmax, ystep, 0.0)
        rows.append(np.concatenate([sp, np.full((len(sp), 1), k)], 1))
    ubis, ri_col, rj_col = [], [], []
    for i in range(NY):
        for j in range(NY):
            if not mask[i, j]:
                continue
            g = owner[i, j]
            w = rng.normal(scale=1e-3, size=3)
            Wx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
            other = (g + 1) % grains
            for gg in ([g, other] if rng.random() < extra else [g]):
                ubis.append(np.linalg.inv((np.eye(3) + Wx) @ ubs[gg]))
                ri_col.append(i)
                rj_col.append(j)
    ubis.append(np.full((3, 3), np.nan))
    ri_col.append(c)
    rj_col.append(c)
    spots = np.concatenate(rows)
    # one detector peak per (grain, reflection, solution, dtyi)
    _, lab = np.unique(spots[:, [8, 6, 7, 4]], axis=0, return_inverse=True)
    lab = lab.ravel()
    cnt = np.bincount(lab).astype(float)
    mean = lambda k: np.bincount(lab, weights=spots[:, k]) / cnt
    sc, fc, om, dty, dtyi, xpos = (mean(k) for k in range(6))
    n = len(cnt)
    sc = sc + rng.normal(scale=0.05, size=n)
    fc = fc + rng.normal(scale=0.05, size=n)
    dtyi = np.round(dtyi)
    pd, r, dist = geometry.pack_xlylzl(pars)
    v1 = (fc - pd[1]) * pd[3]
    v2 = (sc - pd[0]) * pd[2]
    yl = r[4] * v1 + r[5] * v2 + dist[1]
    zl = r[7] * v1 + r[8] * v2 + dist[2]
    eta = np.degrees(np.arctan2(-yl, zl))
    geo = {k: float(pars[k]) for k in ("distance", "y_center", "y_size", "tilt_y", "z_center",
                                       "z_size", "tilt_z", "tilt_x", "o11", "o12", "o21", "o22",
                                       "t_x", "t_y", "t_z", "wedge", "chi", "wavelength")}
    refine_points = [(i, j) for i in range(NY) for j in range(NY)]
    return {
        "refine_points": np.array(refine_points, np.int64),
        "all_pbpmap_ubis": np.ascontiguousarray(np.moveaxis(np.array(ubis), 0, -1)),
        "ri_col": np.array(ri_col, np.int64), "rj_col": np.array(rj_col, np.int64),
        "sx_grid": sx_grid.astype(float), "sy_grid": sy_grid.astype(float), "mask": mask,
        "sc": sc, "fc": fc, "eta": eta, "sum_intensity": cnt * rng.uniform(0.9, 1.1, n),
        "sinomega": np.sin(np.radians(om)), "cosomega": np.cos(np.radians(om)), "omega": om,
        "dty": dty, "dtyi": dtyi.astype(np.int64), "xpos": xpos, "ystep": ystep, "y0": 0.0,
        "B0": B0, "geometry": geo,
        "truth": {"ubs": ubs, "eps": epss[1] if grains == 2 else epss, "epss": epss,
                  "owner": owner}}
