"""Input generators shared by the CPU and GPU tests."""
import numpy as np

from imaged11_amd import simulate


def boundary_pairs(seed, n, tol, hmax):
    """g-vectors whose drlv2 under `ubi` sits very close to tol^2:
    h = integer + a vector of length tol*(1 + d), |d| log-uniform in
    [1e-17, 1e-9], either sign."""
    rng = np.random.default_rng(seed)
    U = simulate.random_rotations(1, rng)[0]
    ub = simulate.ub_from_cell(U, *rng.uniform(2.5, 12.0, 3))
    ubi = np.ascontiguousarray(np.linalg.inv(ub))
    ints = rng.integers(-hmax, hmax + 1, size=(n, 3)).astype(float)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    d = 10.0 ** rng.uniform(-17, -9, n) * rng.choice([-1.0, 1.0], n)
    h = ints + v * (tol * (1.0 + d))[:, None]
    gv = np.ascontiguousarray(h @ ub.T)
    return ubi, gv


def row0_boundary_pairs(seed, n, tol, hmax):
    """g-vectors whose row 0 alone sits at the tolerance edge: h = integer +
    (+-tol*(1 + d), 0, 0), |d| log-uniform in [1e-17, 1e-9], either sign,
    so the sieve's row-0 cut-off (gs_sieve_radius) is tested at its edge;
    large hmax makes the rounding of h0 (the radius margin) matter."""
    rng = np.random.default_rng(seed)
    U = simulate.random_rotations(1, rng)[0]
    ub = simulate.ub_from_cell(U, *rng.uniform(2.5, 12.0, 3))
    ubi = np.ascontiguousarray(np.linalg.inv(ub))
    ints = rng.integers(-hmax, hmax + 1, size=(n, 3)).astype(float)
    d = 10.0 ** rng.uniform(-17, -9, n) * rng.choice([-1.0, 1.0], n)
    off = tol * (1.0 + d) * rng.choice([-1.0, 1.0], n)
    h = ints.copy()
    h[:, 0] += off
    gv = np.ascontiguousarray(h @ ub.T)
    return ubi, gv


def scan_boundary_pairs(seed, n, tol, hmax, wedge=1.0, chi=-2.0, wvln=0.2):
    """Single (peak, grain) pairs for the fused compute_gv + assign path
    whose drlv2 lies very close to tol^2: a peak position xl on a ray from
    the grain origin, its g-vector G from the compute_gv formula, and a
    per-row UBI built so that UBI.G = integer + a vector of length
    tol*(1 + d), |d| log-uniform in [1e-17, 1e-6]."""
    from imaged11_amd import geometry
    rng = np.random.default_rng(seed)
    mat = geometry.wedge_chi_matrix(wedge, chi).reshape(3, 3)
    om = rng.uniform(0, 2 * np.pi, n)
    so, co = np.sin(om), np.cos(om)
    t = rng.uniform(-500, 500, size=(n, 3))
    u = np.stack([co * t[:, 0] - so * t[:, 1], so * t[:, 0] + co * t[:, 1], t[:, 2]], 1)
    o = u @ mat.T
    nh = rng.normal(size=(n, 3))
    nh[:, 0] = np.abs(nh[:, 0]) + 2.0
    nh /= np.linalg.norm(nh, axis=1)[:, None]
    xl = o + nh * rng.uniform(1.2e5, 1.8e5, n)[:, None]
    d = xl - o
    k = (d / np.linalg.norm(d, axis=1)[:, None] - np.array([1.0, 0, 0])) / wvln
    v = k @ mat
    G = np.stack([co * v[:, 0] + so * v[:, 1], -so * v[:, 0] + co * v[:, 1], v[:, 2]], 1)
    U0 = simulate.random_rotations(n, rng) * rng.uniform(3.0, 9.0, size=(n, 1, 3))
    h0 = np.einsum("nij,nj->ni", U0, G)
    ints = np.rint(h0) + rng.integers(-hmax, hmax + 1, size=(n, 3))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    dd = 10.0 ** rng.uniform(-17, -6, n) * rng.choice([-1.0, 1.0], n)
    target = ints + dirs * (tol * (1.0 + dd))[:, None]
    corr = (target - h0)[:, :, None] * G[:, None, :] / (G * G).sum(1)[:, None, None]
    ubis = np.ascontiguousarray(U0 + corr)
    return (np.ascontiguousarray(xl), np.ascontiguousarray(so), np.ascontiguousarray(co), ubis,
            np.ascontiguousarray(t), np.ascontiguousarray(mat.ravel()))
