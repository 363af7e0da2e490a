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
