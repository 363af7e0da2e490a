"""gs_score_and_refine_dev_noalloc, the one-grain refinement that allocates
nothing (its block partials live in the library's fixed per-GPU scratch
array), against gs_score_and_refine_dev, through the C ABI.

Same kernels and grid, so n, the mean drlv2 and the refined UBI must be
bit-identical to the allocating entry point; two calls in a row, the
second smaller, show that the scratch left by the first call does not leak
into the second.  The oracle's score_and_refine is the reference check of
both.
"""
import ctypes

import numpy as np
import pytest
import torch

from imaged11_amd import _lib, simulate

pytestmark = pytest.mark.gpu

TOL = 0.05


def _T(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def _call(fn_name, ubi, gv, dev):
    """One device call: (n, mean, ubi after) read back after a sync."""
    lib = _lib.load()
    u = _T(ubi, dev)
    g = _T(gv, dev)
    n = torch.full((1,), -7, dtype=torch.int32, device=dev)
    m = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    fn = getattr(lib, fn_name)
    st = fn(_lib.dptr(u), _lib.dptr(g), len(gv), TOL, _lib.dptr(n), _lib.dptr(m),
            _lib.stream_ptr())
    _lib.check(st, fn_name)
    torch.cuda.synchronize()
    return int(n.item()), float(m.item()), u.cpu().numpy()


def _same_bits(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.int64),
                          np.asarray(b, np.float64).view(np.int64))


@pytest.mark.parametrize("sizes", [(300000, 40000), (40000, 900), (900, 0)])
def test_noalloc_equals_allocating_entry_point(dev, orc, sizes):
    print("noalloc: ng =", sizes, flush=True)
    for call, ng in enumerate(sizes):
        n_lat = int(ng * 0.6)
        ubi, gv, _ = simulate.lattice_gvectors(seed=100 + call + ng % 97, n_lattice=n_lat,
                                               n_random=ng - n_lat)
        start = ubi @ (np.eye(3) + 1e-3 * np.random.default_rng(ng).normal(size=(3, 3)))
        want = _call("gs_score_and_refine_dev", start, gv, dev)
        got = _call("gs_score_and_refine_dev_noalloc", start, gv, dev)
        assert got[0] == want[0]
        assert _same_bits(got[1], want[1])
        assert _same_bits(got[2], want[2])
        # and the reference semantics (the oracle's score_and_refine)
        u_o = np.ascontiguousarray(start.copy())
        n_o, m_o = orc.score_and_refine(u_o, gv, TOL)
        assert got[0] == n_o
        if n_o:
            assert abs(got[1] - m_o) <= 1e-12 * abs(m_o)
        else:
            assert got[1] == 0.0
        assert np.abs(got[2] - u_o).max() <= 1e-10 * np.abs(u_o).max()


def test_noalloc_rejects_bad_arguments(dev):
    lib = _lib.load()
    u = torch.eye(3, dtype=torch.float64, device=dev)
    g = torch.zeros((4, 3), dtype=torch.float64, device=dev)
    n = torch.zeros(1, dtype=torch.int32, device=dev)
    m = torch.zeros(1, dtype=torch.float64, device=dev)
    st = lib.gs_score_and_refine_dev_noalloc(_lib.dptr(u), _lib.dptr(g), -1, TOL, _lib.dptr(n),
                                             _lib.dptr(m), _lib.stream_ptr())
    assert st == 1 and "negative" in _lib.last_error()
    st = lib.gs_score_and_refine_dev_noalloc(_lib.dptr(u), ctypes.c_void_p(0), 4, TOL,
                                             _lib.dptr(n), _lib.dptr(m), _lib.stream_ptr())
    assert st == 1 and "null" in _lib.last_error()
    lib.gs_clear_error()
