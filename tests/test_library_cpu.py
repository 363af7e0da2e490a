"""CPU-side checks of the product library: the C ABI loads and exports
every declared symbol, the host logic of the drop-in (f2py argument rules),
the accuracy of the device sincos algorithm (its host build), the
simulator's invariants.  No kernel is launched."""
import math
import os

import mpmath
import numpy as np
import pytest
import torch

from imaged11_amd import _lib, geometry, simulate
from imaged11_amd import cImageD11 as gpu

from helpers import boundary_pairs


def test_exports_every_header_symbol():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert {"score", "score_and_refine", "score_and_assign", "refine_assigned", "compute_gv",
            "compute_xlylzl", "verify_rounding", "gs_score_many", "gs_refine_lists"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, "no ctypes signature for %s" % n
    assert lib.gs_version() == 1


def test_single_hip_runtime():
    _lib.load()
    assert len(_lib.hip_runtimes_loaded()) == 1


def _host_sincos(x):
    x = np.ascontiguousarray(x, np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    _lib.load().gs_selftest_sincos_host(_lib.hptr(x), _lib.hptr(s), _lib.hptr(c), len(x))
    return s, c


def test_device_sincos_correctly_rounded():
    """The device sincos algorithm (compiled for the host) returns the
    correctly rounded sin and cos, checked against 200-bit mpmath, on
    omega grids, random angles and awkward arguments."""
    mpmath.mp.prec = 200
    rng = np.random.default_rng(1)
    omega = np.concatenate([(np.arange(720) + 0.5) * 0.25, np.arange(-360, 361, 15.0),
                            rng.uniform(-720, 720, 1500)])
    x = np.concatenate([np.pi / 180.0 * omega,
                        [1e-300, -1e-300, 5e-324, 1e-8, np.pi / 2, np.pi, 3 * np.pi / 2, 1e5,
                         -123456.789, 800000.0]])
    s, c = _host_sincos(x)
    for xi, si, ci in zip(x, s, c):
        m = mpmath.mpf(float(xi))
        assert si == float(mpmath.sin(m)), xi
        assert ci == float(mpmath.cos(m)), xi


def test_device_sincos_special():
    s, c = _host_sincos(np.array([0.0, -0.0, np.inf, np.nan]))
    assert s[0] == 0.0 and not math.copysign(1, s[0]) < 0 and c[0] == 1.0
    assert s[1] == 0.0 and math.copysign(1, s[1]) < 0 and c[1] == 1.0
    assert np.isnan(s[2]) and np.isnan(c[2]) and np.isnan(s[3])


def test_device_sincos_vs_glibc_sincos(orc):
    """glibc's sincos (what the reference's compute_gv calls) is not
    correctly rounded everywhere: the two agree except in the last bit of a
    small fraction of arguments."""
    rng = np.random.default_rng(2)
    omega = rng.uniform(-360, 360, 200000)
    so, co = orc.sincos_omega(omega, 1.0)
    s, c = _host_sincos(np.pi / 180.0 * omega * 1.0)
    diff = (s != so) | (c != co)
    assert diff.mean() < 0.01
    assert np.abs(s - so).max() <= 2.0 ** -53 and np.abs(c - co).max() <= 2.0 ** -53


def _same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


@pytest.mark.parametrize("n", [0, 1, 1000, 300001])
def test_parity_trig_is_glibc_sincos(orc, monkeypatch, n):
    """The parity mode's host trig (libgscore's own code, no oracle) equals
    the oracle's glibc sincos of the reference expression bit for bit, the
    oracle's being pinned against the compiled reference's compute_gv; with
    several host threads (300001 peaks over up to 8 threads) and with one."""
    rng = np.random.default_rng(n)
    omega = rng.uniform(-720, 720, n)
    omega[: min(n, 4)] = [0.0, -0.0, 90.0, 180.0][: min(n, 4)]
    lib = _lib.load()
    for threads in ("8", "1"):
        monkeypatch.setenv("OMP_NUM_THREADS", threads)
        for sign in (1.0, -1.0):
            so = np.full(n, np.nan)
            co = np.full(n, np.nan)
            _lib.check(lib.gs_sincos_omega_host(_lib.hptr(omega), sign, _lib.hptr(so),
                                                _lib.hptr(co), n), "sincos_omega_host")
            want_s, want_c = orc.sincos_omega(omega, sign)
            assert _same_bits(so, want_s) and _same_bits(co, want_c)


def test_parity_wedge_chi_matrix_is_the_references(orc):
    lib = _lib.load()
    for wedge, chi in [(0.0, 0.0), (0.3, -0.7), (-2.5, 11.0), (45.0, 90.0), (1e-9, 179.99)]:
        m = np.zeros(9)
        _lib.check(lib.gs_wedge_chi_matrix_host(wedge, chi, _lib.hptr(m)), "wedge_chi")
        assert _same_bits(m, orc.wedge_chi_matrix(wedge, chi).ravel())


def test_parity_trig_rejects_bad_arguments():
    lib = _lib.load()
    x = np.zeros(4)
    assert lib.gs_sincos_omega_host(_lib.hptr(x), 1.0, None, _lib.hptr(x), 4) == 1
    assert lib.gs_sincos_omega_host(_lib.hptr(x), 1.0, _lib.hptr(x), _lib.hptr(x), -1) == 1
    assert lib.gs_wedge_chi_matrix_host(0.0, 0.0, None) == 1
    lib.gs_clear_error()


def test_dropin_argument_rules():
    """f2py rules of src/_cImageD11.pyf: intent(inout) arrays must have the
    exact dtype and be C-contiguous; shapes are checked; checked before any
    device work."""
    gv = np.zeros((10, 3))
    with pytest.raises(ValueError):
        gpu.score_and_refine(np.eye(3, dtype=np.float32), gv, 0.1)
    with pytest.raises(ValueError):
        gpu.score_and_refine(np.asfortranarray(np.arange(9.0).reshape(3, 3)), gv, 0.1)
    with pytest.raises(ValueError):
        gpu.score_and_refine([[1, 0, 0], [0, 1, 0], [0, 0, 1]], gv, 0.1)
    with pytest.raises(ValueError):
        gpu.score(np.eye(3), np.zeros((10, 2)), 0.1)
    with pytest.raises(ValueError):
        gpu.score(np.eye(2), gv, 0.1)
    with pytest.raises(ValueError):
        gpu.score_and_assign(np.eye(3), gv, 0.1, np.ones(10, np.float32), np.zeros(10, np.int32), 1)
    with pytest.raises(ValueError):
        gpu.score_and_assign(np.eye(3), gv, 0.1, np.ones(10), np.zeros(10, np.int64), 1)
    with pytest.raises(ValueError):
        gpu.score_and_assign(np.eye(3), gv, 0.1, np.ones(9), np.zeros(10, np.int32), 1)
    with pytest.raises(ValueError):
        gpu.compute_gv(gv, np.zeros(10), 1, 0.2, 0, 0, (0, 0, 0), np.zeros((10, 3), np.float32))
    with pytest.raises(ValueError):
        gpu.compute_gv(gv, np.zeros(10), 1, 0.2, 0, 0, (0, 0), np.zeros((10, 3)))
    with pytest.raises(ValueError):
        gpu.compute_xlylzl(np.zeros(5), np.zeros(5), np.zeros(4), np.eye(3), np.zeros(3),
                           np.zeros((5, 3))[:, ::-1])
    with pytest.raises(ValueError):
        gpu.compute_xlylzl(np.zeros(5), np.zeros(4), np.zeros(4), np.eye(3), np.zeros(3),
                           np.zeros((5, 3)))


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure")
def test_dropin_fails_loudly_without_gpu():
    with pytest.raises(_lib.GscoreError):
        gpu.score(np.eye(3), np.zeros((4, 3)), 0.1)
    with pytest.raises(_lib.GscoreError):
        gpu.verify_rounding(20)


def test_omp_thread_knob():
    gpu.cimaged11_omp_set_num_threads(3)
    assert gpu.cimaged11_omp_get_max_threads() == 3
    gpu.cimaged11_omp_set_num_threads(1)


def test_geometry_packing_matches_ctransform_formula():
    pars = {"z_center": 980.0, "y_center": 1010.0, "z_size": 48.0, "y_size": 49.0,
            "distance": 100100.0, "tilt_x": 0.01, "tilt_y": -0.02, "tilt_z": 0.03,
            "o11": 1, "o12": 0, "o21": 0, "o22": -1}
    p, r, dist = geometry.pack_xlylzl(pars)
    assert p.tolist() == [980.0, 1010.0, 48.0, 49.0] and dist.tolist() == [100100.0, 0, 0]
    R = r.reshape(3, 3)
    np.testing.assert_allclose(R @ R.T * np.diag([1, 1, 1]), np.diag([1, 1, 1]), atol=1e-15)
    # Rz first, then Ry, then Rx: a detector vector along y picks up tilt_z first
    rz = geometry.detector_rotation_matrix(0, 0, 0.03)
    assert rz[1, 0] > 0 and rz[0, 1] < 0
    m = geometry.wedge_chi_matrix(3.0, 4.0).reshape(3, 3)
    np.testing.assert_allclose(m @ m.T, np.eye(3), atol=1e-15)


def test_simulator_invariants():
    wl = simulate.scoring_workload(K=10, N=5000, seed=3)
    pk = wl["peaks"]
    assert len(pk["sc"]) == 5000 and (pk["grain"] == -1).sum() == 500
    assert set(np.unique(pk["omega"] / 0.25 - 0.5)) <= set(np.arange(720.0))
    idx, off = simulate.refine_lists(pk["grain"], 10, median=40, lo=20, hi=80)
    L = np.diff(off)
    assert L.min() >= 20 and L.max() <= 80 and off[-1] == len(idx)
    ubi, gv, lat = simulate.lattice_gvectors(seed=1)
    assert gv.shape == (20000, 3) and lat.sum() == 12000


@pytest.mark.parametrize("form", ["fma", "fract"])
@pytest.mark.parametrize("tol,hmax", [(0.05, 3), (0.05, 40), (0.01, 10), (0.3, 200),
                                      (0.5, 5), (0.8, 20), (1e-4, 2)])
def test_certified_decision_is_sound(tol, hmax, form):
    """Whenever the certified path decides, it decides as the reference
    arithmetic does; undecided pairs are the rare exception.  fma: the
    assign kernel's form; fract: the score kernel's (drlv2 - 3/4)."""
    lib = _lib.load()
    fn = lib.gs_selftest_certified_host if form == "fma" else lib.gs_selftest_certified_fract_host
    total_amb = 0
    for seed in range(4):
        ubi, gv = boundary_pairs(seed, 40000, tol, hmax)
        gmax = np.abs(gv).max()
        out = np.zeros(len(gv), np.int32)
        fn(_lib.hptr(ubi), _lib.hptr(gv), len(gv), tol, gmax, _lib.hptr(out))
        dec, exact = out & 3, (out & 4) > 0
        decided = dec != 2
        assert np.array_equal(dec[decided] == 1, exact[decided])
        # the cases straddle tol^2 and the window is exercised; for
        # tol >= 0.01 some near-boundary pairs are decided as well
        assert 0 < exact.sum() < len(gv)
        assert (~decided).sum() > 0
        assert tol < 0.01 or decided.sum() > 0
        total_amb += int((~decided).sum())


@pytest.mark.parametrize("form", ["fma", "fract"])
def test_certified_window_special(form):
    lib = _lib.load()
    fn = lib.gs_selftest_certified_host if form == "fma" else lib.gs_selftest_certified_fract_host
    ubi = np.eye(3)
    gv = np.array([[np.nan, 0, 0], [0.5, 0.5, 0.5], [2.0 ** 51 + 0.5, 0, 0], [0.01, 0, 0]])
    out = np.zeros(4, np.int32)
    # huge gmax -> everything undecided (exact path)
    fn(_lib.hptr(ubi), _lib.hptr(gv), 4, 0.1, 2.0 ** 52, _lib.hptr(out))
    assert (out[1:] & 3 == 2).all() and out[0] == 0
    fn(_lib.hptr(ubi), _lib.hptr(gv), 4, 0.1, np.inf, _lib.hptr(out))
    assert (out[1:] & 3 == 2).all()
    # ties at half integers: (h - round h)^2 = 1/4 either way
    g = np.array([[0.5, 0.0, 0.0], [1.5, 2.5, -0.5], [0.25, 0.0, 0.0]])
    out = np.zeros(3, np.int32)
    fn(_lib.hptr(ubi), _lib.hptr(g), 3, 0.5 + 1e-9, 10.0, _lib.hptr(out))
    assert out[0] == 1 + 4 and out[2] == 1 + 4 and out[1] & 4 == 0


@pytest.mark.parametrize("tol,hmax", [(0.05, 2), (0.05, 30), (0.25, 10), (0.01, 3)])
def test_scan_fast_path_is_sound(tol, hmax):
    """The fused compute_gv + assign fast path only ever skips pairs that
    the reference arithmetic also puts outside the tolerance."""
    from helpers import scan_boundary_pairs
    lib = _lib.load()
    for seed in range(3):
        xl, so, co, ubis, t, mat = scan_boundary_pairs(seed, 30000, tol, hmax)
        out = np.zeros(len(xl), np.int32)
        lib.gs_selftest_scan_host(_lib.hptr(xl), _lib.hptr(so), _lib.hptr(co), _lib.hptr(ubis),
                                  _lib.hptr(t), len(xl), _lib.hptr(mat), 0.2, tol,
                                  _lib.hptr(out))
        skipped = (out & 3) == 1
        exact_in = (out & 4) > 0
        assert not np.any(skipped & exact_in)
        assert 0 < exact_in.sum() < len(xl)
        assert skipped.sum() > 0 and (~skipped).sum() > 0


def test_smoke_workload_generates():
    """smoke()'s input builds on the CPU (the simulator raises when the
    grains give fewer spots than asked for)."""
    import __graft_entry__ as g
    from imaged11_amd import simulate
    wl = simulate.scoring_workload(**g.SMOKE_WORKLOAD)
    assert len(wl["peaks"]["sc"]) == g.SMOKE_WORKLOAD["N"]


def test_gauss_newton_ubis_closed_form():
    """The cofactor solve of the per-grain Gauss-Newton systems (the
    device function of gs_gauss_newton_ubis, built for the host) against
    numpy's solver; a singular grain keeps its UBI, or gets NaN without
    one."""
    lib = _lib.load()
    rng = np.random.default_rng(8)
    K = 50
    G = rng.normal(size=(K, 40, 3))
    h = rng.normal(size=(K, 40, 3))
    A = np.einsum("kni,knj->kij", h, G)
    M = np.einsum("kni,knj->kij", G, G)
    M[3] = np.outer([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])  # singular
    sums = np.concatenate([A.reshape(K, 9), M[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]], 1)
    prev = rng.normal(size=(K, 3, 3))
    got = np.empty((K, 3, 3))
    lib.gs_selftest_gn_ubis_host(_lib.hptr(sums), _lib.hptr(prev), K, _lib.hptr(got))
    nan = np.empty((K, 3, 3))
    lib.gs_selftest_gn_ubis_host(_lib.hptr(sums), None, K, _lib.hptr(nan))
    assert np.all(np.isnan(nan[3])) and np.array_equal(nan[:3], got[:3])
    for k in range(K):
        if k == 3:
            assert np.array_equal(got[k], prev[k])
        else:
            want = np.linalg.solve(M[k], A[k].T).T
            np.testing.assert_allclose(got[k], want, rtol=1e-11, atol=1e-12)


def test_pack_geometry_host_matches_autograd():
    """The closed-form packing and Jacobian of pack_geometry_host against
    pack_geometry's values and torch autograd's derivatives of them, at
    random geometries (wedge and chi non-zero), both trig modes; the
    values against geometry.pack_xlylzl and the wedge/chi matrix too."""
    import torch
    from imaged11_amd import adjoint, geometry
    rng = np.random.default_rng(21)
    names = list(adjoint.FREE) + ["z_size", "y_size"]
    for trial in range(6):
        fixed = geometry.full_parameters({"o11": -1.0 if trial % 2 else 1.0, "o22": -1.0})
        vals = {"distance": rng.uniform(5e3, 2e5), "y_center": rng.uniform(500, 1500),
                "z_center": rng.uniform(500, 1500), "tilt_x": rng.normal(scale=0.02),
                "tilt_y": rng.normal(scale=0.02), "tilt_z": rng.normal(scale=0.02),
                "wavelength": rng.uniform(0.1, 0.3), "wedge": rng.normal(scale=2.0),
                "chi": rng.normal(scale=2.0), "z_size": rng.uniform(40, 60),
                "y_size": rng.uniform(40, 60)}
        parity = trial >= 3
        geo, jac = adjoint.pack_geometry_host(vals, fixed, parity=parity)
        assert jac.shape == (adjoint.GEO, len(names))

        def f(*xs):
            ps = {k: x for k, x in zip(names, xs)}
            return adjoint.pack_geometry(ps, fixed, parity=parity)
        xs = tuple(torch.tensor(vals[k], dtype=torch.float64) for k in names)
        want = f(*xs).numpy()
        np.testing.assert_allclose(geo, want, rtol=0, atol=1e-15)
        q = dict(fixed, **vals)
        p, r, dist = geometry.pack_xlylzl(q)
        np.testing.assert_array_equal(geo[:4], p)
        np.testing.assert_allclose(geo[4:13], r, rtol=0, atol=1e-15)
        np.testing.assert_array_equal(geo[13:16], dist)
        assert geo[16] == vals["wavelength"]
        np.testing.assert_allclose(geo[17:], geometry.wedge_chi_matrix(vals["wedge"], vals["chi"]),
                                   rtol=0, atol=1e-15)
        J = torch.autograd.functional.jacobian(f, xs)
        J = np.stack([j.numpy() for j in J], 1)
        np.testing.assert_allclose(jac, J, rtol=1e-13, atol=1e-16)
        # and against central differences of the host packing itself
        for j, k in enumerate(names):
            h = 1e-6 * max(abs(vals[k]), 1.0)
            up = adjoint.pack_geometry_host(dict(vals, **{k: vals[k] + h}), fixed)[0]
            dn = adjoint.pack_geometry_host(dict(vals, **{k: vals[k] - h}), fixed)[0]
            np.testing.assert_allclose(jac[:, j], (up - dn) / (2 * h), rtol=0, atol=1e-8)
    import pytest
    with pytest.raises(ValueError):
        adjoint.pack_geometry_host({"omegasign": 1.0}, {})


def test_adam_matches_torch():
    """The host Adam of refine_joint against torch.optim.Adam over 40 steps
    of random gradients (8 parameters)."""
    import torch
    from imaged11_amd import adjoint
    rng = np.random.default_rng(22)
    x0 = rng.normal(size=8)
    grads = rng.normal(size=(40, 8)) * np.logspace(-6, 3, 8)
    opt = adjoint.Adam(8, 0.05)
    x = x0.copy()
    t = torch.tensor(x0, requires_grad=True)
    topt = torch.optim.Adam([t], lr=0.05, foreach=False)
    for g in grads:
        x = opt.step(x, g)
        t.grad = torch.tensor(g)
        topt.step()
        np.testing.assert_allclose(x, t.detach().numpy(), rtol=1e-14, atol=1e-15)


def test_ctypes_signatures_match_the_header():
    """Every ctypes declaration has as many arguments as the prototype in
    include/gscore.h (a mismatch would only show when the call is made)."""
    import re
    src = re.sub(r"/\*.*?\*/", "", open(_lib.HEADER).read(), flags=re.S)
    arity = {}
    for m in re.finditer(r"\b(?:int|void|double|char|long long)\s*\*?\s*(\w+)\s*\(([^;]*?)\)\s*;", src,
                         re.S):
        params = m.group(2).strip()
        arity[m.group(1)] = 0 if params in ("", "void") else len(params.split(","))
    for name, (_, args) in _lib.SIGNATURES.items():
        assert name in arity, name
        assert len(args) == arity[name], (name, len(args), arity[name])


def test_mode_constants_match_the_header():
    """The score-mode values the Python layer passes are the header's."""
    import re
    text = open(os.path.join(os.path.dirname(__file__), "..", "include", "gscore.h")).read()
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define GS_SCORE_(\w+) (\d+)", text)}
    assert defs == {"CERTIFIED": _lib.SCORE_CERTIFIED, "EXACT": _lib.SCORE_EXACT,
                    "SIEVE": _lib.SCORE_SIEVE}
    from imaged11_amd import batched
    assert batched.score_mode(0.05) == _lib.SCORE_SIEVE
    assert batched.score_mode(-0.1) == _lib.SCORE_SIEVE
    assert batched.score_mode(0.2) == _lib.SCORE_CERTIFIED


def _schedule(off, n_idx, piece_len):
    """gs_joint_schedule through ctypes on host arrays: (piece_off,
    piece_grain, slots)."""
    import ctypes
    lib = _lib.load()
    lib.gs_clear_error()
    off = np.ascontiguousarray(off, np.int64)
    K = len(off) - 1
    po = np.zeros(K + 1, np.int64)
    n = ctypes.c_longlong(-1)
    _lib.check(lib.gs_joint_schedule(_lib.hptr(off), K, n_idx, piece_len, _lib.hptr(po), None,
                                     None, 0, ctypes.byref(n)), "schedule query")
    grain = np.full(max(n.value, 1), -7, np.int32)
    slots = np.full(max(n.value, 1), -7, np.int32)
    _lib.check(lib.gs_joint_schedule(_lib.hptr(off), K, n_idx, piece_len, _lib.hptr(po),
                                     _lib.hptr(grain), _lib.hptr(slots), n.value, ctypes.byref(n)),
               "schedule")
    return po, grain[:n.value], slots[:n.value]


@pytest.mark.parametrize("piece_len", [1, 7, 256])
def test_joint_schedule_covers_every_list(piece_len):
    """The pieces of the adjoint pass: each grain's clipped list is covered
    by consecutive pieces of at most piece_len entries, in grain order, and
    the launch order is a permutation sorted by length (longest first, ties
    in piece order).  Lists with bounds outside [0, n_idx] are clipped the
    way the kernels clip them; empty lists own no piece."""
    rng = np.random.default_rng(8)
    lengths = rng.integers(0, 3 * piece_len + 2, 40)
    lengths[[0, 5, 6]] = 0
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n_idx = int(off[-1])
    off[-1] = n_idx + 10  # past the end: clipped to n_idx
    off[-2] = min(off[-2], off[-3])  # a list shorter than nothing: empty
    po, grain, slots = _schedule(off, n_idx, piece_len)
    K = len(off) - 1
    o1 = np.clip(off[1:], 0, n_idx)
    o0 = np.minimum(np.clip(off[:-1], 0, None), o1)
    ln = o1 - o0
    assert po[0] == 0 and np.array_equal(np.diff(po), -(-ln // piece_len))
    assert np.array_equal(grain, np.repeat(np.arange(K), np.diff(po)))
    # the pieces of grain g tile [o0, o1) in order
    plen = np.empty(len(grain), np.int64)
    for g in range(K):
        ps = np.arange(po[g], po[g + 1])
        q0 = o0[g] + (ps - po[g]) * piece_len
        plen[ps] = np.minimum(piece_len, o1[g] - q0)
        assert plen[ps].sum() == ln[g] and (plen[ps] >= 1).all()
    assert np.array_equal(np.sort(slots), np.arange(len(grain)))
    want = sorted(range(len(grain)), key=lambda p: (-plen[p], p))
    assert np.array_equal(slots, want)


def test_joint_schedule_edges():
    lib = _lib.load()
    # a piece length far beyond every list: one piece per non-empty list,
    # and the sort's buckets sized by the longest list, not piece_len
    po, grain, slots = _schedule(np.array([0, 5, 5, 12], np.int64), 12, 2**31 - 1)
    assert np.array_equal(po, [0, 1, 1, 2]) and np.array_equal(grain, [0, 2])
    assert np.array_equal(slots, [1, 0])
    po, grain, slots = _schedule(np.zeros(1, np.int64), 0, 16)  # K = 0
    assert np.array_equal(po, [0]) and len(grain) == 0
    po, grain, slots = _schedule(np.array([0, 0, 0]), 0, 16)  # empty lists
    assert np.array_equal(po, [0, 0, 0]) and len(slots) == 0
    import ctypes
    off = np.array([0, 40], np.int64)
    po = np.zeros(2, np.int64)
    g = np.zeros(8, np.int32)
    n = ctypes.c_longlong(0)
    lib.gs_clear_error()
    assert lib.gs_joint_schedule(_lib.hptr(off), 1, 40, 4, _lib.hptr(po), _lib.hptr(g),
                                 _lib.hptr(g), 8, ctypes.byref(n)) != 0
    assert n.value == 10 and b"capacity" in lib.gs_last_error()
    lib.gs_clear_error()
    assert lib.gs_joint_schedule(_lib.hptr(off), 1, 40, 0, _lib.hptr(po), None, None, 0,
                                 ctypes.byref(n)) != 0
    lib.gs_clear_error()
    # workspace: per piece 48 doubles, per block of the pass 20
    assert lib.gs_joint_workspace_doubles(0) == 0
    assert lib.gs_joint_workspace_doubles(10) == 10 * 48 + 3 * 20


def test_device_sincos_table_path_equals_series():
    """The table path of the device sincos settles almost every argument
    and, where it does, returns what the ~2^-80-accurate series returns:
    the whole function is bit-identical to the series alone on omega
    grids, wide ranges, tiny arguments and arguments next to multiples of
    pi/2 and to the table's nodes."""
    import ctypes
    lib = _lib.load()
    rng = np.random.default_rng(5)
    n = 200000
    for x in (np.pi / 180.0 * rng.uniform(-720, 720, n), rng.uniform(-823549, 823549, n),
              np.exp(rng.uniform(np.log(1e-300), 0.0, n)) * rng.choice([-1, 1], n),
              (np.arange(n) % 2000 - 1000) * (np.pi / 2) * (1 + rng.normal(scale=1e-15, size=n)),
              rng.integers(-101, 102, n) / 128.0 + rng.integers(-50, 50, n) * np.pi / 2):
        x = np.ascontiguousarray(x)
        s, c = _host_sincos(x)
        s2, c2 = np.zeros_like(x), np.zeros_like(x)
        settled = ctypes.c_int(0)
        lib.gs_selftest_sincos_series_host(_lib.hptr(x), _lib.hptr(s2), _lib.hptr(c2), n,
                                           ctypes.byref(settled))
        assert np.array_equal(s.view(np.int64), s2.view(np.int64))
        assert np.array_equal(c.view(np.int64), c2.view(np.int64))
        assert settled.value > 0.97 * n
