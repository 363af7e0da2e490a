"""The refinegrains row on the CPU: the oracle's restatement pinned, the
simplex (product and oracle) against the reference's own trajectories,
and the kernel's per-entry arithmetic (host build) against the oracle."""
import json
import math
import os

import numpy as np
import pytest

from imaged11_amd import _lib, geometry, parameters, simplex, simulate
from imaged11_amd import refinegrains as RG
from oracle import refinegrains_oracle as R

G = os.path.join(os.path.dirname(__file__), "golden")


# ---- simplex --------------------------------------------------------------

def _problems():
    import importlib.util
    spec = importlib.util.spec_from_file_location("msg", os.path.join(G, "make_simplex_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return {name: f for name, f, *_ in mod.PROBLEMS}


GOLD = json.load(open(os.path.join(G, "simplex_golden.json")))


@pytest.mark.parametrize("impl", ["oracle", "product"])
@pytest.mark.parametrize("name", sorted(GOLD))
def test_simplex_matches_reference_trajectory(impl, name):
    """Every objective call (point and order), the result, the error and
    the iteration count equal the reference simplex's, bit for bit."""
    d = GOLD[name]
    f = _problems()[name]
    calls = []

    def rec(a):
        v = f(a)
        calls.append([list(map(float, a)), float(v)])
        return v
    cls = R.Simplex if impl == "oracle" else simplex.Simplex
    res, err, it = cls(rec, list(d["guess"]), list(d["increments"])).minimize(**d["kwargs"])
    assert calls == d["calls"]
    assert list(map(float, res)) == d["result"]
    assert err == d["error"] and it == d["iter"]


def test_simplex_batch_lockstep_equals_sequential():
    """40 problems stepped together make, per problem, exactly the calls of
    the sequential oracle simplex, and return its results."""
    f = _problems()
    rng = np.random.default_rng(0)
    B = 40
    gs = rng.normal(size=(B, 3)) * 2
    fn = lambda b, a: f["stepped"](a) if b % 2 else f["bowl_5_iters"](a)
    seq = []
    for b in range(B):
        calls = []

        def rec(a, b=b, calls=calls):
            calls.append(list(map(float, a)))
            return fn(b, list(a))
        seq.append((calls, R.Simplex(rec, list(gs[b]), [0.3] * 3).minimize(maxiters=80)))
    got = [[] for _ in range(B)]

    def ev(pts, which):
        out = []
        for p, b in zip(pts, which):
            got[b].append(list(map(float, p)))
            out.append(fn(b, list(p)))
        return out
    sb = simplex.SimplexBatch(ev, gs, [0.3, 0.3, 0.3])
    P, E, I = sb.minimize(maxiters=80)
    for b in range(B):
        calls, (res, err, it) = seq[b]
        assert got[b] == calls
        assert list(P[b]) == list(res) and E[b] == err and I[b] == it
        assert list(sb.last[b]) == calls[-1]
    assert sb.rounds < sum(len(c) for c, _ in seq)


def test_simplex_batch_nan_ties_and_limits():
    """Objectives with NaN regions and flat plateaus (ties in every
    comparison) stepped together equal the sequential oracle simplex call
    for call; more than 64 variables is refused."""
    import math
    rng = np.random.default_rng(3)
    B = 24
    gs = rng.normal(size=(B, 3))

    def fn(b, a):
        x, y, z = a
        if b % 3 == 0 and x > 0.7:
            return float("nan")
        if b % 3 == 1:
            return float(math.floor(2 * (x * x + y * y + z * z)))  # plateaus
        return (x - 0.3) ** 2 + 2 * (y + 0.1) ** 2 + 0.5 * z * z
    seq = []
    for b in range(B):
        calls = []

        def rec(a, b=b, calls=calls):
            calls.append(list(map(float, a)))
            return fn(b, list(a))
        seq.append((calls, R.Simplex(rec, list(gs[b]), [0.4] * 3).minimize(maxiters=60)))
    got = [[] for _ in range(B)]

    def ev(pts, which):
        out = []
        for p, b in zip(pts, which):
            got[b].append(list(map(float, p)))
            out.append(fn(b, list(p)))
        return out
    P, E, I = simplex.SimplexBatch(ev, gs, [0.4] * 3).minimize(maxiters=60)
    for b in range(B):
        calls, (res, err, it) = seq[b]
        assert got[b] == calls, b
        assert list(P[b]) == list(res) and I[b] == it
        assert E[b] == err or (math.isnan(E[b]) and math.isnan(err))
    with pytest.raises(RuntimeError):
        simplex.SimplexBatch(lambda p, w: np.zeros(len(w)), np.zeros((2, 65)),
                             np.ones(65)).minimize(maxiters=2)


# ---- transform chain -------------------------------------------------------

@pytest.mark.parametrize("wedge,chi", [(0.0, 0.0), (1.5, -2.0), (-3.0, 0.0)])
def test_numpy_chain_equals_compute_gv(ref, wedge, chi):
    """compute_g_vectors(compute_tth_eta_from_xyz(xyz, omega*sign, t)) is
    the compiled reference's compute_gv with translation t."""
    rng = np.random.default_rng(1)
    n = 4000
    xyz = np.stack([rng.uniform(1e5, 2e5, n), rng.uniform(-9e4, 9e4, n),
                    rng.uniform(-9e4, 9e4, n)])
    om = rng.uniform(-180, 360, n)
    for sign in (1.0, -1.0):
        t = np.array([40.0, -70.0, 12.0])
        p = dict(t_x=t[0], t_y=t[1], t_z=t[2], wedge=wedge, chi=chi)
        tth, eta = R.compute_tth_eta_from_xyz(xyz, om * sign, **p)
        g = R.compute_g_vectors(tth, eta, om * sign, 0.2, wedge, chi).T
        want = np.zeros((n, 3))
        ref.compute_gv(np.ascontiguousarray(xyz.T), om, sign, 0.2, wedge, chi, t, want)
        np.testing.assert_allclose(g, want, rtol=0, atol=1e-13 * np.abs(want).max())


@pytest.mark.parametrize("w,c", [(5.0, 10.0), (5.0, 0.0), (0.0, 10.0), (0.0, 0.0)])
def test_uncompute_g_vectors_inverts(w, c):
    """The reference's own test (test/test_uncomputegv.py): one of the two
    solutions gives back tth, eta and omega."""
    tth = np.array([60, 10, 15, 20, 25, 1.9555], float)
    eta = np.array([20, 10, 120, -20, 340, -73], float)
    omega = np.array([60, 90, 180, 60, 97, 131], float)
    g = R.compute_g_vectors(tth, eta, omega, 0.5, wedge=w, chi=c)
    t2, e2, o2 = R.uncompute_g_vectors(g, 0.5, wedge=w, chi=c)
    angmod = lambda x: np.degrees(np.arctan2(np.sin(np.radians(x)), np.cos(np.radians(x))))
    for i in range(len(tth)):
        best = np.argmin(np.abs(angmod(np.array(o2)[:, i] - omega[i])))
        assert abs(tth[i] - t2[i]) < 1e-5
        assert abs(angmod(np.array(e2)[best, i] - eta[i])) < 1e-5
        assert abs(angmod(np.array(o2)[best, i] - omega[i])) < 1e-5


def test_uncompute_vectors_match_oracle():
    for w, c in [(0.0, 0.0), (1.5, -2.0), (0.0, 3.0)]:
        rb, pre = geometry.uncompute_vectors(0.3, w, c)
        rb2, pre2 = R.uncompute_vectors(0.3, w, c)
        assert np.array_equal(rb, rb2)
        assert (pre is None) == (pre2 is None)
        if pre is not None:
            assert np.array_equal(pre, pre2)


# ---- the kernel's arithmetic on the host -----------------------------------

def _host_grain_gv(lib, xl, so, co, omega, pars, trans, mats, idx, owner, slop):
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    rb, pre = geometry.uncompute_vectors(wvln, wedge, chi)
    M = len(idx)
    tth, eta, omc, gv = np.zeros(M), np.zeros(M), np.zeros(M), np.zeros((M, 3))
    P = lambda a: None if a is None else a.ctypes.data
    st = lib.gs_grain_gv_host(P(xl), P(so), P(co), P(omega), len(xl), sign, wvln, wedge, chi,
                              P(trans), P(mats), len(trans), _lib.darr(rb),
                              None if pre is None else _lib.darr(pre.ravel()), slop, P(idx),
                              P(owner), M, P(tth), P(eta), P(omc) if mats is not None else None,
                              P(gv))
    assert st == 0
    return gv, tth, eta, omc


@pytest.mark.parametrize("wedge,chi", [(0.0, 0.0), (1.5, -2.0)])
def test_grain_gv_host_matches_oracle(orc, wedge, chi):
    """gs_grain_gv's per-entry code (host build) against refinegrains'
    compute_gv restated in numpy, observed and omega-float."""
    lib = _lib.load()
    w = simulate.refinegrains_workload(K=8, N=6000, seed=3, wedge=wedge, chi=chi)
    p, pk = w["pars"], w["peaks"]
    rg = R.RefineGrains(orc, p, pk["sc"], pk["fc"], pk["omega"], w["ubi"], w["trans"],
                        tolerance=0.05, OmFloat=True, OmSlop=0.25)
    rg.assignlabels()
    xl = R.compute_xyz_lab([rg.sc, rg.fc], **p).T.copy()
    so, co = orc.sincos_omega(pk["omega"], p["omegasign"])
    moved = 0.0
    for g in range(8):
        gr = rg.grains[g]
        rg.set_translation(g)
        rg.OMEGA_FLOAT = False
        rg.compute_gv(gr)
        obs, tth_w, eta_w = rg.gv.copy(), rg.tth.copy(), rg.eta.copy()
        rg.OMEGA_FLOAT = True
        rg.compute_gv(gr)
        flo, omc_w = rg.gv.copy(), gr.omega_calc.copy()
        mat = gr.ubi.copy()
        orc.score_and_refine(mat, np.ascontiguousarray(obs), rg.tolerance)
        idx = gr.ind.astype(np.int32)
        own = np.zeros(len(idx), np.int32)
        tr = np.ascontiguousarray(gr.translation[None, :])
        gv, tth, eta, _ = _host_grain_gv(lib, xl, so, co, pk["omega"], p, tr, None, idx, own,
                                         0.25)
        scale = np.abs(obs).max()
        assert np.abs(gv - obs).max() <= 1e-14 * scale
        assert np.abs(tth - tth_w).max() < 1e-12 and np.abs(eta - eta_w).max() < 1e-12
        gv, _, _, omc = _host_grain_gv(lib, xl, so, co, pk["omega"], p, tr,
                                       np.ascontiguousarray(mat[None]), idx, own, 0.25)
        assert np.abs(gv - flo).max() <= 1e-14 * scale
        assert np.abs(omc - omc_w).max() < 1e-11
        moved = max(moved, np.abs(omc_w - pk["omega"][gr.ind]).max())
    assert 0.05 < moved <= 0.25  # the float step moved peaks, within the slop


def test_grain_gv_host_bad_entries():
    lib = _lib.load()
    xl = np.array([[1e5, 10.0, 20.0]])
    so, co, om = np.array([0.1]), np.array([math.sqrt(1 - 0.01)]), np.array([5.7])
    pars = dict(wavelength=0.2, omegasign=1.0, wedge=0.0, chi=0.0)
    tr = np.zeros((1, 3))
    idx = np.array([0, 1, -1, 0], np.int32)
    own = np.array([0, 0, 0, 1], np.int32)
    gv, tth, eta, _ = _host_grain_gv(lib, xl, so, co, om, pars, tr, None, idx, own, 0.25)
    assert np.isfinite(gv[0]).all()
    assert np.isnan(gv[1:]).all() and np.isnan(tth[1:]).all()
    # a singular fitted UBI gives a NaN angle (the reference's inv raises),
    # hence NaN x, y components (z does not depend on omega)
    gv, _, _, omc = _host_grain_gv(lib, xl, so, co, om, pars, tr, np.zeros((1, 3, 3)),
                                   idx[:1], own[:1], 0.25)
    assert np.isnan(omc[0]) and np.isnan(gv[0, :2]).all()


# ---- lattice symmetry ------------------------------------------------------

def test_lattice_symmetry_projection():
    rng = np.random.default_rng(2)
    U = simulate.random_rotations(5, rng)
    cells = np.array([[4.05, 4.05, 4.05, 90, 90, 90]] * 5, float)
    B = RG.b_matrix(cells)
    ubi = np.linalg.inv(U @ B) @ (np.eye(3) + rng.normal(scale=1e-3, size=(5, 3, 3)))
    # the oracle and the product agree, and triclinic is the identity
    for sym in ("cubic", "tetragonal", "orthorhombic", "hexagonal", "monoclinic_b"):
        a = RG.symmetrize(ubi, RG.LATTICES[sym])
        for k in range(5):
            b = R.symmetrize(ubi[k], getattr(R, sym))
            np.testing.assert_allclose(a[k], b, rtol=1e-12, atol=1e-12)
    assert RG.symmetrize(ubi, RG.triclinic) is ubi
    c = RG.symmetrize(ubi, RG.cubic)
    cc = RG.ubi_to_cell(c)
    np.testing.assert_allclose(cc[:, 3:], 90.0, atol=1e-9)
    np.testing.assert_allclose(cc[:, 0], cc[:, 1], rtol=1e-12)
    np.testing.assert_allclose(cc[:, 0], RG.ubi_to_cell(ubi)[:, :3].mean(1), rtol=1e-12)
    # the orientation is kept: U of the projection equals U of the input
    np.testing.assert_allclose(RG.ubi_to_u(c), RG.ubi_to_u(ubi), atol=1e-12)
    # a round trip through (U, cell) is the identity
    np.testing.assert_allclose(RG.u_to_ubi(RG.ubi_to_u(ubi), RG.ubi_to_cell(ubi)), ubi,
                               rtol=1e-12, atol=1e-12)


def test_cell_and_orientation_closed_forms():
    """UBI = (U B(cell))^-1 for known cells (cubic to triclinic) and random
    rotations: ubi_to_cell returns the cell, ubi_to_u the rotation, and
    u_to_ubi rebuilds the UBI (the product and the oracle)."""
    from imaged11_amd import unitcell
    rng = np.random.default_rng(31)
    cells = ([4.05] * 3 + [90.0] * 3, [3.0, 4.0, 5.0, 90.0, 90.0, 90.0],
             [3.1, 3.1, 5.2, 90.0, 90.0, 120.0], [3.1, 4.2, 5.3, 80.0, 95.0, 101.0])
    U = simulate.random_rotations(len(cells), rng)
    for k, cell in enumerate(cells):
        B = unitcell.unitcell(cell, "P").B
        ubi = np.linalg.inv(U[k] @ B)
        for mod in (RG, R):
            np.testing.assert_allclose(np.ravel(mod.ubi_to_cell(ubi)), cell, rtol=1e-12,
                                       atol=1e-10)
            np.testing.assert_allclose(np.reshape(mod.ubi_to_u(ubi), (3, 3)), U[k],
                                       atol=1e-12)
            np.testing.assert_allclose(np.reshape(mod.u_to_ubi(U[k], np.array(cell)), (3, 3)),
                                       ubi, rtol=1e-12, atol=1e-12)


def test_b_matrix_is_the_unitcell_b():
    from imaged11_amd import unitcell
    for cell in ([4.05] * 3 + [90] * 3, [3.1, 4.2, 5.3, 80.0, 95.0, 101.0]):
        want = unitcell.unitcell(cell, "P").B
        np.testing.assert_allclose(RG.b_matrix(np.array(cell, float)), want, rtol=1e-12,
                                   atol=1e-14)


def test_parameters_varylist():
    p = parameters.parameters(a=1.0, b=2.0, c=3.0)
    p.stepsizes.update(a=0.1, c=0.3)
    p.set_varylist(["c", "a"])
    assert p.get_variable_values() == [3.0, 1.0]
    assert p.get_variable_stepsizes() == [0.3, 0.1]
    p.set_variable_values([5.0, 6.0])
    assert p.get("c") == 5.0 and p.get("a") == 6.0
    with pytest.raises(ValueError):
        p.set_variable_values([1.0])
    with pytest.raises(KeyError):
        p.set_varylist(["zz"])


# ---- the oracle driver -----------------------------------------------------

def test_oracle_refinepositions_recovers_translations(orc):
    """Noise-free spots (exact omega, no jitter): gof is ~0 at the true
    translations, and refinepositions from 20-micron-off starting points
    comes back to them."""
    w = simulate.refinegrains_workload(K=4, N=3000, seed=5, spurious=0.0, omega_step=0.0,
                                       jitter_px=0.0)
    p, pk = w["pars"], w["peaks"]
    rng = np.random.default_rng(9)
    start = w["trans"] + rng.uniform(-20, 20, size=w["trans"].shape)
    rg = R.RefineGrains(orc, p, pk["sc"], pk["fc"], pk["omega"], w["ubi"], w["trans"],
                        tolerance=0.05, OmFloat=False)
    rg.assignlabels()
    rg.varylist = []
    for g in range(4):
        # gof reads the translation from the parameters (as the reference)
        rg.set_translation(g)
        rg.grains_to_refine = [g]
        assert rg.gof([]) < 1e-12 * 1e6
    for g in range(4):
        rg.grains[g].translation = start[g].copy()
    res = rg.refinepositions({"t_x": 0.2, "t_y": 0.2, "t_z": 0.2}, maxiters=400)
    for g in range(4):
        assert np.abs(rg.grains[g].translation - w["trans"][g]).max() < 2.0, (g, res[g])


def test_point_groups_and_unique_orientation():
    """makeuniq's groups: the reference's own sym_u.test() assertions
    (operator matrices, determinants, group orders, the cubic unique choice
    of a 90-degree rotation), the cubic group against its closed form (the
    24 signed permutation matrices of determinant +1), and find_uniq_u
    against a brute-force maximum of the trace."""
    import itertools
    from imaged11_amd import refinegrains as RG
    assert np.allclose(RG.symmetry_operator("x,y,z"), np.identity(3))
    assert np.allclose(RG.symmetry_operator("-y,x,z"), [[0, 1, 0], [-1, 0, 0], [0, 0, 1]])
    assert np.allclose(RG.symmetry_operator("-y,y-x,z"), [[0, -1, 0], [-1, 1, 0], [0, 0, 1]])
    for op in ("x,y,z", "-y,x-y,z", "-y,x,z"):
        assert np.linalg.det(RG.symmetry_operator(op)) == 1.0
    orders = {"cubic": 24, "hexagonal": 12, "trigonal": 6, "tetragonal": 8, "orthorhombic": 4,
              "monoclinic_a": 2, "monoclinic_b": 2, "monoclinic_c": 2, "triclinic": 1,
              "rhombohedralP": 6}
    for name, n in orders.items():
        assert len(RG.point_group(name)) == n, name
    closed = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            m = np.zeros((3, 3))
            for i, j in enumerate(perm):
                m[i, j] = signs[i]
            if np.linalg.det(m) > 0:
                closed.append(m)
    cub = RG.point_group("cubic")
    assert len(closed) == 24
    for m in closed:
        assert sum(np.allclose(m, c) for c in cub) == 1
    assert np.allclose(RG.find_uniq_u(np.array([[0., 1, 0], [-1, 0, 0], [0, 0, 1]]), cub),
                       np.identity(3))
    with pytest.raises(ValueError):
        RG.point_group("nonsense")
    with pytest.raises(ValueError):
        RG.symmetry_operator("2x,y,z")
    rng = np.random.default_rng(8)
    from imaged11_amd import simulate
    for U in simulate.random_rotations(50, rng):
        ubi = np.linalg.inv(U * 4.05)
        got = RG.find_uniq_u(ubi, cub)
        traces = [np.trace(m @ ubi) for m in closed] + [np.trace(ubi)]
        assert np.isclose(np.trace(got), max(traces), rtol=0, atol=1e-12)
        assert any(np.allclose(got, m @ ubi) for m in closed)
