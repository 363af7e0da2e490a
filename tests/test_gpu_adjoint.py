"""The adjoint of the joint objective: value against the oracle, gradient
against central finite differences (h = 1e-6 relative) for every physical
geometry parameter and for UBI and translation entries, the g-space
gradient against the normal equations of score_and_refine, determinism,
and the joint Gauss-Newton/Adam refinement."""
import numpy as np
import pytest
import torch

from imaged11_amd import adjoint, geometry, simulate

pytestmark = pytest.mark.gpu


def T(a, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


@pytest.fixture(scope="module")
def setup(dev):
    wl = simulate.scoring_workload(K=120, N=60000, seed=41, translations=True)
    pk, pars = wl["peaks"], wl["pars"]
    pars = dict(pars, wedge=0.7, chi=-0.4)  # exercise the wedge/chi path too
    idx, off = simulate.refine_lists(pk["grain"], 120, seed=42, median=200, lo=20, hi=1000)
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                pars["omegasign"], T(idx, dev, torch.int32),
                                T(off, dev, torch.int64))
    params = {k: torch.tensor(float(pars[k]), dtype=torch.float64, device=dev)
              for k in adjoint.FREE}
    ubis = T(wl["ubi"], dev)
    trans = T(wl["trans"], dev)
    geo = adjoint.pack_geometry(params, pars)
    prob.assign_hkl(geo, ubis, trans)
    return dict(wl=wl, pars=pars, idx=idx, off=off, prob=prob, params=params, ubis=ubis,
                trans=trans, geo=geo)


def test_pack_geometry_matches_host_packing(setup):
    p, r, dist = geometry.pack_xlylzl(setup["pars"])
    geo = setup["geo"].cpu().numpy()
    np.testing.assert_allclose(geo[:4], p, rtol=0, atol=0)
    np.testing.assert_allclose(geo[4:13], r, rtol=0, atol=1e-15)
    np.testing.assert_allclose(geo[13:16], dist, rtol=0, atol=0)
    assert geo[16] == setup["pars"]["wavelength"]
    np.testing.assert_allclose(geo[17:], geometry.wedge_chi_matrix(0.7, -0.4), atol=1e-15)


def test_loss_matches_oracle(setup, orc):
    s = setup
    loss, *_ = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"])
    geo = s["geo"].cpu().numpy()
    pk = s["wl"]["peaks"]
    so, co = s["prob"].so.cpu().numpy(), s["prob"].co.cpu().numpy()
    want = orc.joint_loss(pk["sc"], pk["fc"], so, co, geo[:4], geo[4:13], geo[13:16], geo[16],
                          geo[17:], s["wl"]["ubi"], s["wl"]["trans"], s["idx"], s["off"],
                          s["prob"].hkl.cpu().numpy())
    assert abs(float(loss) - want) <= 1e-10 * abs(want)


def _loss_at(s, params=None, ubis=None, trans=None):
    geo = adjoint.pack_geometry(params or s["params"], s["pars"])
    loss, *_ = s["prob"].evaluate(geo, s["ubis"] if ubis is None else ubis,
                                  s["trans"] if trans is None else trans)
    return float(loss)


def test_geometry_gradient_finite_differences(setup):
    s = setup
    params = {k: v.clone().requires_grad_(True) for k, v in s["params"].items()}
    loss = adjoint.joint_loss(s["prob"], s["ubis"], s["trans"], params, s["pars"])
    loss.backward()
    floors = {"tilt_x": 1e-3, "tilt_y": 1e-3, "tilt_z": 1e-3, "wedge": 1.0, "chi": 1.0}
    for k in adjoint.FREE:
        v = float(s["params"][k])
        h = 1e-6 * max(abs(v), floors.get(k, 0.0))
        up = {kk: vv.clone() for kk, vv in s["params"].items()}
        dn = {kk: vv.clone() for kk, vv in s["params"].items()}
        up[k] = up[k] + h
        dn[k] = dn[k] - h
        fd = (_loss_at(s, up) - _loss_at(s, dn)) / (2 * h)
        g = float(params[k].grad)
        assert abs(g - fd) <= 1e-5 * max(abs(fd), abs(g)) + 1e-9 * abs(float(loss)), (k, g, fd)


def _grain_loss_at(s, g, ubis=None, trans=None):
    """Loss of grain g alone: a UBI or translation entry of grain g moves
    only that term, so its finite difference is not buried in the rounding
    noise of the other grains' terms."""
    s["prob"].evaluate(s["geo"], s["ubis"] if ubis is None else ubis,
                       s["trans"] if trans is None else trans)
    return float(s["prob"].loss_g[g])


def test_ubi_and_translation_gradient_finite_differences(setup):
    s = setup
    _, gu, gt, _, _ = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"])
    gu, gt = gu.clone(), gt.clone()
    rng = np.random.default_rng(0)
    for _ in range(6):
        g, i, j = int(rng.integers(0, 120)), int(rng.integers(0, 3)), int(rng.integers(0, 3))
        h = 1e-6 * abs(float(s["ubis"][g, i, j])) + 1e-9
        up, dn = s["ubis"].clone(), s["ubis"].clone()
        up[g, i, j] += h
        dn[g, i, j] -= h
        fd = (_grain_loss_at(s, g, ubis=up) - _grain_loss_at(s, g, ubis=dn)) / (2 * h)
        assert abs(float(gu[g, i, j]) - fd) <= 1e-6 * abs(fd) + 1e-8
        # translations move g-vectors by ~|t|/distance: a larger step keeps
        # the difference well above rounding (the loss is smooth in t)
        c = int(rng.integers(0, 3))
        h = 1e-2
        up, dn = s["trans"].clone(), s["trans"].clone()
        up[g, c] += h
        dn[g, c] -= h
        lg = _grain_loss_at(s, g)
        fd = (_grain_loss_at(s, g, trans=up) - _grain_loss_at(s, g, trans=dn)) / (2 * h)
        assert abs(float(gt[g, c]) - fd) <= 1e-5 * abs(fd) + 1e-13 * abs(lg) / h, (g, c)


def test_gspace_gradient_vanishes_at_score_and_refine_solution(setup, orc):
    """g-space gradient UB.H - R is zero at UB = R.H^-1, the matrix that
    score_and_refine (src/closest.c:267-345) returns for the same peaks."""
    s = setup
    pk = s["wl"]["peaks"]
    geo = s["geo"].cpu().numpy()
    so, co = s["prob"].so.cpu().numpy(), s["prob"].co.cpu().numpy()
    xl = np.zeros((len(pk["sc"]), 3))
    orc.compute_xlylzl(pk["sc"], pk["fc"], geo[:4], geo[4:13], geo[13:16], xl)
    ubis_new = s["wl"]["ubi"].copy()
    hkl = s["prob"].hkl.cpu().numpy()
    for g in range(120):
        sl = slice(s["off"][g], s["off"][g + 1])
        gv = orc.compute_gv_trig(xl[s["idx"][sl]], so[s["idx"][sl]], co[s["idx"][sl]], geo[17:],
                                 geo[16], s["wl"]["trans"][g])
        u = ubis_new[g].copy()
        n, _ = orc.score_and_refine(u, gv, 1.0)  # tol 1: every list peak kept
        assert n == s["off"][g + 1] - s["off"][g]
        assert np.array_equal(np.rint(gv @ ubis_new[g].T), hkl[sl])
        ubis_new[g] = u
    _, gub, _, _, _ = s["prob"].evaluate(s["geo"], T(ubis_new, s["ubis"].device), s["trans"],
                                         gspace=True)
    gub = gub.cpu().numpy()
    # scale: |R| ~ sum |h||g|
    scale = np.array([np.abs(hkl[s["off"][g]:s["off"][g + 1]]).sum() for g in range(120)])
    assert np.all(np.abs(gub).reshape(120, -1).max(1) <= 1e-11 * scale)
    # away from the solution it equals UB.H - R computed in numpy
    _, gub0, _, _, _ = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], gspace=True)
    g = 7
    sl = slice(s["off"][g], s["off"][g + 1])
    gv = orc.compute_gv_trig(xl[s["idx"][sl]], so[s["idx"][sl]], co[s["idx"][sl]], geo[17:],
                             geo[16], s["wl"]["trans"][g])
    h = hkl[sl].astype(float)
    R = gv.T @ h
    H = h.T @ h
    UB = np.linalg.inv(s["wl"]["ubi"][g])
    np.testing.assert_allclose(gub0[g].cpu().numpy(), UB @ H - R, rtol=1e-9, atol=1e-9 * np.abs(R).max())


def test_adjoint_deterministic(setup):
    s = setup
    a = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True)
    b = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True)
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_joint_refinement_noisy_data(setup):
    """50 steps (Gauss-Newton UBIs + Adam geometry) from a perturbed
    detector on jittered, omega-binned spots: the fit ends at least as good
    as the fit at the true geometry.  (With this noise the true geometry is
    not the minimum, so the parameters themselves are not asserted here.)"""
    s = setup
    truth = {k: float(v) for k, v in s["params"].items()}
    names = ("distance", "y_center", "z_center", "tilt_z")
    dev = s["ubis"].device
    # the fit at the true geometry: Gauss-Newton UBIs, then the loss
    _, _, _, _, sums = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True)
    best_loss = float(s["prob"].evaluate(s["geo"], adjoint.gauss_newton_ubis(sums),
                                         s["trans"])[0])
    start = dict(truth, distance=truth["distance"] + 150.0, y_center=truth["y_center"] + 0.6,
                 z_center=truth["z_center"] - 0.4, tilt_z=truth["tilt_z"] + 2e-4)
    params = {k: torch.tensor(start[k], dtype=torch.float64, device=dev) for k in names}
    scales = {"distance": 100.0, "y_center": 1.0, "z_center": 1.0, "tilt_z": 1e-3}
    ubis, final, hist = adjoint.refine_joint(s["prob"], s["ubis"], s["trans"], params,
                                             dict(s["pars"]), steps=50, lr=0.05, scales=scales)
    assert hist[-1] < hist[0]
    assert hist[-1] <= 1.05 * best_loss


def test_gauss_newton_kernel_against_solver(setup, dev):
    """gs_gauss_newton_ubis on the device against numpy's solver of each
    grain's normal equations, on the sums of a real evaluation and on
    random ones with a singular grain; in place gives the same bits."""
    s = setup
    _, _, _, _, sums = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True)
    rng = np.random.default_rng(9)
    K = 300
    G = rng.normal(size=(K, 30, 3))
    h = rng.normal(size=(K, 30, 3))
    A = np.einsum("kni,knj->kij", h, G)
    M = np.einsum("kni,knj->kij", G, G)
    M[5] = np.outer([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])  # singular
    rand = np.concatenate([A.reshape(K, 9), M[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]], 1)
    for sm, prev in ((sums, s["ubis"]), (T(rand, dev), T(rng.normal(size=(K, 3, 3)), dev))):
        got = adjoint.gauss_newton_ubis(sm, prev).cpu().numpy()
        n = sm.cpu().numpy()
        Ah = n[:, :9].reshape(-1, 3, 3)
        Mh = n[:, [9, 10, 11, 10, 12, 13, 11, 13, 14]].reshape(-1, 3, 3)
        for k in range(len(n)):
            if np.linalg.det(Mh[k]) == 0.0:
                assert np.array_equal(got[k], prev[k].cpu().numpy())
                continue
            want = np.linalg.solve(Mh[k], Ah[k].T).T
            np.testing.assert_allclose(got[k], want, rtol=1e-10,
                                       atol=1e-12 * np.abs(want).max())
        inplace = prev.clone()
        adjoint.gauss_newton_ubis(sm, inplace, out=inplace)
        assert np.array_equal(inplace.cpu().numpy(), got)
    assert bool(torch.isnan(adjoint.gauss_newton_ubis(T(rand, dev))[5]).all())


def test_refine_joint_follows_the_autograd_loop(setup):
    """refine_joint (closed-form packing Jacobian, host Adam, device
    Gauss-Newton step) against the same algorithm written with torch
    autograd through pack_geometry and torch.optim.Adam: losses, final
    parameters and UBIs agree over 12 steps."""
    s = setup
    names = ("chi", "distance", "tilt_x", "wavelength", "wedge", "y_center")
    start = {k: float(v) for k, v in s["params"].items()}
    start.update(distance=start["distance"] + 120.0, y_center=start["y_center"] - 0.5,
                 tilt_x=start["tilt_x"] + 1e-4, wedge=start["wedge"] + 0.01)
    scales = {"distance": 100.0, "tilt_x": 1e-3, "wavelength": 1e-5, "wedge": 0.01,
              "chi": 0.01, "y_center": 1.0}
    dev = s["ubis"].device
    params = {k: torch.tensor(start[k], dtype=torch.float64, device=dev) for k in names}
    ubis, final, hist = adjoint.refine_joint(s["prob"], s["ubis"], s["trans"], params,
                                             dict(s["pars"]), steps=12, lr=0.05, scales=scales)
    # the autograd loop
    x = {k: torch.zeros((), dtype=torch.float64, requires_grad=True) for k in sorted(names)}
    opt = torch.optim.Adam([x[k] for k in sorted(names)], lr=0.05, foreach=False)
    u = s["ubis"].clone()
    want = []
    for _ in range(12):
        cur = {k: start[k] + scales[k] * x[k] for k in names}
        geo = adjoint.pack_geometry(cur, s["pars"])
        loss, _, _, gg, sums = s["prob"].evaluate(geo.detach(), u, s["trans"], want_sums=True)
        u = adjoint.gauss_newton_ubis(sums, u)
        want.append(float(loss))
        opt.zero_grad()
        geo.backward(gg.cpu())
        opt.step()
    np.testing.assert_allclose(hist, want, rtol=1e-9)
    for k in names:
        v = start[k] + scales[k] * float(x[k])
        assert abs(float(final[k]) - v) <= 1e-9 * scales[k], k
    np.testing.assert_allclose(ubis.cpu().numpy(), u.cpu().numpy(), rtol=0,
                               atol=1e-9 * float(u.abs().max()))
    assert hist[-1] < hist[0]


@pytest.fixture(scope="module")
def exact_setup(dev):
    """Noise-free spots (no pixel jitter, exact omega, no spurious peaks):
    the true geometry is the exact minimum of the joint objective."""
    wl = simulate.scoring_workload(K=60, N=30000, seed=43, translations=True)
    pars = dict(wl["pars"], wedge=0.7, chi=-0.4)
    pk = simulate.simulate_peaks(wl["ub"], wl["trans"], pars, 12000, seed=44, spurious=0.0,
                                 omega_step=0.0, jitter_px=0.0)
    # each grain's list = exactly its own spots
    idx = np.argsort(pk["grain"], kind="stable").astype(np.int32)
    off = np.concatenate([[0], np.cumsum(np.bincount(pk["grain"], minlength=60))])
    assert np.all(np.diff(off) >= 20)
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                pars["omegasign"], T(idx, dev, torch.int32),
                                T(off, dev, torch.int64))
    params = {k: torch.tensor(float(pars[k]), dtype=torch.float64, device=dev)
              for k in adjoint.FREE}
    ubis, trans = T(wl["ubi"], dev), T(wl["trans"], dev)
    prob.assign_hkl(adjoint.pack_geometry(params, pars), ubis, trans)
    return dict(pars=pars, prob=prob, params=params, ubis=ubis, trans=trans)


def test_joint_refinement_recovers_exact_geometry(exact_setup):
    """On noise-free spots the refinement from a perturbed detector drives
    the loss towards zero and the beam centre back towards the truth.
    (Distance trades off against the cell scale of the UBIs, so it is not
    asserted.)"""
    s = exact_setup
    truth = {k: float(v) for k, v in s["params"].items()}
    geo = adjoint.pack_geometry(s["params"], s["pars"])
    assert float(s["prob"].evaluate(geo, s["ubis"], s["trans"])[0]) < 1e-12
    names = ("distance", "y_center", "z_center", "tilt_z")
    dev = s["ubis"].device
    start = dict(truth, distance=truth["distance"] + 150.0, y_center=truth["y_center"] + 0.6,
                 z_center=truth["z_center"] - 0.4, tilt_z=truth["tilt_z"] + 2e-4)
    params = {k: torch.tensor(start[k], dtype=torch.float64, device=dev) for k in names}
    scales = {"distance": 100.0, "y_center": 1.0, "z_center": 1.0, "tilt_z": 1e-3}
    ubis, final, hist = adjoint.refine_joint(s["prob"], s["ubis"], s["trans"], params,
                                             dict(s["pars"]), steps=100, lr=0.05, scales=scales)
    assert hist[-1] < 0.05 * hist[0]
    for k in ("y_center", "z_center"):
        assert abs(float(final[k]) - truth[k]) < 0.5 * abs(start[k] - truth[k]), k


def test_packed_entries_bit_identical_to_indexed(setup, dev):
    """The per-entry copies (no index) give the indexed pass's values in
    the same order: hkl, loss and every gradient bit for bit."""
    s = setup
    pk = s["wl"]["peaks"]
    packed = s["prob"]
    assert packed.rows is not None
    indexed = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                   s["pars"]["omegasign"], T(s["idx"], dev, torch.int32),
                                   T(s["off"], dev, torch.int64))
    indexed.rows = None
    indexed.assign_hkl(s["geo"], s["ubis"], s["trans"])
    assert torch.equal(indexed.hkl, packed.hkl)
    for gspace in (False, True):
        a = packed.evaluate(s["geo"], s["ubis"], s["trans"], gspace=gspace, want_sums=True)
        b = indexed.evaluate(s["geo"], s["ubis"], s["trans"], gspace=gspace, want_sums=True)
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def test_out_of_range_index_keeps_lists_indexed(setup, dev):
    s = setup
    pk = s["wl"]["peaks"]
    idx = s["idx"].copy()
    idx[7] = len(pk["sc"]) + 3
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                s["pars"]["omegasign"], T(idx, dev, torch.int32),
                                T(s["off"], dev, torch.int64))
    assert prob.rows is None
    prob.assign_hkl(s["geo"], s["ubis"], s["trans"])
    loss, *_ = prob.evaluate(s["geo"], s["ubis"], s["trans"])
    assert bool(torch.isfinite(loss))
    assert prob.hkl[7].abs().sum() == 0


def test_parity_mode_inputs_are_the_references(setup, orc, dev):
    """JointProblem(parity=True): sin/cos of omega and the wedge/chi matrix
    are glibc's (the reference's values, bit for bit); the loss then equals
    the oracle's objective on the oracle's own trig within 1e-10 (the
    adjoint's own FMA / reciprocal-square-root arithmetic and summation
    order) and the geometry gradient keeps flowing through mat."""
    s = setup
    pk = s["wl"]["peaks"]
    pars = s["pars"]
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                pars["omegasign"], T(s["idx"], dev, torch.int32),
                                T(s["off"], dev, torch.int64), parity=True)
    so, co = orc.sincos_omega(pk["omega"], pars["omegasign"])
    assert np.array_equal(prob.so.cpu().numpy().view(np.int64), so.view(np.int64))
    assert np.array_equal(prob.co.cpu().numpy().view(np.int64), co.view(np.int64))
    params = {k: v.clone().requires_grad_(True) for k, v in s["params"].items()}
    geo = adjoint.pack_geometry(params, pars, parity=True)
    g = geo.detach().cpu().numpy()
    assert np.array_equal(g[17:].view(np.int64),
                          orc.wedge_chi_matrix(pars["wedge"], pars["chi"]).view(np.int64))
    prob.assign_hkl(geo, s["ubis"], s["trans"])
    loss = adjoint.joint_loss(prob, s["ubis"], s["trans"], params, pars)
    want = orc.joint_loss(pk["sc"], pk["fc"], so, co, g[:4], g[4:13], g[13:16], g[16], g[17:],
                          s["wl"]["ubi"], s["wl"]["trans"], s["idx"], s["off"],
                          prob.hkl.cpu().numpy())
    assert abs(float(loss) - want) <= 1e-10 * abs(want)
    loss.backward()
    assert params["wedge"].grad is not None and float(params["wedge"].grad) != 0.0


def test_gauss_newton_mode_ubi_gradient_matches_direct(setup):
    """With the Gauss-Newton sums requested, dL/dUBI is formed from them,
    2 (UBI . sum G G^T - sum h G^T), instead of being accumulated pair by
    pair: the two agree to rounding, and the loss, translation and
    geometry gradients are the same computation in both modes."""
    s = setup
    l1, gu1, gt1, gg1, _ = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"])
    l2, gu2, gt2, gg2, sums = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"],
                                                want_sums=True)
    assert abs(float(l1) - float(l2)) <= 1e-14 * abs(float(l1))
    scale = float(gu1.abs().max())
    assert float((gu2 - gu1).abs().max()) <= 1e-9 * scale
    torch.testing.assert_close(gt2, gt1, rtol=1e-12, atol=1e-12 * float(gt1.abs().max()))
    torch.testing.assert_close(gg2, gg1, rtol=1e-12, atol=1e-12 * float(gg1.abs().max()))
    # and the sums themselves: sum G G^T is symmetric positive semi-definite
    GG = sums[:, 9:]
    assert bool((GG[:, 0] >= 0).all() and (GG[:, 3] >= 0).all() and (GG[:, 5] >= 0).all())


@pytest.mark.parametrize("piece_len", [16, 100, 4096])
def test_pieces_and_lanes_give_the_same_results(setup, dev, piece_len):
    """The pass over pieces: any piece length and lane count gives the
    same per-grain loss, gradients and sums (to the rounding of a
    different summation order), each grain's geo_part sums to the total
    dL/dgeo, and grain g's loss_g is the sum of its own entries' terms
    (one piece per grain at piece_len 4096 is the plain per-grain sum)."""
    s = setup
    pk = s["wl"]["peaks"]
    ref = s["prob"].evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True)
    ref_lg = s["prob"].loss_g.clone()
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                s["pars"]["omegasign"], T(s["idx"], dev, torch.int32),
                                T(s["off"], dev, torch.int64), piece_len=piece_len)
    lengths = np.diff(s["off"])
    assert prob.n_pieces == int((-(-lengths // piece_len)).sum())
    prob.hkl.copy_(s["prob"].hkl)
    for lanes in (16, 32, 64):
        part = torch.empty((120, adjoint.GEO), dtype=torch.float64, device=dev)
        got = prob.evaluate(s["geo"], s["ubis"], s["trans"], want_sums=True,
                            lanes_per_piece=lanes, geo_part=part)
        for x, y in zip(got, ref):
            torch.testing.assert_close(x, y, rtol=1e-11, atol=1e-11 * float(y.abs().max()))
        torch.testing.assert_close(prob.loss_g, ref_lg, rtol=1e-11, atol=0)
        total = got[3].clone()
        total[16] = total[16] * -(s["geo"][16] ** 2)  # back to d/d(1/wavelength)
        torch.testing.assert_close(part.sum(0), total, rtol=1e-10,
                                   atol=1e-10 * float(total.abs().max()))


@pytest.mark.parametrize("piece_len,lanes", [(1, 16), (7, 32), (512, 64), (64, 16)])
def test_pieces_ragged_and_empty_lists(setup, dev, orc, piece_len, lanes):
    """Ragged lists around the piece boundaries: empty lists first, in the
    middle and last, one entry, exactly piece_len, piece_len + 1.  Every
    grain's loss against the oracle's objective on that grain alone, empty
    grains with all-zero outputs, and the totals the sums of the grains."""
    s = setup
    pk = s["wl"]["peaks"]
    rng = np.random.default_rng(7)
    lens = [0, 0, 1, piece_len, piece_len + 1, 0, 2 * piece_len + 3, 37, 0, 5, 0]
    K = len(lens)
    idx = rng.integers(0, len(pk["sc"]), int(sum(lens))).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    gs = rng.integers(0, 120, K)
    ubis = s["ubis"][torch.as_tensor(gs, device=dev)].contiguous()
    trans = s["trans"][torch.as_tensor(gs, device=dev)].contiguous()
    prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                s["pars"]["omegasign"], T(idx, dev, torch.int32),
                                T(off, dev, torch.int64), piece_len=piece_len)
    assert prob.n_pieces == sum(-(-n // piece_len) for n in lens)
    prob.assign_hkl(s["geo"], ubis, trans)
    part = torch.empty((K, adjoint.GEO), dtype=torch.float64, device=dev)
    loss, gu, gt, gg, sums = prob.evaluate(s["geo"], ubis, trans, want_sums=True,
                                           lanes_per_piece=lanes, geo_part=part)
    lg = prob.loss_g.cpu().numpy()
    geo = s["geo"].cpu().numpy()
    so, co = prob.so.cpu().numpy(), prob.co.cpu().numpy()
    hkl = prob.hkl.cpu().numpy()
    u_h, t_h = ubis.cpu().numpy(), trans.cpu().numpy()
    for g in range(K):
        if lens[g] == 0:
            assert lg[g] == 0.0
            for x in (gu[g], gt[g], sums[g], part[g]):
                assert bool((x == 0).all())
            continue
        want = orc.joint_loss(pk["sc"], pk["fc"], so, co, geo[:4], geo[4:13], geo[13:16],
                              geo[16], geo[17:], u_h[g:g + 1], t_h[g:g + 1], idx,
                              off[g:g + 2], hkl)
        assert abs(lg[g] - want) <= 1e-10 * abs(want) + 1e-300, (g, lg[g], want)
    assert abs(float(loss) - lg.sum()) <= 1e-12 * abs(lg.sum())
    total = gg.clone()
    total[16] = total[16] * -(s["geo"][16] ** 2)
    torch.testing.assert_close(part.sum(0), total, rtol=1e-10,
                               atol=1e-10 * float(total.abs().max()))


def test_no_grains_and_no_entries(setup, dev):
    """K = 0, and grains whose lists are all empty: zero loss and
    gradients, no pieces, nothing launched over pieces."""
    s = setup
    pk = s["wl"]["peaks"]
    for off in (np.zeros(1, np.int64), np.zeros(4, np.int64)):
        K = len(off) - 1
        prob = adjoint.JointProblem(T(pk["sc"], dev), T(pk["fc"], dev), T(pk["omega"], dev),
                                    s["pars"]["omegasign"], T(np.zeros(0), dev, torch.int32),
                                    T(off, dev, torch.int64))
        assert prob.n_pieces == 0
        ubis = s["ubis"][:K].contiguous()
        trans = s["trans"][:K].contiguous()
        prob.assign_hkl(s["geo"], ubis, trans)
        loss, gu, gt, gg, sums = prob.evaluate(s["geo"], ubis, trans, want_sums=True)
        assert float(loss) == 0.0 and bool((gg == 0).all())
        assert gu.shape == (K, 3, 3) and bool((gu == 0).all()) and bool((sums == 0).all())


def test_joint_adjoint_rejects_bad_arguments(setup, dev):
    """The C entry point refuses what it cannot run: negative sizes, a
    piece length below 1, missing outputs, per-entry arrays shorter than the
    lists; every refusal is a status and a message, nothing launched."""
    from imaged11_amd import _lib
    from imaged11_amd._lib import dptr, stream_ptr
    lib = _lib.load()
    p = setup["prob"]
    K = p.K
    (a, b, c, d), npk, idx = p._peaks()
    geo = setup["geo"].contiguous()
    ubis, trans = setup["ubis"], setup["trans"]
    gu = torch.empty((K, 3, 3), dtype=torch.float64, device=dev)
    gt = torch.empty((K, 3), dtype=torch.float64, device=dev)
    out = torch.empty(27, dtype=torch.float64, device=dev)

    def call(**kw):
        args = dict(npk=npk, K=K, n_idx=p.idx.numel(), n_pieces=p.n_pieces,
                    piece_len=p.piece_len, grad_ubi=dptr(gu), out=dptr(out))
        args.update(kw)
        lib.gs_clear_error()
        return lib.gs_joint_adjoint(
            dptr(a), dptr(b), dptr(c), dptr(d), args["npk"], dptr(geo), dptr(ubis),
            dptr(trans), args["K"], idx, args["n_idx"], dptr(p.off), dptr(p.hkl),
            dptr(p.piece_off), dptr(p.piece_grain), dptr(p.slots), args["n_pieces"],
            args["piece_len"], 0, dptr(p.loss_g), args["grad_ubi"], dptr(gt), None, None,
            args["out"], 16, dptr(p.ws), stream_ptr())
    assert call() == 0
    for bad in (dict(K=-1), dict(n_idx=-1), dict(n_pieces=-1), dict(piece_len=0),
                dict(grad_ubi=None), dict(out=None), dict(npk=p.idx.numel() - 1)):
        assert call(**bad) != 0, bad
        assert _lib.last_error(), bad
    lib.gs_clear_error()
