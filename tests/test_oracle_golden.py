"""The oracle (clean-room C restatement) against the golden vectors written
by the reference C (tests/golden/make_golden.py).  Bit-exact throughout."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_verify_rounding(orc):
    s = load("scoring_cases.npz")
    got = [orc.verify_rounding(n) for n in (0, 1, 20, 12345)]
    assert got == s["verify_rounding"].tolist() == [0, 0, 0, 0]


def test_score(orc):
    s = load("scoring_cases.npz")
    for i, u in enumerate(s["ubis"]):
        for j, t in enumerate(s["tols"]):
            assert orc.score(u, s["gv"], t) == s["score"][i, j]


def test_score_and_refine_three_calls(orc):
    s = load("scoring_cases.npz")
    for i, u0 in enumerate(s["ubis"]):
        for j, t in enumerate(s["tols"]):
            u = u0.copy()
            for k in range(3):
                n, m = orc.score_and_refine(u, s["gv"], t)
                assert n == s["sar_n"][i, j, k]
                assert m == s["sar_mean"][i, j, k]
                assert np.array_equal(u, s["sar_ubi"][i, j, k])


@pytest.mark.parametrize("conv,d0,l0", [("rg", 1.0, -1), ("nb", 1.0, 0), ("fop", 2.0, -1)])
@pytest.mark.parametrize("tname,tol", [("t05", 0.05), ("t2", 0.2)])
def test_score_and_assign_sequences(orc, conv, d0, l0, tname, tol):
    s = load("scoring_cases.npz")
    gv, ubis = s["gv"], s["ubis"]
    d = np.full(len(gv), d0)
    lab = np.full(len(gv), l0, np.int32)
    ns = [orc.score_and_assign(u, gv, tol, d, lab, g) for g, u in enumerate(ubis)]
    ns.append(orc.score_and_assign(ubis[3], gv, tol, d, lab, 2))
    key = "saa_%s_%s" % (conv, tname)
    assert ns == s[key + "_n"].tolist()
    assert np.array_equal(d, s[key + "_drlv2"])
    assert np.array_equal(lab, s[key + "_labels"])


def test_refine_assigned(orc):
    s = load("scoring_cases.npz")
    lab = s["saa_rg_t2_labels"]
    for g, u0 in enumerate(s["ubis"]):
        u = u0.copy()
        n, m = orc.refine_assigned(u, s["gv"], lab, g)
        assert n == s["ra_n"][g] and m == s["ra_mean"][g]
        assert np.array_equal(u, s["ra_ubi"][g])


def test_compute_xlylzl(orc):
    s = load("geometry_cases.npz")
    for r, want in zip(s["xl_r"], s["xl_out"]):
        xl = np.zeros_like(want)
        orc.compute_xlylzl(s["xl_sc"], s["xl_fc"], s["xl_p"], r, s["xl_dist"], xl)
        assert np.array_equal(xl, want)


def test_compute_gv(orc):
    s = load("geometry_cases.npz")
    outs = iter(s["gv_out"])
    for wedge, chi, sign, wvln, tx, ty, tz in s["gv_sets"]:
        for x, w in ((s["gv_xyz"], s["gv_omega"]), (s["gv_xyz_rand"], s["gv_omega_rand"])):
            g = np.zeros((len(w), 3))
            orc.compute_gv(x, w, sign, wvln, wedge, chi, np.array([tx, ty, tz]), g)
            assert np.array_equal(g, next(outs))


def test_small_workload_pipeline(orc):
    w = load("workload_small.npz")
    n = len(w["sc"])
    xl = np.zeros((n, 3))
    orc.compute_xlylzl(w["sc"], w["fc"], w["p"], w["r"], w["dist"], xl)
    assert np.array_equal(xl, w["xl"])
    sign, wvln, wedge, chi = w["gonio"]
    gv = np.zeros((n, 3))
    orc.compute_gv(xl, w["omega"], sign, wvln, wedge, chi, np.zeros(3), gv)
    assert np.array_equal(gv, w["gv"])
    assert np.array_equal(orc.score_many(w["ubis"], gv, 0.05), w["counts"])
    # assignlabels with per-grain translations, via the trig-given driver
    so, co = orc.sincos_omega(w["omega"], sign)
    mat = orc.wedge_chi_matrix(wedge, chi)
    d = np.ones(n)
    lab = np.full(n, -1, np.int32)
    cnt = orc.gv_assign_many(xl, so, co, mat, wvln, w["ubis"], w["trans"],
                             np.arange(len(w["ubis"])), 0.05, d, lab)
    assert np.array_equal(cnt, w["assign_n"])
    assert np.array_equal(d, w["assign_drlv2"])
    assert np.array_equal(lab, w["assign_labels"])


def test_ken_simul_counts(orc):
    """fight_over_peaks + refine_assigned on the reference's ken_simul data
    reproduce the per-grain Npeaks recorded in its test.idx exactly."""
    k = load("ken_simul.npz")
    gv, ubis = k["gv"], k["ubis"]
    d = np.full(len(gv), 2.0)
    lab = np.full(len(gv), -1, np.int32)
    ns = [orc.score_and_assign(u, gv, 0.01, d, lab, i) for i, u in enumerate(ubis)]
    assert ns == k["fop_n"].tolist()
    assert np.array_equal(lab, k["fop_labels"]) and np.array_equal(d, k["fop_drlv2"])
    ra = [orc.refine_assigned(u.copy(), gv, lab, i) for i, u in enumerate(ubis)]
    assert [a for a, _ in ra] == k["idx_npeaks"].tolist()
    assert [b for _, b in ra] == k["ra_mean"].tolist()
