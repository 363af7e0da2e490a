"""The native decision loop of indexer.scorethem (gs_index_decide, host
code in libgscore) against the oracle's restatement of the reference's
loop (ImageD11/indexing.py:778-838), on synthetic hits: the same accepted
hits in the same order, the same peak labels, scores, hits left and count
of non-unique rejections.  No device work: it runs without a GPU."""
import ctypes

import numpy as np
import pytest

from imaged11_amd import _lib
from oracle import indexer_oracle as IO


def _native(hits, ga, trial, slot, lists, off, minpks, uniqueness, max_grains, label_base):
    lib = _lib.load()
    nh = len(hits)
    hi = np.ascontiguousarray([h[1] for h in hits], np.int32)
    hj = np.ascontiguousarray([h[2] for h in hits], np.int32)
    npk = np.ascontiguousarray(trial, np.int32)
    slot = np.ascontiguousarray(slot, np.int32)
    lists = np.ascontiguousarray(lists if len(lists) else np.zeros(1), np.int32)
    off = np.ascontiguousarray(off, np.int64)
    ga = np.ascontiguousarray(ga, np.int32)
    acc = np.empty(max(nh, 1), np.int32)
    na, stop, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    a = lambda x: x.ctypes.data
    rc = lib.gs_index_decide(nh, a(hi), a(hj), a(npk), minpks, a(slot), a(lists), a(off),
                             len(off) - 1, uniqueness, max_grains, label_base, a(ga), len(ga),
                             a(acc), ctypes.byref(na), ctypes.byref(stop), ctypes.byref(nu))
    assert rc == 0, lib.gs_last_error()
    return list(acc[:na.value]), ga, stop.value, nu.value


def _case(seed, npeaks=400, nh=300, minpks=5, uniqueness=0.5, max_grains=10 ** 6,
          label_base=0, empty_every=0):
    rng = np.random.default_rng(seed)
    hits = [[0.0, int(rng.integers(npeaks)), int(rng.integers(npeaks))] for _ in range(nh)]
    for k in rng.choice(nh, nh // 20, replace=False):  # some i == j hits
        hits[k][2] = hits[k][1]
    trial = rng.integers(0, 3 * minpks, nh)
    ga = np.full(npeaks, -1, np.int32)
    ga[rng.choice(npeaks, npeaks // 10, replace=False)] = 99  # assigned before
    cand = np.nonzero((trial > minpks) & (np.array([h[1] != h[2] for h in hits])))[0]
    slot = np.full(nh, -1, np.int64)
    slot[cand] = np.arange(len(cand))
    lists, off, ubis, npk = [], [0], [], []
    for c, h in enumerate(cand):
        n = 0 if empty_every and c % empty_every == 0 else int(rng.integers(3, 40))
        lists.extend(sorted(rng.choice(npeaks, n, replace=False).tolist()))
        off.append(len(lists))
        ubis.append(rng.normal(size=(3, 3)))
        npk.append(int(trial[h]) + int(rng.integers(0, 3)))
    lists, off = np.array(lists, np.int64), np.array(off, np.int64)

    def follow(h, _npk):
        c = slot[h]
        return ubis[c], npk[c], lists[off[c]:off[c + 1]]
    want_ga = ga.copy()
    w_ubis, w_scores, w_acc, w_left, w_nu = IO.decide(
        hits, want_ga, lambda h: int(trial[h]), follow, minpks, uniqueness, max_grains,
        label_base)
    acc, got_ga, stop, nu = _native(hits, ga, trial, slot, lists, off, minpks, uniqueness,
                                    max_grains, label_base)
    assert acc == w_acc
    assert np.array_equal(got_ga, want_ga)
    assert stop == len(w_left)
    assert nu == w_nu
    assert [npk[slot[h]] for h in acc] == w_scores
    return len(acc)


@pytest.mark.parametrize("seed", range(6))
def test_decide_matches_reference_loop(seed):
    assert _case(seed) > 0


def test_decide_stops_at_max_grains_and_labels_from_base():
    n = _case(11, max_grains=3, label_base=7)
    assert n == 3


def test_decide_empty_lists_never_unique():
    _case(12, empty_every=2)


@pytest.mark.parametrize("uniqueness", [0.0, 0.9, 1.0])
def test_decide_uniqueness_thresholds(uniqueness):
    _case(13, uniqueness=uniqueness)


def test_decide_rejects_bad_peak_numbers():
    lib = _lib.load()
    hits = [[0.0, 0, 5]]
    hi, hj = np.array([0], np.int32), np.array([5], np.int32)
    npk, slot = np.array([10], np.int32), np.array([0], np.int32)
    lists, off = np.array([1, 2], np.int32), np.array([0, 2], np.int64)
    ga = np.full(4, -1, np.int32)  # peak 5 is outside [0, 4)
    acc = np.empty(1, np.int32)
    na, stop, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    a = lambda x: x.ctypes.data
    rc = lib.gs_index_decide(1, a(hi), a(hj), a(npk), 5, a(slot), a(lists), a(off), 1, 0.5, 10,
                             0, a(ga), 4, a(acc), ctypes.byref(na), ctypes.byref(stop),
                             ctypes.byref(nu))
    assert rc != 0 and b"outside" in lib.gs_last_error()
    lib.gs_clear_error()
    assert hits


def test_oracle_refine_fixed_point_on_a_noiseless_lattice():
    """indexer.refine's restatement: on exact lattice g-vectors the refined
    UBI is the input, every ring peak is scored and the fit is ~0; peaks
    off the rings (ra = -1) are not used."""
    from imaged11_amd import simulate
    ubi, gv, _ = simulate.lattice_gvectors(seed=21, n_lattice=500, n_random=0, noise=0.0)
    ra = np.zeros(len(gv), np.int32)
    ra[::7] = -1
    u, n, fit = IO.refine(gv, ra, ubi, 0.01)
    np.testing.assert_allclose(u, ubi, rtol=0, atol=1e-12)
    assert n == int((ra > -1).sum())
    assert fit < 1e-12
