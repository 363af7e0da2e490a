"""Batched GPU indexer: ``ImageD11/indexing.py``'s ``indexer`` for the
ring-pair search (``assigntorings``, ``find``, ``scorethem``, ``score``,
``getind``).

The reference runs ``find`` as a Python loop over ring-1 peaks, each a
numpy dot product plus one ``closest`` call, and ``scorethem`` as a Python
loop over the hits, each an ``orient`` plus one ``score`` over all peaks.
Here:

* ``assigntorings`` is one kernel (``gs_ring_assign``);
* ``find`` is one kernel over all ring-1 x ring-2 peak pairs
  (``gs_find_hits``, or ``gs_find_all`` for a negative cosine tolerance);
* ``scorethem`` builds every hit's trial UBI in one kernel
  (``gs_orient_hits``) and scores all of them in one launch of the
  many-grain scorer (``gs_score_many``).  The follow-up of every hit above
  ``minpks`` (orientation variants within the cosine tolerance, the best
  of them, ``score_and_refine``, the ``getind`` peak list) does not depend
  on earlier decisions either, so it is also computed for all such hits at
  once: variants by ``gs_orient_hits``, their scores by ``gs_score_many``,
  the refinement as one ``gs_refine_lists`` iteration on the peak lists
  ``gs_select_many`` extracts, and the ``getind`` lists by
  ``gs_select_many`` again.  The reference's decision loop then visits the
  hits in its order (last first), skips the same ones and applies the
  uniqueness rule to those precomputed results, in native host code
  (``gs_index_decide``: the loop is sequential by definition).

Results (hits, grains, scores, peak assignments) are the reference's for
the same g-vectors, unit-cell tables and parameters; see the parity tests.
"""
import logging
import math

import numpy as np
import torch

from . import _lib, batched, cImageD11
from ._lib import dptr, stream_ptr
from . import unitcell as unitcell_mod


def _dev_f64(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


def _dev_i32(a, device):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


class HitList:
    """The hits of a search: the reference's list of [diff, i, j] rows
    (cosine difference, ring-1 peak, ring-2 peak; ``ImageD11/indexing.py:
    693-728``), held as three numpy columns so that the batched steps take
    them without a Python loop.  Indexing a row gives the reference's
    [float, int, int] list; a slice or an index array gives a HitList; it
    compares equal to a list of the same rows and supports the list
    operations the reference applies to its hits (len, iteration, pop,
    append)."""

    __slots__ = ("diff", "i", "j")

    def __init__(self, diff=(), i=(), j=()):
        self.diff = np.ascontiguousarray(diff, dtype=np.float64)
        self.i = np.ascontiguousarray(i, dtype=np.int64)
        self.j = np.ascontiguousarray(j, dtype=np.int64)
        if not (self.diff.shape == self.i.shape == self.j.shape and self.i.ndim == 1):
            raise ValueError("HitList: three 1-d columns of one length")

    @classmethod
    def from_rows(cls, rows):
        rows = list(rows)
        return cls([r[0] for r in rows], [r[1] for r in rows], [r[2] for r in rows])

    def __len__(self):
        return self.i.shape[0]

    def __getitem__(self, k):
        if isinstance(k, (slice, np.ndarray, list)):
            return HitList(self.diff[k], self.i[k], self.j[k])
        return [float(self.diff[k]), int(self.i[k]), int(self.j[k])]

    def __iter__(self):
        return iter(self.tolist())

    def tolist(self):
        return [list(r) for r in zip(self.diff.tolist(), self.i.tolist(), self.j.tolist())]

    def __eq__(self, other):
        if isinstance(other, HitList):
            return (np.array_equal(self.diff, other.diff) and np.array_equal(self.i, other.i)
                    and np.array_equal(self.j, other.j))
        try:
            return self.tolist() == [list(r) for r in other]
        except TypeError:
            return NotImplemented

    __hash__ = None

    def pop(self):
        if len(self) == 0:
            raise IndexError("pop from empty HitList")
        row = self[len(self) - 1]
        self.diff, self.i, self.j = self.diff[:-1], self.i[:-1], self.j[:-1]
        return row

    def append(self, row):
        d, i, j = row
        self.diff = np.append(self.diff, float(d))
        self.i = np.append(self.i, int(i))
        self.j = np.append(self.j, int(j))

    def __repr__(self):
        return "HitList(%r)" % (self.tolist(),)


def _hit_columns(hits):
    """(ring-1 peaks, ring-2 peaks) of a HitList or of a list of rows."""
    if isinstance(hits, HitList):
        return hits.i, hits.j
    return (np.array([h[1] for h in hits], dtype=np.int64),
            np.array([h[2] for h in hits], dtype=np.int64))


def ring_cosines(uc, ring_1, ring_2):
    """Distinct cosines between the hkls of two rings, largest first, with
    +-1 removed and values within 1e-5 merged (ImageD11/indexing.py:
    656-676)."""
    h1s = uc.ringhkls[uc.ringds[int(ring_1)]]
    h2s = uc.ringhkls[uc.ringds[int(ring_2)]]
    cos = sorted(uc.anglehkls(a, b)[1] for a in h1s for b in h2s)
    out = []
    while cos:
        c = cos.pop()
        if abs(c - 1.0) < 1e-5 or abs(c + 1.0) < 1e-5:
            continue
        if not out or abs(out[-1] - c) > 1e-5:
            out.append(c)
    return out


def find_hits(gv, i1, i2, cs, cosine_tol):
    """The hits of indexer.find for ring-1 peaks i1 and ring-2 peaks i2
    (ImageD11/indexing.py:693-728) as a HitList of [diff, i, j] rows in the
    reference's order.  gv: (N,3) float64 device tensor; i1, i2: int index
    arrays."""
    lib = _lib.load()
    _lib.require_gpu()
    dev = gv.device
    i1 = np.asarray(i1, dtype=np.int64)
    i2 = np.asarray(i2, dtype=np.int64)
    cs = np.ascontiguousarray(cs, dtype=np.float64)
    N1, N2, nv = len(i1), len(i2), len(cs)
    if N1 == 0 or N2 == 0:
        return HitList()
    n1 = torch.empty((N1, 3), dtype=torch.float64, device=dev)
    n2 = torch.empty((N2, 3), dtype=torch.float64, device=dev)
    d1, d2 = _dev_i32(i1, dev), _dev_i32(i2, dev)
    N = gv.shape[0]
    _lib.check(lib.gs_unit_rows(dptr(gv), N, dptr(d1), N1, dptr(n1), stream_ptr()), "unit_rows")
    _lib.check(lib.gs_unit_rows(dptr(gv), N, dptr(d2), N2, dptr(n2), stream_ptr()), "unit_rows")
    dcs = _dev_f64(cs, dev)
    tol = float(cosine_tol)
    if tol > 0:
        ib = torch.empty(N1, dtype=torch.int32, device=dev)
        best = torch.empty(N1, dtype=torch.float64, device=dev)
        _lib.check(lib.gs_find_hits(dptr(n1), N1, dptr(n2), N2, dptr(dcs), nv, dptr(ib),
                                    dptr(best), stream_ptr()), "find_hits")
        ib, best = ib.cpu().numpy(), best.cpu().numpy()
        sel = np.nonzero(best < tol)[0]
        return HitList(best[sel], i1[sel], i2[ib[sel]])
    mtol = -tol
    counter = torch.zeros(1, dtype=torch.int64, device=dev)
    _lib.check(lib.gs_find_all(dptr(n1), N1, dptr(n2), N2, dptr(dcs), nv, mtol, dptr(counter),
                               None, 0, stream_ptr()), "find_all")
    n = int(counter.item())
    if n == 0:
        return HitList()
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    _lib.check(lib.gs_find_all(dptr(n1), N1, dptr(n2), N2, dptr(dcs), nv, mtol, dptr(counter),
                               dptr(keys), n, stream_ptr()), "find_all")
    keys = torch.sort(keys).values.cpu().numpy()
    j = keys % N2
    i = keys // (nv * N2)
    return HitList(np.zeros(n), i1[i], i2[j])


def orient_hits(uc, ring_1, ring_2, gv, hits, entries=None, with_cosines=False):
    """Trial UBI of every hit (a HitList or rows [diff, i, j]) as unitcell.orient(ring_1,
    gv[i], ring_2, gv[j]) makes it with crange <= 0, in one kernel; with
    `entries`, the UBI from table entry entries[h] instead (the crange > 0
    variants).  Returns (nh,3,3) UBIs, and the observed cosines when
    with_cosines."""
    lib = _lib.load()
    dev = gv.device
    _, c2ab, matrs = uc.getanglehkls(ring_1, ring_2)
    ci, cj = _hit_columns(hits)
    hi, hj = _dev_i32(ci, dev), _dev_i32(cj, dev)
    nh = len(hits)
    ubis = torch.empty((nh, 3, 3), dtype=torch.float64, device=dev)
    cos = torch.empty(nh, dtype=torch.float64, device=dev) if with_cosines else None
    bsel = _dev_i32(entries, dev) if entries is not None else None
    dc = _dev_f64(np.asarray(c2ab, dtype=np.float64), dev)
    dbt = _dev_f64(np.asarray(matrs, dtype=np.float64).reshape(-1, 3, 3), dev)
    if nh:
        _lib.check(lib.gs_orient_hits(dptr(gv), gv.shape[0], dptr(hi), dptr(hj), nh, dptr(dc),
                                      len(c2ab), dptr(dbt),
                                      dptr(bsel) if bsel is not None else None, dptr(ubis),
                                      None, dptr(cos) if cos is not None else None,
                                      stream_ptr()), "orient_hits")
    return (ubis, cos) if with_cosines else ubis


def select_lists(ubis, gv, thr, counts=None, sort=True):
    """CSR lists of the peaks each UBI indexes with the reference
    arithmetic (drlv2 < thr).  ubis (K,3,3) and gv (N,3) device tensors;
    returns (idx int32, off int64 (K+1,)) on the device.

    ``gs_select_many`` counts in one pass and writes the lists in a second.
    counts (host ints, one per UBI): the expected list lengths, e.g. the
    scores of the same UBIs at the same tolerance, or upper bounds; the
    lists are then written in the first pass into that much room, and the
    pass's own counts decide: equal counts are the result, smaller ones
    are packed together on the device, and a longer list than its room
    (a bound too small) sends the whole call to a second pass with the
    counts the first one found.  sort=True puts each list in peak order
    (the order of the reference's selection, and a fixed order for the
    sums of a refinement); sort=False leaves the order of the atomic slots,
    for callers that use a list as a set."""
    lib = _lib.load()
    dev = gv.device
    K, N = ubis.shape[0], gv.shape[0]
    cnt = torch.empty(K, dtype=torch.int32, device=dev)

    def fill(off, total, count=False):
        # count=True: the pass also gives the counts (cnt), so it runs
        # even when there is no room at all
        idx = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        if total or count:
            _lib.check(lib.gs_select_many(dptr(ubis), K, dptr(gv), N, float(thr), dptr(off),
                                          dptr(cnt), dptr(idx), stream_ptr()), "select_many")
        return idx

    def offsets(c):
        off_h = np.zeros(K + 1, np.int64)
        np.cumsum(c, out=off_h[1:])
        return torch.as_tensor(off_h, device=dev), int(off_h[-1])

    room = None if counts is None else np.asarray(counts, dtype=np.int64).reshape(K)
    if room is not None and K:
        off, total = offsets(room)
        idx = fill(off, total, count=True)
        got = cnt.cpu().numpy().astype(np.int64)
        if np.array_equal(got, room):
            pass
        elif np.all(got <= room):
            # pack: entry q of list g moves from off[g] + q to new_off[g] + q
            new_off, new_total = offsets(got)
            owner = torch.repeat_interleave(torch.arange(K, device=dev),
                                            torch.as_tensor(room, device=dev), output_size=total)
            q = torch.arange(total, device=dev) - off[owner]
            dest = torch.where(q < cnt.to(torch.int64)[owner], new_off[owner] + q,
                               torch.full_like(q, new_total))
            packed = torch.empty(new_total + 1, dtype=torch.int32, device=dev)
            packed[dest] = idx[:total]
            idx, off, total = packed, new_off, new_total
        else:
            off, total = offsets(got)
            idx = fill(off, total)
    else:
        _lib.check(lib.gs_select_many(dptr(ubis), K, dptr(gv), N, float(thr), None, dptr(cnt),
                                      None, stream_ptr()), "select_many")
        off = torch.zeros(K + 1, dtype=torch.int64, device=dev)
        torch.cumsum(cnt.to(torch.int64), 0, out=off[1:])
        total = int(off[-1].item())
        idx = fill(off, total)
    if sort and total:
        owner = torch.repeat_interleave(torch.arange(K, device=dev, dtype=torch.int64),
                                        (off[1:] - off[:-1]), output_size=total)
        key = torch.sort(owner * N + idx[:total].to(torch.int64)).values
        idx = (key - owner * N).to(torch.int32)
    return idx[:total], off


class _Followups:
    """Every qualifying hit's follow-up: slot[h] = its row (or -1), and per
    row the refined UBI, the score, and the getind peak list idx[off[c] ..
    off[c+1])."""

    def __init__(self, slot, ubi, npk, idx, off):
        self.slot, self.ubi, self.npk, self.idx, self.off = slot, ubi, npk, idx, off

    def __call__(self, h, _trial=None, _npk=None):
        c = self.slot[h]
        return self.ubi[c].copy(), int(self.npk[c]), self.idx[self.off[c]:self.off[c + 1]]


class indexer:
    """Finds grain orientations from pairs of peaks on two powder rings
    (ImageD11/indexing.py:313-1183, the parts on the search path).  Same
    parameters and attributes: gv, ds, ra, ga, hits, ubis, scores."""

    def __init__(self, unitcell=None, gv=None, cosine_tol=0.002, minpks=10, hkl_tol=0.01,
                 ring_1=1, ring_2=2, ds_tol=0.005, wavelength=-1, uniqueness=0.5,
                 eta_range=0.0, max_grains=100, device="cuda"):
        self.device = torch.device(device)
        self.unitcell = unitcell
        self.gv = None
        self.ra = None
        if gv is not None:
            gv = np.asarray(gv)
            assert gv.shape[1] == 3
            self.gv = gv.astype(float)
            self.ds = np.sqrt((gv * gv).sum(axis=1))
            self.ga = np.zeros(len(self.ds), np.int32) - 1
            self.gvflat = np.ascontiguousarray(gv, float)
            self.wedge = 0.0
        self.cosine_tol = cosine_tol
        self.wavelength = wavelength
        self.hkl_tol = hkl_tol
        self.ring_1 = ring_1
        self.ring_2 = ring_2
        self.uniqueness = uniqueness
        self.minpks = minpks
        self.ds_tol = ds_tol
        self.max_grains = max_grains
        self.eta_range = eta_range
        self.ubis = []
        self.scores = []
        self.hits = HitList()
        self.omega = None
        self.wedge = 0.0
        self._gv_dev = None

    # parameters given as strings, as from ImageD11's parameter files
    _TYPES = {"cosine_tol": float, "hkl_tol": float, "ds_tol": float, "wavelength": float,
              "uniqueness": float, "eta_range": float, "minpks": int, "ring_1": int,
              "ring_2": int, "max_grains": int}

    def set_parameters(self, pars):
        """Set search parameters from a dict (values may be strings)."""
        for k, v in pars.items():
            if k in self._TYPES:
                setattr(self, k, self._TYPES[k](v))

    def _gvd(self):
        if self._gv_dev is None or self._gv_dev.shape[0] != len(self.gv):
            self._gv_dev = _dev_f64(self.gv, self.device)
        return self._gv_dev

    def out_of_eta_range(self, eta):
        """True for peaks within eta_range of eta = 0 or +-180
        (ImageD11/indexing.py:426-434); eta is first brought into
        [-180, 180] by whole turns as mod_360(eta, 0) does (:172-183)."""
        e = float(eta)
        while e < -180:
            e += 360
        while e > 180:
            e -= 360
        er = abs(self.eta_range)
        if -er < e < er:
            return True
        if e < -180.0 + er or e > 180.0 - er:
            return True
        return False

    def readgvfile(self, filename, quiet=True):
        """Read a .gve file: cell line, '#' header (wavelength, wedge), then
        rows xr yr zr xc yc ds eta omega (ImageD11/indexing.py:1184-1248)."""
        with open(filename) as f:
            self.unitcell = unitcell_mod.cellfromstring(f.readline())
            while True:
                line = f.readline()
                if line == "":
                    raise ValueError("%s: no column header line" % filename)
                if line[0] != "#":
                    continue
                if "wavelength" in line:
                    self.wavelength = float(line.split()[-1])
                    continue
                if "wedge" in line:
                    self.wedge = float(line.split()[-1])
                    continue
                if "ds h k l" in line:
                    continue
                if "omega" in line and "xc" in line:
                    break
            rows = []
            for line in f:
                v = [float(x) for x in line.split()]
                if len(v) == 0:
                    continue
                if self.out_of_eta_range(v[6]):
                    continue
                rows.append(v[:8])
        rows = np.array(rows, float).reshape(-1, 8)
        self.xr, self.yr, self.zr = rows[:, 0], rows[:, 1], rows[:, 2]
        self.xp, self.yp = rows[:, 3], rows[:, 4]
        self.ds = rows[:, 5].copy()
        self.eta = rows[:, 6].copy()
        self.omega = rows[:, 7].copy()
        self.omega_fullrange = find_omega_ranges(self.omega)
        if self.wavelength > 0:
            self.tth = np.arcsin(self.ds * self.wavelength / 2) * 360 / math.pi
        else:
            self.tth = np.zeros(len(self.ds))
        self.gv = np.ascontiguousarray(rows[:, :3])
        self.allgv = self.gv.copy()
        self.ga = np.zeros(len(self.ds), np.int32) - 1
        self.gvflat = np.ascontiguousarray(self.gv, float)
        self.gv = self.gvflat.copy()
        self._gv_dev = None

    def assigntorings(self):
        """ra[k] = nearest ring within ds_tol, -1 otherwise
        (ImageD11/indexing.py:437-457), on the GPU."""
        lib = _lib.load()
        _lib.require_gpu()
        limit = np.amax(self.ds)
        self.unitcell.makerings(limit, tol=self.ds_tol)
        dsr = np.asarray(self.unitcell.ringds, dtype=np.float64)
        n = len(self.ds)
        ds = _dev_f64(self.ds, self.device)
        rd = _dev_f64(dsr, self.device)
        ra = torch.empty(n, dtype=torch.int32, device=self.device)
        _lib.check(lib.gs_ring_assign(dptr(ds), n, dptr(rd), len(dsr), float(self.ds_tol),
                                      dptr(ra), stream_ptr()), "ring_assign")
        self.ra = ra.cpu().numpy()
        self.na = np.bincount(self.ra[self.ra >= 0], minlength=len(dsr)).astype(np.int32)

    def find(self):
        """Candidate peak pairs (ImageD11/indexing.py:633-728)."""
        if self.ra is None:
            self.assigntorings()
        iall = np.arange(self.gv.shape[0])
        i1 = iall[(self.ra == self.ring_1) & (self.ga == -1)]
        i2 = iall[(self.ra == self.ring_2) & (self.ga == -1)]
        self.hits = HitList()
        if len(i1) == 0 or len(i2) == 0:
            logging.info("no peaks left for those rings")
            return
        cs = ring_cosines(self.unitcell, self.ring_1, self.ring_2)
        self.cosangles = np.array(cs, "d")
        self.hits = find_hits(self._gvd(), i1, i2, self.cosangles, self.cosine_tol)
        logging.info("Number of trial orientations generated %d", len(self.hits))

    def score(self, UBI, tol=None):
        return cImageD11.score(UBI, self._gvd(), self.hkl_tol if tol is None else tol)

    def getind(self, UBI, tol=None, drlv2tmp=None, labelstmp=None):
        """Boolean mask of the peaks UBI indexes: score_and_assign from
        drlv2 = 1, labels = 0, label 1 (ImageD11/indexing.py:1060-1088)."""
        if tol is None:
            tol = self.hkl_tol
        g = self._gvd()
        n = g.shape[0]
        drlv2 = torch.ones(n, dtype=torch.float64, device=g.device)
        labels = torch.zeros(n, dtype=torch.int32, device=g.device)
        cImageD11.score_and_assign(UBI, g, tol, drlv2, labels, 1)
        return (labels == 1).cpu().numpy()

    def scorethem(self, fitb4=False, batched_followup=True):
        """Decide which hits become grains (ImageD11/indexing.py:762-874).

        Every hit's trial UBI is built and scored in two launches.  With
        batched_followup (default) the follow-up of every hit above minpks
        (orientation variants, score_and_refine, the getind peak list) is
        also computed for all of them at once, since none of it depends on
        earlier decisions; the reference's decision loop (gs_index_decide,
        native) then only reads those results.  Ends, as the reference does,
        with `refine` of the best grain (scorelastrefined, fitlastrefined).  batched_followup=False runs the follow-ups one hit
        at a time inside the loop, as the reference does."""
        if self.hits is None or len(self.hits) == 0:
            logging.info("No hits to score")
            return
        if fitb4:
            raise NotImplementedError("fitb4 (marked as not working in the reference)")
        g = self._gvd()
        tol = float(self.hkl_tol)
        trial = orient_hits(self.unitcell, self.ring_1, self.ring_2, g, self.hits)
        npk_all = batched.score_many(trial, g, tol).cpu().numpy()
        if batched_followup:
            follow = self._followups(g, tol, npk_all)
            self._decide(follow, npk_all)
            self._refine_best()
            return
        follow = self._followup_one
        trial = trial.cpu().numpy()
        ng = 0
        self.nuniq = 0
        h = len(self.hits)
        while h > 0 and ng < self.max_grains:
            h -= 1
            _, i, j = self.hits[h]
            if self.ga[i] > -1 or self.ga[j] > -1 or i == j:
                continue
            if npk_all[h] <= self.minpks:
                continue
            UBI, npk, ind = follow(h, trial[h], int(npk_all[h]))
            ga = self.ga[ind]
            with np.errstate(invalid="ignore", divide="ignore"):
                uniq = np.sum(np.where(ga == -1, 1, 0)) * 1.0 / ga.shape[0]
            if uniq > self.uniqueness:
                self.ga[ind] = len(self.scores) + 1
                self.ubis.append(UBI)
                self.scores.append(npk)
                ng += 1
            else:
                self.nuniq += 1
        # the reference pops every visited hit off its list
        self.hits = self.hits[:h]
        self._refine_best()

    def _decide(self, follow, npk_all):
        """The decision loop over precomputed follow-ups, in native code
        (gs_index_decide): the same visiting order, skips and uniqueness
        rule as the reference's loop (ImageD11/indexing.py:778-838)."""
        import ctypes
        lib = _lib.load()
        hits = self.hits
        nh = len(hits)
        ci, cj = _hit_columns(hits)
        hi = np.ascontiguousarray(ci, dtype=np.int32)
        hj = np.ascontiguousarray(cj, dtype=np.int32)
        npk_h = np.ascontiguousarray(npk_all, dtype=np.int32)
        if follow is None:
            slot = np.full(nh, -1, np.int32)
            lists, off, ncand = np.zeros(1, np.int32), np.zeros(1, np.int64), 0
        else:
            slot = np.ascontiguousarray(follow.slot, dtype=np.int32)
            lists = np.ascontiguousarray(follow.idx, dtype=np.int32)
            off = np.ascontiguousarray(follow.off, dtype=np.int64)
            ncand = len(off) - 1
        ga = np.ascontiguousarray(self.ga, dtype=np.int32)
        accepted = np.empty(max(nh, 1), np.int32)
        na, stop, nu = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.gs_clear_error()
        a = lambda x: x.ctypes.data
        _lib.check(lib.gs_index_decide(
            nh, a(hi), a(hj), a(npk_h), int(self.minpks), a(slot), a(lists), a(off), ncand,
            float(self.uniqueness), int(min(self.max_grains, 2 ** 31 - 1)), len(self.scores),
            a(ga), len(ga), a(accepted), ctypes.byref(na), ctypes.byref(stop),
            ctypes.byref(nu)), "index_decide")
        self.ga[:] = ga
        if na.value:
            rows = follow.slot[accepted[:na.value]]
            self.ubis.extend(follow.ubi[rows])  # fancy indexing copies: one (3,3) array each
            self.scores.extend(follow.npk[rows].tolist())
        self.nuniq = nu.value
        # the reference pops every visited hit off its list
        self.hits = hits[:stop.value]

    def _refine_best(self):
        """The end of the reference's scorethem (ImageD11/indexing.py:
        839-874): refine the best-scoring grain for the log."""
        logging.info("Number of orientations with more than %d peaks is %d",
                     self.minpks, len(self.ubis))
        if len(self.ubis) > 0:
            best = int(np.argmax(self.scores))
            self.refine(self.ubis[best])
            logging.info("Indexes %d peaks, with <drlv2>=%f", self.scorelastrefined,
                         self.fitlastrefined)

    def refine(self, UBI):
        """Refine an orientation on the ring-assigned peaks it indexes and
        rescore it (ImageD11/indexing.py:1099-1167): R = sum g h^T and H =
        sum h h^T over the peaks with drlv2 < hkl_tol^2 on a ring, UB =
        R H^-1, then scorelastrefined / fitlastrefined of the new UBI.
        Returns the refined UBI (the input when H or UB is singular).

        The reference rounds with floor(h + 1/2), the kernels half to even;
        they differ only at exact half-integers, where drlv2 >= 1/4 exceeds
        any hkl_tol^2 below 1/4, so the selected peaks are the same."""
        g = self._gvd()
        tol = float(self.hkl_tol)
        if not tol * tol < 0.25:
            raise ValueError("indexer.refine needs hkl_tol < 0.5")
        ring = torch.as_tensor(np.asarray(self.ra) > -1, device=g.device)
        u = torch.as_tensor(np.array(UBI, dtype=np.float64).reshape(1, 3, 3), device=g.device)

        def on_rings(ubi):
            idx, off = select_lists(ubi, g, tol * tol)
            idx = idx[ring[idx.long()]]
            return idx, torch.tensor([0, idx.numel()], dtype=torch.int64, device=g.device)
        idx, off = on_rings(u)
        if idx.numel() == 0:
            raise Exception("No contributing reflections for" + str(UBI))
        uo, _, _ = batched.refine_lists(u, g, idx, off, tol, n_iter=1)
        idx, off = on_rings(uo)
        self.scorelastrefined = int(idx.numel())
        if idx.numel() == 0:
            raise ValueError("No contributing reflections for ubi:\n{}\nafter "
                             "refinement".format(UBI))
        _, _, mean = batched.refine_lists(uo, g, idx, off, tol, n_iter=1)
        self.fitlastrefined = float(np.sqrt(float(mean[0, 0])))
        return uo[0].cpu().numpy()

    def _followup_one(self, h, UBI, npk):
        """The reference's follow-up of one hit above minpks
        (ImageD11/indexing.py:801-837): the orientation variants within
        cosine_tol, the best-scoring one if it scores at least as well,
        score_and_refine, then getind.  Returns (UBI, npk, peak mask)."""
        g = self._gvd()
        tol = float(self.hkl_tol)
        uc = self.unitcell
        _, i, j = self.hits[h]
        UBI = UBI.copy()
        uc.orient(self.ring_1, self.gv[i, :], self.ring_2, self.gv[j, :], verbose=0,
                  crange=abs(self.cosine_tol))
        if len(uc.UBIlist) > 1:
            cand = torch.as_tensor(np.array(uc.UBIlist), device=g.device)
            npks = batched.score_many(cand, g, tol).cpu().numpy()
            choice = int(np.argmax(npks))
            if npks[choice] >= npk:
                UBI = uc.UBIlist[choice].copy()
                npk = int(npks[choice])
        cImageD11.score_and_refine(UBI, g, tol)
        return UBI, npk, self.getind(UBI)

    def _followups(self, g, tol, npk_all):
        """The follow-up of every hit above minpks at once (a _Followups,
        or None when no hit qualifies)."""
        hits = self.hits if isinstance(self.hits, HitList) else HitList.from_rows(self.hits)
        hi, hj = hits.i, hits.j
        cand = np.nonzero((npk_all > self.minpks) & (hi != hj))[0]
        slot = np.full(len(hits), -1, np.int64)
        slot[cand] = np.arange(len(cand))
        if len(cand) == 0:
            return None
        uc = self.unitcell
        chits = hits[cand]
        ubi, cos = orient_hits(uc, self.ring_1, self.ring_2, g, chits, with_cosines=True)
        ubi = ubi.cpu().numpy()
        npk = npk_all[cand].astype(np.int64)
        crange = abs(self.cosine_tol)
        if crange > 0:
            # unitcell.orient with crange: every table entry within crange,
            # reduced to distinct orientations, the best one taken when it
            # scores at least as well as the trial (ImageD11/indexing.py:
            # 803-825, ImageD11/unitcell.py:620-666)
            _, c2ab, _ = uc.getanglehkls(self.ring_1, self.ring_2)
            near = abs(np.asarray(c2ab)[None, :] - cos.cpu().numpy()[:, None]) < crange
            rows, cols = np.nonzero(near)
            var = orient_hits(uc, self.ring_1, self.ring_2, g, chits[rows],
                              entries=cols).cpu().numpy()
            starts = np.searchsorted(rows, np.arange(len(cand) + 1))
            lists, owners = [], []
            # a candidate with one variant keeps its trial orientation
            for c in np.nonzero(np.diff(starts) > 1)[0]:
                vs = list(var[starts[c]:starts[c + 1]])
                if len(vs) > 1:
                    vs = unitcell_mod.ubi_equiv(vs, [np.linalg.inv(v) for v in vs])
                if len(vs) > 1:
                    lists.extend(vs)
                    owners.extend([c] * len(vs))
            if lists:
                owners = np.array(owners)
                sc = batched.score_many(torch.as_tensor(np.array(lists), device=g.device), g,
                                        tol).cpu().numpy()
                uniq, firsts, counts = np.unique(owners, return_index=True, return_counts=True)
                for c, f, n in zip(uniq, firsts, counts):
                    choice = int(np.argmax(sc[f:f + n]))
                    if sc[f + choice] >= npk[c]:
                        ubi[c] = lists[f + choice]
                        npk[c] = sc[f + choice]
        # score_and_refine of each chosen UBI over all peaks = one refine
        # iteration on the list of peaks it indexes
        dubi = torch.as_tensor(np.ascontiguousarray(ubi), device=g.device)
        # npk are the scores of these UBIs at tol, i.e. the lengths of
        # their drlv2 < tol^2 lists
        idx, off = select_lists(dubi, g, tol * tol, counts=npk)
        refined, _, _ = batched.refine_lists(dubi, g, idx, off, tol, n_iter=1)
        # getind: score_and_assign from drlv2 = 1, labels = 0, label 1; the
        # decision loop uses each list as a set, so any order does, and the
        # refined UBI's list is about as long as the trial's
        room = npk + np.maximum(16, npk // 4)
        idx2, off2 = select_lists(refined, g, min(tol * tol, 1.0), counts=room, sort=False)
        return _Followups(slot, refined.cpu().numpy(), npk, idx2.cpu().numpy(),
                          off2.cpu().numpy())

    def fight_over_peaks(self):
        """Give each peak to the grain indexing it best, then refine each
        grain on its peaks (ImageD11/indexing.py:876-905)."""
        from . import drivers
        g = self._gvd()
        ubis = torch.as_tensor(np.array(self.ubis), device=g.device)
        labels, drlv2, gas = drivers.fight_over_peaks(ubis.reshape(-1, 3, 3), g, self.hkl_tol)
        self.ga = labels.cpu().numpy()
        self.gas = gas.cpu().numpy()
        self.drlv2 = drlv2.cpu().numpy()
        return self.ga

    def score_all_pairs(self, n=None, rmulmax=None, rings_to_use=None):
        """Search every pair of rings in turn (ImageD11/indexing.py:577-631):
        the rings holding peaks (or those of rings_to_use that do), without
        those whose multiplicity exceeds rmulmax, paired with each other in
        both orders and with themselves, sorted by (product of the
        multiplicities, product of the ring occupations, ring_1, ring_2) so
        that the sparsest pairs come first; for each pair `find` and, when
        it gives hits, `scorethem`.  Each pair's peaks exclude those earlier
        pairs assigned, so the pairs run in this order; within a pair every
        hit is found, oriented and scored in one launch each.  n: stop after
        the n-th pair that gave hits (the reference's count)."""
        self.assigntorings()
        if rings_to_use is not None:
            rings = [r for r in rings_to_use if (self.ra == r).sum() > 0]
        else:
            rings = sorted(set(int(r) for r in self.ra if r >= 0))
        uc = self.unitcell
        mults = {r: len(uc.ringhkls[uc.ringds[r]]) for r in rings}
        if rmulmax is not None:
            rings = [r for r in rings if mults[r] <= rmulmax]
        pairs = sorted((int(mults[r1] * mults[r2]), int(self.na[r1] * self.na[r2]), r1, r2)
                       for r1 in rings for r2 in rings)
        self.tried = 0
        self.npairs = len(pairs)
        self.stop = False
        k = 0
        for _, _, r1, r2 in pairs:
            k += 1
            self.ring_1 = r1
            self.ring_2 = r2
            self.find()
            if len(self.hits) == 0:
                continue
            self.scorethem()
            self.tried += 1
            if self.stop:
                break
            if n is not None and k > n:
                break
            logging.info("Tried r1=%d r2=%d attempt %d of %d, got %d grains", r1, r2,
                         self.tried, len(pairs), len(self.ubis))
        logging.info("Tested %d pairs and found %d grains so far", self.tried, len(self.ubis))

    def saveubis(self, filename):
        """Write the UBIs found (ImageD11/indexing.py:1170-1174)."""
        write_ubi_file(filename, self.ubis)

    def saveindexing(self, filename, tol=None):
        """Write the indexing report (ImageD11/indexing.py:907-1058): each
        grain's peak count and mean drlv from `fight_over_peaks` followed
        by `refine_assigned` of a copy of its UBI (the UBIs are not
        changed), its UBI, cell, U and B, and per assigned peak its hkl,
        drlv, detector position and observed/calculated omega, eta, tth;
        then every ring-assigned peak with its grain.  Peak assignment and
        refine_assigned run on the GPU, the text formatting on the host."""
        from . import transform
        self.gv = np.ascontiguousarray(self.gv)
        self.fight_over_peaks()
        g = self._gvd()
        labels = torch.as_tensor(np.ascontiguousarray(self.ga, np.int32), device=g.device)
        npeaks = len(self.ra)
        allind = np.arange(npeaks)
        tthcalc = np.zeros(npeaks, float)
        etacalc = np.zeros(npeaks, float)
        omegacalc = np.zeros(npeaks, float)
        xp = getattr(self, "xp", None)
        yp = getattr(self, "yp", None)
        uinverses = []
        with open(filename, "w") as f:
            for i, ubi in enumerate(self.ubis):
                ubi = np.asarray(ubi, float)
                uinverses.append(np.linalg.inv(ubi))
                npk, mdrlv = cImageD11.refine_assigned(ubi.copy(), g, labels, i)
                assert npk == self.gas[i]
                f.write("Grain: %d   Npeaks=%d   <drlv>=%f\n" % (i, self.gas[i], np.sqrt(mdrlv)))
                f.write("UBI:\n" + str(ubi) + "\n")
                f.write("Cell pars: ")
                cellpars = ubitocellpars(ubi)
                for v in cellpars[:3]:
                    f.write("%10.6f " % v)
                for v in cellpars[3:]:
                    f.write("%10.3f " % v)
                f.write("\n")
                f.write("U:\n" + str(ubitoU(ubi)) + "\n")
                f.write("B:\n" + str(ubitoB(ubi)) + "\n")
                ind = allind[self.ga == i]
                h = np.dot(ubi, self.gv[ind].T)
                hint = np.floor(h + 0.5)
                gint = np.dot(uinverses[-1], hint)
                f.write("Peak   (  h       k       l      )   drlv             x       y ")
                if self.wavelength < 0:
                    f.write("\n")
                else:
                    f.write("   Omega_obs Omega_calc   Eta_obs Eta_calc   tth_obs tth_calc\n")
                    tc, ec, oc = transform.uncompute_g_vectors(gint, self.wavelength,
                                                               wedge=self.wedge)
                for col, j in enumerate(ind):
                    f.write("%-6d ( % 6.4f % 6.4f % 6.4f ) % 12.8f "
                            % (j, h[0, col], h[1, col], h[2, col], np.sqrt(self.drlv2[j])))
                    f.write(" % 7.1f % 7.1f " % (xp[j] if xp is not None else 0.0,
                                                 yp[j] if yp is not None else 0.0))
                    if self.wavelength < 0:
                        f.write("\n")
                    else:
                        to = math.asin(self.wavelength * self.ds[j] / 2) * 360 / math.pi
                        eo = mod_360(self.eta[j], 0)
                        oo = self.omega[j]
                        w = int(np.argmin([abs(ec[0][col] - eo), abs(ec[1][col] - eo)]))
                        ec1 = ec[w][col]
                        oc1 = mod_360(oc[w][col], oo)
                        f.write("  % 9.4f % 9.4f     % 9.4f % 9.4f   % 9.4f % 9.4f"
                                % (oo, oc1, eo, ec1, to, tc[col]))
                        etacalc[j] = ec1
                        omegacalc[j] = oc1
                        tthcalc[j] = tc[col]
                    if self.ra[j] == -1:
                        f.write(" *** was not assigned to ring\n")
                    else:
                        f.write("\n")
                f.write("\n\n")
            in_rings = allind[np.asarray(self.ra) > -1]
            f.write("\n\nAnd now listing via peaks which were assigned to rings\n")
            nleft = 0
            npk = 0
            omega = self.omega if self.omega is not None else np.zeros(npeaks)
            eta = getattr(self, "eta", np.zeros(npeaks))
            tth = getattr(self, "tth", np.zeros(npeaks))
            for peak in in_rings:
                gp = self.gv[peak, :]
                f.write("\nPeak= %-5d Ring= %-5d gv=[ % -6.4f % -6.4f % -6.4f ]   omega= % 9.4f"
                        "   eta= % 9.4f   tth= % 9.4f\n"
                        % (peak, self.ra[peak], gp[0], gp[1], gp[2], omega[peak], eta[peak],
                           tth[peak]))
                if self.ga[peak] != -1:
                    m = self.ga[peak]
                    hi = np.dot(self.ubis[m], gp)
                    hint = np.floor(hi + 0.5).astype(int)
                    f.write("Grain %-5d (%3d,%3d,%3d)" % (m, hint[0], hint[1], hint[2]))
                    f.write("  ( % -6.4f % -6.4f % -6.4f )  " % (hi[0], hi[1], hi[2]))
                    f.write(" omega= % 9.4f   eta= %9.4f   tth= %9.4f\n"
                            % (omegacalc[peak], etacalc[peak], tthcalc[peak]))
                    npk += 1
                else:
                    if len(self.ubis) > 0:
                        f.write("Peak not assigned\n")
                    else:
                        f.write("Peak not assigned, no grains found\n")
                    nleft += 1
            f.write("\n\nTotal number of peaks was %d\n" % (self.gv.shape[0]))
            f.write("Peaks assigned to grains %d\n" % npk)
            f.write("Peaks assigned to rings but remaining unindexed %d\n" % nleft)
            f.write("Peaks not assigned to rings at all %d\n"
                    % int(np.sum(np.asarray(self.ra) == -1)))


# ---- file helpers and the column-file drivers --------------------------------

def write_ubi_file(filename, ubilist):
    """3x3 matrices, one row per line ("%f %f %f"), a blank line after each
    matrix (ImageD11/indexing.py:107-114)."""
    with open(filename, "w") as f:
        for u in ubilist:
            f.write("%f %f %f\n" % (u[0][0], u[0][1], u[0][2]))
            f.write("%f %f %f\n" % (u[1][0], u[1][1], u[1][2]))
            f.write("%f %f %f\n\n" % (u[2][0], u[2][1], u[2][2]))


def readubis(ubifile):
    """The 3x3 matrices of a file written by write_ubi_file, skipping '#'
    lines (ImageD11/indexing.py:89-104)."""
    out, rows = [], []
    with open(ubifile) as f:
        for line in f:
            if line.startswith("#"):
                continue
            vals = [float(x) for x in line.split()]
            if len(vals) == 3:
                rows.append(vals)
            if len(rows) == 3:
                out.append(np.array(rows))
                rows = []
    return out


def ubitocellpars(ubi):
    """(a, b, c, alpha, beta, gamma) of a UBI (ImageD11/indexing.py:117-128)."""
    g = np.dot(ubi, np.transpose(ubi))
    a, b, c = math.sqrt(g[0, 0]), math.sqrt(g[1, 1]), math.sqrt(g[2, 2])
    return (a, b, c, math.degrees(math.acos(g[1, 2] / b / c)),
            math.degrees(math.acos(g[0, 2] / a / c)), math.degrees(math.acos(g[0, 1] / a / b)))


def ubitoU(ubi):
    """The orientation U with inv(UBI) = U.B (ImageD11/indexing.py:131-138)."""
    from .refinegrains import ubi_to_u
    return ubi_to_u(ubi)


def ubitoB(ubi):
    """B from the Cholesky factor of UBI.UBI^T (ImageD11/indexing.py:166-169)."""
    g = np.dot(ubi, np.transpose(ubi))
    return np.transpose(np.linalg.inv(np.linalg.cholesky(g)))


def mod_360(theta, target):
    """theta plus the multiple of 360 that brings it within 180 of target
    (ImageD11/indexing.py:172-183)."""
    diff = theta - target
    while diff < -180:
        theta = theta + 360
        diff = theta - target
    while diff > 180:
        theta = theta - 360
        diff = theta - target
    return theta


def find_omega_ranges(omega):
    """Whole degrees of omega that hold a peak (ImageD11/indexing.py:1254-1264)."""
    iomega = np.round(np.asarray(omega) % 360).astype(int)
    return int((np.bincount(iomega) > 0).sum())


def indexer_from_colfile(colfile, **kwds):
    """An indexer on a column file's g-vectors with the unit cell and
    wavelength of its parameters (ImageD11/indexing.py:278-288)."""
    uc = unitcell_mod.unitcell_from_parameters(colfile.parameters)
    w = float(colfile.parameters.get("wavelength"))
    gv = np.array((colfile.gx, colfile.gy, colfile.gz), float)
    kwds.update({"unitcell": uc, "wavelength": w, "gv": gv.T})
    ind = indexer(**kwds)
    if "omega" in colfile.titles:
        ind.omega_fullrange = find_omega_ranges(colfile.omega)
    ind.colfile = colfile
    return ind


def indexer_from_colfile_and_ucell(colfile, ucell, **kwds):
    """indexer_from_colfile with a given unit cell; the wavelength from
    kwds or the column file's parameters (ImageD11/indexing.py:291-308)."""
    w = kwds.get("wavelength")
    if w is None:
        w = colfile.parameters.get("wavelength")
        if w is None:
            raise KeyError("Wavelength must be provided either in kwds or colfile.parameters")
    gv = np.array((colfile.gx, colfile.gy, colfile.gz), float)
    kwds.update({"unitcell": ucell, "wavelength": float(w), "gv": gv.T})
    ind = indexer(**kwds)
    if "omega" in colfile.titles:
        ind.omega_fullrange = find_omega_ranges(colfile.omega)
    ind.colfile = colfile
    return ind


def index(colf, npk_tol=((400, 0.01), (200, 0.02)), cosine_tol=math.cos(math.radians(90 - 0.1)),
          ds_tol=0.004, max_grains=1000, rmulmax=10, rings_to_use=None, maxpairs=None,
          log_level=3):
    """score_all_pairs for each (minpks, hkl_tol) of npk_tol on a column
    file's peaks (ImageD11/indexing.py:1267-1306).  Returns the indexer."""
    ind = indexer_from_colfile(colf, ds_tol=ds_tol, max_grains=max_grains,
                               cosine_tol=cosine_tol)
    ind.assigntorings()
    for minpks, tol in npk_tol:
        ind.minpks = minpks
        ind.hkl_tol = tol
        ind.score_all_pairs(n=maxpairs, rmulmax=rmulmax, rings_to_use=rings_to_use)
    return ind


def do_index(cf, dstol=0.005, hkl_tols=(0.01, 0.025, 0.05), fracs=(0.9, 0.7),
             cosine_tol=math.cos(math.radians(90 - 0.25)), max_grains=1000, forgen=(),
             foridx=(), unitcell=None, **kwds):
    """Index a column file (ImageD11/indexing.py:1309-1433): keep the peaks
    on the rings in foridx, then for each frac (minpks = frac x the peaks
    expected per grain) and each hkl_tol, search every pair (i <= j) of the
    rings in forgen, sparsest first.  Returns (grains, indexer)."""
    from .grain import grain as Grain
    for ringid in forgen:
        if ringid not in foridx:
            raise ValueError("All rings in forgen must be in foridx!")
    if cosine_tol < 0:
        import warnings
        warnings.warn("cosine_tol given as a negative number, are you sure about that?")
    logging.info("Indexing %d peaks", cf.nrows)

    def make(colfile):
        if unitcell is not None:
            ind = indexer_from_colfile_and_ucell(colfile, ucell=unitcell, **kwds)
        else:
            ind = indexer_from_colfile(colfile)
        ind.ds_tol = dstol
        ind.assigntorings()
        return ind
    ind = make(cf)
    pkmask = np.zeros(cf.nrows, bool)
    for i in foridx:
        pkmask |= ind.ra == i
    cf_for_indexing = cf.copyrows(pkmask)
    cf_for_indexing.parameters = cf.parameters
    ind = make(cf_for_indexing)
    omega_range = getattr(ind, "omega_fullrange", -1)
    if omega_range < 0:
        omega_range = 180
    n_peaks_expected = 0
    rings = []
    for i, dstar in enumerate(ind.unitcell.ringds):
        if ind.na[i] > 0 and i in foridx:
            multiplicity = len(ind.unitcell.ringhkls[ind.unitcell.ringds[i]])
            n_peaks_expected += int(multiplicity * omega_range / 180.0)
            if i in forgen:
                rings.append((ind.na[i], multiplicity, i))
    rings.sort()
    print("{} peaks expected".format(n_peaks_expected))
    print("Trying these rings (counts, multiplicity, ring number): {}".format(rings))
    ind.cosine_tol = cosine_tol
    ind.max_grains = max_grains
    threadb4 = cImageD11.cimaged11_omp_get_max_threads()
    try:
        cImageD11.cimaged11_omp_set_num_threads(1)
        for frac in fracs:
            ind.minpks = n_peaks_expected * frac
            for ind.hkl_tol in hkl_tols:
                for i in range(len(rings)):
                    for j in range(i, len(rings)):
                        ind.ring_1 = rings[i][2]
                        ind.ring_2 = rings[j][2]
                        ind.find()
                        if len(ind.hits) > 0:
                            ind.scorethem()
                print(frac, ind.hkl_tol, len(ind.ubis))
    finally:
        cImageD11.cimaged11_omp_set_num_threads(threadb4)
    grains = [Grain(ubi) for ubi in ind.ubis]
    if unitcell is not None:
        for g in grains:
            g.ref_unitcell = unitcell
    for ginc, g in enumerate(grains):
        try:
            g.name = g.ref_unitcell.name + ":" + str(ginc)
        except (NameError, KeyError, AttributeError):
            g.name = str(ginc)
    return grains, ind
