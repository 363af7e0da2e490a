"""Batched device API: many grains per launch on HBM-resident torch tensors.

Where ImageD11 loops in Python over grains and calls one C routine per
grain (``indexing.py:818-820`` for score, ``refinegrains.py:718-734`` for
compute_gv + score_and_assign, ``refinegrains.py:459-498`` for
score_and_refine), these functions run the whole loop in one kernel launch
with the same per-call semantics.  Every function takes CUDA (HIP) tensors,
enqueues on torch's current stream and returns without synchronising,
except the parity mode's host trig (``parity=True``), which returns once
its values are on the device.
"""
import numpy as np
import torch

from . import _lib, geometry
from ._lib import dptr, darr, hptr, stream_ptr


def _need(t, dtype, name, shape=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a GPU tensor" % name)
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if shape is not None:
        if len(shape) != t.dim() or any(s is not None and s != d for s, d in zip(shape, t.shape)):
            raise ValueError("%s: expected shape %s, got %s" % (name, shape, tuple(t.shape)))
    return t


def _ready():
    lib = _lib.load()
    lib.gs_clear_error()
    return lib


from ._lib import SIEVE_TOL_MAX, score_mode  # noqa: E402  (the default score kernel)


def score_many(ubis, gv, tol, peaks_per_lane=0, out=None, exact=False, mode=None):
    """counts[g] = score(ubis[g], gv, tol) for every grain g.

    ubis (K,3,3) float64, gv (N,3) float64 -> int32 (K,).  exact=True uses
    the reference's arithmetic for every pair; by default the certified
    kernel score_mode(tol) picks runs; every mode returns the reference's
    counts."""
    lib = _ready()
    _need(ubis, torch.float64, "ubis", (None, 3, 3))
    _need(gv, torch.float64, "gv", (None, 3))
    K = ubis.shape[0]
    if out is None:
        out = torch.empty(K, dtype=torch.int32, device=gv.device)
    _need(out, torch.int32, "out", (K,))
    if mode is None:
        mode = _lib.SCORE_EXACT if exact else score_mode(tol)
    _lib.check(lib.gs_score_many(dptr(ubis), K, dptr(gv), gv.shape[0], float(tol), dptr(out),
                                 int(peaks_per_lane), mode, stream_ptr()), "score_many")
    return out


def boundary_pairs(ubis, gv, tol, rel=1e-12):
    """The boundary-peak audit of g-vectors computed with the device's
    sin/cos: per grain, the number of peaks with tol^2 (1 - rel) <= drlv2 <
    tol^2 (1 + rel), the band in which a last-bit difference between the
    device's and the host libm's sin/cos (hence in g) could move a peak
    across the tolerance sphere.  Zero everywhere means the counts (and
    labels) cannot depend on which sin/cos made the g-vectors.  Two
    score_many launches: count(drlv2 < tol_hi^2) - count(drlv2 < tol_lo^2)."""
    hi = score_many(ubis, gv, float(tol) * float(np.sqrt(1.0 + rel)))
    lo = score_many(ubis, gv, float(tol) * float(np.sqrt(1.0 - rel)))
    return hi - lo


def assign_many(ubis, gv, tol, drlv2, labels, grain_labels=None, out=None, exact=False):
    """score_and_assign(ubis[g], gv, tol, drlv2, labels, grain_labels[g]) for
    g = 0..K-1 in order; drlv2 (N,) float64 and labels (N,) int32 are
    updated in place.  grain_labels defaults to 0..K-1.  Returns the int32
    (K,) per-call counts.  exact=True evaluates every pair with the
    reference arithmetic instead of the certified fast path (same results)."""
    lib = _ready()
    _need(ubis, torch.float64, "ubis", (None, 3, 3))
    _need(gv, torch.float64, "gv", (None, 3))
    N = gv.shape[0]
    K = ubis.shape[0]
    _need(drlv2, torch.float64, "drlv2", (N,))
    _need(labels, torch.int32, "labels", (N,))
    if grain_labels is None:
        grain_labels = torch.arange(K, dtype=torch.int32, device=gv.device)
    _need(grain_labels, torch.int32, "grain_labels", (K,))
    if out is None:
        out = torch.empty(K, dtype=torch.int32, device=gv.device)
    mode = _lib.SCORE_EXACT if exact else _lib.SCORE_CERTIFIED
    _lib.check(lib.gs_assign_many(dptr(ubis), dptr(grain_labels), K, dptr(gv), N, float(tol),
                                  dptr(drlv2), dptr(labels), dptr(out), mode, stream_ptr()),
               "assign_many")
    return out


def xlylzl(sc, fc, p, r, dist, out=None):
    """compute_xlylzl on the device: (N,) sc, fc -> (N,3) lab positions."""
    lib = _ready()
    _need(sc, torch.float64, "sc", (None,))
    n = sc.shape[0]
    _need(fc, torch.float64, "fc", (n,))
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float64, device=sc.device)
    _lib.check(lib.gs_compute_xlylzl_dev(dptr(sc), dptr(fc), darr(p), darr(np.ravel(r)), darr(dist),
                                         dptr(out), n, stream_ptr()), "xlylzl")
    return out


def gvectors(xl, omega, omegasign, wvln, wedge, chi, t=(0.0, 0.0, 0.0), out=None,
             parity=False):
    """compute_gv on the device: (N,3) lab positions, (N,) omega -> (N,3).
    parity=True takes sin/cos from the host's glibc (see sincos_omega):
    the reference's g-vectors, bit for bit."""
    lib = _ready()
    _need(xl, torch.float64, "xlylzl", (None, 3))
    n = xl.shape[0]
    _need(omega, torch.float64, "omega", (n,))
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float64, device=xl.device)
    if parity:
        so, co = sincos_omega(omega, omegasign, parity=True)
        return gvectors_trig(xl, so, co, wedge_chi_matrix(wedge, chi), wvln, t=t, out=out)
    _lib.check(lib.gs_compute_gv_dev(dptr(xl), dptr(omega), float(omegasign), float(wvln),
                                     float(wedge), float(chi), darr(t), dptr(out), n, stream_ptr()),
               "gvectors")
    return out


def gvectors_trig(xl, so, co, mat, wvln, t=(0.0, 0.0, 0.0), out=None):
    """compute_gv arithmetic with the per-peak sin/cos and wedge/chi matrix
    supplied (the parity mode's, or any other)."""
    lib = _ready()
    _need(xl, torch.float64, "xlylzl", (None, 3))
    n = xl.shape[0]
    _need(so, torch.float64, "so", (n,))
    _need(co, torch.float64, "co", (n,))
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float64, device=xl.device)
    _lib.check(lib.gs_compute_gv_trig_dev(dptr(xl), dptr(so), dptr(co), darr(np.ravel(mat)),
                                          float(wvln), darr(t), dptr(out), n, stream_ptr()),
               "gvectors_trig")
    return out


def sincos_omega(omega, omegasign, parity=False):
    """Per-peak sin, cos of ((PI/180)*omega)*omegasign.

    parity=False (the fast mode): the device's correctly rounded sincos.
    parity=True: glibc's sincos on the host, the function the reference's
    compute_gv calls (``src/cdiffraction.c:146-147``), so the values are
    the reference's bit for bit; omega goes down to the host and (so, co)
    come back up (16 B/peak over PCIe, once per peak table).  The two
    differ in the last bit for a few tenths of a percent of angles, where
    glibc is not correctly rounded."""
    lib = _ready()
    _need(omega, torch.float64, "omega", (None,))
    so = torch.empty_like(omega)
    co = torch.empty_like(omega)
    fn = lib.gs_sincos_omega_parity if parity else lib.gs_sincos_omega
    _lib.check(fn(dptr(omega), float(omegasign), dptr(so), dptr(co), omega.shape[0],
                  stream_ptr()), "sincos_omega")
    return so, co


def wedge_chi_matrix(wedge, chi):
    """The (3,3) wedge/chi matrix of compute_gv (``src/cdiffraction.c:119-142``)
    from glibc's sincos on the host: the reference's values, bit for bit.
    Used by both modes (the per-peak omega trig is what the modes choose)."""
    m = np.zeros(9)
    _lib.check(_ready().gs_wedge_chi_matrix_host(float(wedge), float(chi), hptr(m)),
               "wedge_chi_matrix")
    return m.reshape(3, 3)


def sf2gv(sc, fc, omega, pars, t=(0.0, 0.0, 0.0), keep=False, parity=False):
    """Raw peaks (sc, fc, omega) to g-vectors in one fused pass, with the
    detector/goniometer of an ImageD11 parameter dict (``Ctransform.sf2gv``,
    ``ImageD11/transform.py``).  keep=True also returns the lab positions
    and per-peak sin/cos: (gv, xl, so, co).  parity=True: sin/cos of omega
    and the wedge/chi matrix from the host's glibc (see sincos_omega), so
    the g-vectors are the reference's compute_xlylzl + compute_gv, bit for
    bit; otherwise the device sincos in the same pass."""
    lib = _ready()
    _need(sc, torch.float64, "sc", (None,))
    n = sc.shape[0]
    _need(fc, torch.float64, "fc", (n,))
    _need(omega, torch.float64, "omega", (n,))
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    gv = torch.empty((n, 3), dtype=torch.float64, device=sc.device)
    if parity:
        so, co = sincos_omega(omega, sign, parity=True)
        xl = torch.empty((n, 3), dtype=torch.float64, device=sc.device) if keep else None
        _lib.check(lib.gs_sf2gv_trig(dptr(sc), dptr(fc), dptr(so), dptr(co), darr(p), darr(r),
                                     darr(dist), darr(np.ravel(wedge_chi_matrix(wedge, chi))),
                                     float(wvln), darr(t), dptr(xl) if keep else None, dptr(gv),
                                     n, stream_ptr()), "sf2gv")
        return (gv, xl, so, co) if keep else gv
    xl = so = co = None
    if keep:
        xl = torch.empty((n, 3), dtype=torch.float64, device=sc.device)
        so = torch.empty(n, dtype=torch.float64, device=sc.device)
        co = torch.empty(n, dtype=torch.float64, device=sc.device)
    _lib.check(lib.gs_sf2gv(dptr(sc), dptr(fc), dptr(omega), darr(p), darr(r), darr(dist),
                            float(sign), float(wvln), float(wedge), float(chi), darr(t),
                            dptr(xl) if keep else None, dptr(so) if keep else None,
                            dptr(co) if keep else None, dptr(gv), n, stream_ptr()), "sf2gv")
    if keep:
        return gv, xl, so, co
    return gv


def list_order(off):
    """Grain visiting order for the persistent refine: longest list first."""
    lengths = off[1:] - off[:-1]
    return torch.argsort(lengths, descending=True, stable=True).to(torch.int32)


def gather_rows(gv, idx):
    """rows[q] = gv[idx[q]] for every list entry, in list order, a NaN row
    where idx[q] is outside [0, len(gv)) (never kept by the refinement, as
    such an index is never kept)."""
    lib = _ready()
    _need(gv, torch.float64, "gv", (None, 3))
    _need(idx, torch.int32, "idx", (None,))
    rows = torch.empty((idx.numel(), 3), dtype=torch.float64, device=gv.device)
    _lib.check(lib.gs_gather_rows(dptr(gv), gv.shape[0], dptr(idx), idx.numel(), dptr(rows),
                                  stream_ptr()), "gather_rows")
    return rows


def refine_lists(ubis, gv, idx, off, tol, n_iter=2, order=None, lanes_per_grain=0,
                 waves_per_cu=0, inplace=False, gather=None, long_min=0, mid_min=0):
    """n_iter successive score_and_refine calls per grain on CSR peak lists.

    ubis (K,3,3) float64; gv (N,3); idx int32 peak indices, or None when
    row q of gv is list entry q; off int64 (K+1,) offsets.  Returns
    (ubis_out, n int32 (n_iter,K), mean float64 (n_iter,K)); n/mean of
    iteration i are those the i-th call returns (kept peaks and mean drlv2
    of the UBI entering that call).

    gather (default: when idx is given and n_iter >= 2) packs the list rows
    once with `gather_rows`, so every iteration streams contiguous rows
    instead of fetching a cache line per entry; the results are
    bit-identical either way.  long_min > 0: grains whose list has at
    least that many entries get a whole wave each, mid_min > 0: those with
    at least that many get 32 lanes each (gs_refine_lists_long), so the
    longest lists stop setting the time."""
    lib = _ready()
    _need(ubis, torch.float64, "ubis", (None, 3, 3))
    _need(gv, torch.float64, "gv", (None, 3))
    K = ubis.shape[0]
    _need(off, torch.int64, "off", (K + 1,))
    if idx is not None:
        _need(idx, torch.int32, "idx", (None,))
        if gather is None:
            gather = n_iter >= 2
        if gather:
            gv, idx = gather_rows(gv, idx), None
    n_idx = gv.shape[0] if idx is None else idx.numel()
    if order is None:
        order = list_order(off)
    _need(order, torch.int32, "order", (K,))
    u = ubis if inplace else ubis.clone()
    n = torch.empty((n_iter, K), dtype=torch.int32, device=gv.device)
    m = torch.empty((n_iter, K), dtype=torch.float64, device=gv.device)
    queue = torch.empty(3, dtype=torch.int32, device=gv.device)
    _lib.check(lib.gs_refine_lists_long(dptr(u), K, dptr(gv), gv.shape[0],
                                        None if idx is None else dptr(idx), n_idx,
                                        dptr(off), dptr(order), float(tol), int(n_iter), dptr(n),
                                        dptr(m), int(lanes_per_grain), int(waves_per_cu),
                                        int(long_min), int(mid_min), dptr(queue), stream_ptr()),
               "refine_lists")
    return u, n, m


def _mat_arg(mat):
    return None if mat is None else darr(np.ravel(mat))


def gv_assign_many(xl, so, co, wvln, ubis, trans, tol, drlv2, labels, grain_labels=None,
                   wedge=0.0, chi=0.0, mat=None, out=None):
    """For g = 0..K-1 in order: compute_gv with grain g's translation, then
    score_and_assign with label grain_labels[g] (refinegrains.assignlabels,
    ``ImageD11/refinegrains.py:718-734``), fused in one launch.  xl (N,3)
    lab positions, so/co (N,) sin/cos of omega*omegasign; drlv2, labels
    updated in place.  Returns the int32 (K,) per-call counts."""
    lib = _ready()
    _need(xl, torch.float64, "xlylzl", (None, 3))
    N = xl.shape[0]
    _need(so, torch.float64, "so", (N,))
    _need(co, torch.float64, "co", (N,))
    _need(ubis, torch.float64, "ubis", (None, 3, 3))
    K = ubis.shape[0]
    _need(trans, torch.float64, "trans", (K, 3))
    _need(drlv2, torch.float64, "drlv2", (N,))
    _need(labels, torch.int32, "labels", (N,))
    if grain_labels is None:
        grain_labels = torch.arange(K, dtype=torch.int32, device=xl.device)
    _need(grain_labels, torch.int32, "grain_labels", (K,))
    if out is None:
        out = torch.empty(K, dtype=torch.int32, device=xl.device)
    _lib.check(lib.gs_gv_assign_many(dptr(xl), dptr(so), dptr(co), N, float(wvln), float(wedge),
                                     float(chi), _mat_arg(mat), dptr(ubis), dptr(trans),
                                     dptr(grain_labels), K, float(tol), dptr(drlv2), dptr(labels),
                                     dptr(out), stream_ptr()), "gv_assign_many")
    return out


def lists_from_labels(labels, K):
    """CSR peak lists of grains 0..K-1 from per-peak labels (peaks in
    ascending order inside each list, label -1 and others dropped):
    returns (idx int32, off int64 (K+1,), owner int32)."""
    lab = labels.long()
    keep = (lab >= 0) & (lab < K)
    peaks = torch.nonzero(keep, as_tuple=False).flatten()
    order = torch.argsort(lab[peaks], stable=True)
    idx = peaks[order].to(torch.int32)
    counts = torch.bincount(lab[peaks], minlength=K)
    off = torch.zeros(K + 1, dtype=torch.int64, device=labels.device)
    off[1:] = torch.cumsum(counts, 0)
    owner = torch.repeat_interleave(torch.arange(K, dtype=torch.int32, device=labels.device),
                                    counts)
    return idx, off, owner


def gv_lists(xl, so, co, wvln, trans, idx, owner, wedge=0.0, chi=0.0, mat=None):
    """g-vector of every list entry with its grain's translation: (n_idx, 3)."""
    lib = _ready()
    _need(xl, torch.float64, "xlylzl", (None, 3))
    N = xl.shape[0]
    _need(so, torch.float64, "so", (N,))
    _need(co, torch.float64, "co", (N,))
    _need(trans, torch.float64, "trans", (None, 3))
    _need(idx, torch.int32, "idx", (None,))
    _need(owner, torch.int32, "owner", (idx.numel(),))
    out = torch.empty((idx.numel(), 3), dtype=torch.float64, device=xl.device)
    _lib.check(lib.gs_gv_lists(dptr(xl), dptr(so), dptr(co), N, float(wvln), float(wedge),
                               float(chi), _mat_arg(mat), dptr(trans), trans.shape[0], dptr(idx),
                               dptr(owner), idx.numel(), dptr(out), stream_ptr()), "gv_lists")
    return out


def grain_gv(xl, so, co, omega, pars, trans, idx, owner, mats=None, slop=0.25, angles=False):
    """refinegrains.compute_gv (``ImageD11/refinegrains.py:366-456``) for
    every list entry at once: g-vector of peak idx[q] seen from request
    owner[q] (translation trans[owner]).  With ``mats`` (R,3,3), the
    omega-float branch with those fitted UBIs and slop.  Returns (gv (M,3),
    tth, eta, omega_calc), the last three None unless ``angles``
    (omega_calc also None without mats)."""
    lib = _ready()
    _need(xl, torch.float64, "xlylzl", (None, 3))
    N = xl.shape[0]
    _need(so, torch.float64, "so", (N,))
    _need(co, torch.float64, "co", (N,))
    _need(trans, torch.float64, "trans", (None, 3))
    R = trans.shape[0]
    _need(idx, torch.int32, "idx", (None,))
    M = idx.numel()
    _need(owner, torch.int32, "owner", (M,))
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    rb = pre = None
    if mats is not None:
        _need(mats, torch.float64, "mats", (R, 3, 3))
        _need(omega, torch.float64, "omega", (N,))
        rbv, prev = geometry.uncompute_vectors(wvln, wedge, chi)
        rb = darr(rbv)
        pre = None if prev is None else darr(prev.ravel())
    dev = xl.device
    gv = torch.empty((M, 3), dtype=torch.float64, device=dev)
    tth = eta = omc = None
    if angles:
        tth = torch.empty(M, dtype=torch.float64, device=dev)
        eta = torch.empty(M, dtype=torch.float64, device=dev)
        if mats is not None:
            omc = torch.empty(M, dtype=torch.float64, device=dev)
    opt = lambda t: None if t is None else dptr(t)
    _lib.check(lib.gs_grain_gv(dptr(xl), dptr(so), dptr(co), opt(omega), N, float(sign),
                               float(wvln), float(wedge), float(chi), dptr(trans), opt(mats), R,
                               rb, pre, float(slop), dptr(idx), dptr(owner), M, opt(tth),
                               opt(eta), opt(omc), dptr(gv), stream_ptr()), "grain_gv")
    return gv, tth, eta, omc
