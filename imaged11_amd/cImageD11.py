"""Drop-in for ``ImageD11.cImageD11`` on the grain-scoring path.

Same names, arguments, return values and in-place behaviour as the f2py
wrappers of ``src/_cImageD11.pyf:106-170,551-583`` (re-exported by
``ImageD11/cImageD11.py:10``), so callers such as ``indexing.py`` and
``refinegrains.py`` work unchanged once these names are bound in place of
the extension's:

    n = score(ubi, gv, tol)
    n, mean = score_and_refine(ubi, gv, tol)          # ubi updated in place
    n = score_and_assign(ubi, gv, tol, drlv2, labels, label)
    npk, mean = refine_assigned(ubi, gv, labels, label)
    compute_gv(xlylzl, omega, omegasign, wvln, wedge, chi, t, gv)
    compute_xlylzl(s, f, p, r, dist, xlylzl)
    verify_rounding(n)
    compute_geometry(xlylzl, omega, omegasign, wvln, wedge, chi, t, out)
    compute_xlylzl_xpos_variable(s, f, p, r, dist, xpos, xlylzl)
    ibest, best = closest(x, v)                       # indexer.find
    quickorient(ubi, bt)                              # unitcell.orient

Argument handling follows f2py: ``intent(in)`` arrays of another dtype or
layout are converted (copied); ``intent(inout)`` arrays (``gv`` of
compute_gv, ``xlylzl``, ``drlv2``, ``labels``, and the ``ubi`` of
score_and_refine) must already be C-contiguous with the exact dtype, or
ValueError is raised before anything runs.  ``refine_assigned`` writes the
refined matrix into ``ubi`` only when ``ubi`` is a C-contiguous float64
(3, 3) array, as the reference's ``intent(c,in)`` does.

Every array argument may also be a torch tensor resident on the GPU: the
work then runs on the device arrays directly (no staging), on torch's
current stream, and returns once the results are available.  numpy inputs
go through the library's host-pointer entry points, which stage to the GPU.
All computation runs in the HIP kernels; there is no CPU path.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import GscoreError, dptr, hptr, darr, stream_ptr

__all__ = [
    "verify_rounding", "score", "score_and_refine", "score_and_assign", "refine_assigned",
    "compute_gv", "compute_xlylzl", "compute_geometry", "compute_xlylzl_xpos_variable",
    "closest", "quickorient", "cimaged11_omp_set_num_threads",
    "cimaged11_omp_get_max_threads", "GscoreError",
]

_omp_threads = 1


def cimaged11_omp_set_num_threads(n):
    """Kept for callers that pin host threads (``ImageD11/indexing.py:1402``);
    the kernels run on the GPU, so the value is only recorded."""
    global _omp_threads
    _omp_threads = max(1, int(n))


def cimaged11_omp_get_max_threads():
    """Returns the value recorded by cimaged11_omp_set_num_threads."""
    return _omp_threads


def _is_dev(x):
    return isinstance(x, torch.Tensor) and x.is_cuda


def _in_array(x, dtype, shape, name):
    """intent(in): convert to a C-contiguous array of dtype/shape (copy if needed)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=dtype))
    if shape is not None and not _shape_ok(a.shape, shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, a.shape))
    return a


def _shape_ok(got, want):
    if len(got) != len(want):
        return False
    return all(w is None or g == w for g, w in zip(got, want))


def _inout_array(x, dtype, shape, name):
    """intent(inout): must already be an ndarray of that dtype, C-contiguous."""
    if not isinstance(x, np.ndarray):
        raise ValueError("%s: intent(inout) argument must be a numpy array" % name)
    if x.dtype != np.dtype(dtype):
        raise ValueError("%s: intent(inout) array must have dtype %s, got %s"
                         % (name, np.dtype(dtype), x.dtype))
    if not x.flags.c_contiguous:
        raise ValueError("%s: intent(inout) array is not C-contiguous" % name)
    if shape is not None and not _shape_ok(x.shape, shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, x.shape))
    return x


def _dev_in(x, dtype, shape, name, device=None):
    """A device tensor input converted to a contiguous tensor of dtype."""
    if _is_dev(x):
        t = x.to(dtype=dtype).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(x, dtype=_NP[dtype])),
                            device=device)
    if shape is not None and not _shape_ok(tuple(t.shape), shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, tuple(t.shape)))
    return t


def _dev_inout(x, dtype, shape, name):
    if x.dtype != dtype:
        raise ValueError("%s: intent(inout) tensor must have dtype %s, got %s" % (name, dtype, x.dtype))
    if not x.is_contiguous():
        raise ValueError("%s: intent(inout) tensor is not contiguous" % name)
    if shape is not None and not _shape_ok(tuple(x.shape), shape):
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, tuple(x.shape)))
    return x


_NP = {torch.float64: np.float64, torch.int32: np.int32}


def _lib_ready():
    lib = _lib.load()
    _lib.require_gpu()
    lib.gs_clear_error()
    return lib



def verify_rounding(n):
    """Number of disagreements between the device's MAGIC rounding and
    floor(x + 0.5) on 600 integer-valued points (``src/closest.c:100-119``);
    0 when the rounding is right."""
    lib = _lib_ready()
    r = lib.verify_rounding(int(n))
    _lib.check_host_call("verify_rounding")
    return r


def score(ubi, gv, tol):
    """Count the g-vectors indexed by ubi within tol (``src/closest.c:227-248``)."""
    if _is_dev(gv):
        g = _dev_in(gv, torch.float64, (None, 3), "gv")
        u = _dev_in(ubi, torch.float64, (3, 3), "ubi", device=g.device)
        lib = _lib_ready()
        out = torch.empty(1, dtype=torch.int32, device=g.device)
        _lib.check(lib.gs_score_many(dptr(u), 1, dptr(g), g.shape[0], float(tol), dptr(out), 0,
                                     _lib.score_mode(tol), stream_ptr()), "score")
        return int(out.item())
    u = _in_array(ubi, np.float64, (3, 3), "ubi")
    g = _in_array(gv, np.float64, (None, 3), "gv")
    lib = _lib_ready()
    n = lib.score(hptr(u), hptr(g), float(tol), g.shape[0])
    _lib.check_host_call("score")
    return n


def _device_scalars(device):
    """(int32 view, float64 view) of one 16-byte device result buffer."""
    res = torch.empty(2, dtype=torch.float64, device=device)
    return res, res[0:1].view(torch.int32)[0:1], res[1:2]


def _read_scalars(res):
    r = res.cpu()
    return int(r[0:1].view(torch.int32)[0]), float(r[1])


def score_and_refine(ubi, gv, tol):
    """Score, then least-squares refine ubi in place on the indexed peaks
    (``src/closest.c:267-345``).  Returns (n, mean drlv2 before refinement)."""
    if _is_dev(ubi) or _is_dev(gv):
        g = _dev_in(gv, torch.float64, (None, 3), "gv")
        lib = _lib_ready()
        if not _is_dev(ubi):  # host UBI and results, device peaks: one launch, one wait
            host = _inout_array(ubi, np.float64, (3, 3), "ubi")
            n = ctypes.c_int(0)
            m = ctypes.c_double(0.0)
            _lib.check(lib.gs_score_and_refine_host_ubi(
                hptr(host), dptr(g), g.shape[0], float(tol), ctypes.byref(n), ctypes.byref(m),
                stream_ptr(torch.cuda.current_stream(g.device))), "score_and_refine")
            return n.value, m.value
        u = _dev_inout(ubi, torch.float64, (3, 3), "ubi")
        res, n, m = _device_scalars(g.device)
        _lib.check(lib.gs_score_and_refine_dev(dptr(u), dptr(g), g.shape[0], float(tol), dptr(n),
                                               dptr(m), stream_ptr()), "score_and_refine")
        return _read_scalars(res)
    u = _inout_array(ubi, np.float64, (3, 3), "ubi")
    g = _in_array(gv, np.float64, (None, 3), "gv")
    lib = _lib_ready()
    n = ctypes.c_int(0)
    m = ctypes.c_double(0.0)
    lib.score_and_refine(hptr(u), hptr(g), float(tol), ctypes.byref(n), ctypes.byref(m), g.shape[0])
    _lib.check_host_call("score_and_refine")
    return n.value, m.value


def score_and_assign(ubi, gv, tol, drlv2, labels, label):
    """Assign peaks to this ubi where they fit better than drlv2
    (``src/closest.c:366-397``); drlv2 and labels are updated in place.
    Returns the number of peaks assigned by this call."""
    if _is_dev(gv) or _is_dev(drlv2) or _is_dev(labels):
        g = _dev_in(gv, torch.float64, (None, 3), "gv")
        ng = g.shape[0]
        d = _dev_inout(drlv2, torch.float64, (ng,), "drlv2")
        lab = _dev_inout(labels, torch.int32, (ng,), "labels")
        u = _dev_in(ubi, torch.float64, (3, 3), "ubi", device=g.device)
        lib = _lib_ready()
        gl = torch.tensor([int(label)], dtype=torch.int32, device=g.device)
        out = torch.empty(1, dtype=torch.int32, device=g.device)
        _lib.check(lib.gs_assign_many(dptr(u), dptr(gl), 1, dptr(g), ng, float(tol), dptr(d),
                                      dptr(lab), dptr(out), _lib.SCORE_CERTIFIED, stream_ptr()),
                   "score_and_assign")
        return int(out.item())
    u = _in_array(ubi, np.float64, (3, 3), "ubi")
    g = _in_array(gv, np.float64, (None, 3), "gv")
    ng = g.shape[0]
    d = _inout_array(drlv2, np.float64, (ng,), "drlv2")
    lab = _inout_array(labels, np.int32, (ng,), "labels")
    lib = _lib_ready()
    n = lib.score_and_assign(hptr(u), hptr(g), float(tol), hptr(d), hptr(lab), int(label), ng)
    _lib.check_host_call("score_and_assign")
    return n


def refine_assigned(ubi, gv, labels, label):
    """Least-squares fit on the peaks with labels == label
    (``src/closest.c:416-474``).  Returns (npk, mean drlv2 before
    refinement); the fitted matrix is written into ubi only if ubi is a
    C-contiguous float64 (3, 3) array (or a float64 device tensor)."""
    writable = (isinstance(ubi, np.ndarray) and ubi.dtype == np.float64 and
                ubi.flags.c_contiguous and ubi.shape == (3, 3))
    if _is_dev(gv) or _is_dev(labels) or _is_dev(ubi):
        g = _dev_in(gv, torch.float64, (None, 3), "gv")
        ng = g.shape[0]
        lab = _dev_in(labels, torch.int32, (ng,), "labels", device=g.device)
        lib = _lib_ready()
        if not _is_dev(ubi):  # host UBI and results, device peaks: one launch, one wait
            u = ubi if writable else np.array(_in_array(ubi, np.float64, (3, 3), "ubi"))
            n = ctypes.c_int(0)
            m = ctypes.c_double(0.0)
            _lib.check(lib.gs_refine_assigned_host_ubi(
                hptr(u), dptr(g), ng, dptr(lab), int(label), ctypes.byref(n), ctypes.byref(m),
                stream_ptr(torch.cuda.current_stream(g.device))), "refine_assigned")
            return n.value, m.value
        if _is_dev(ubi) and ubi.dtype == torch.float64 and ubi.is_contiguous():
            u = _dev_inout(ubi, torch.float64, (3, 3), "ubi")
        else:
            u = _dev_in(ubi, torch.float64, (3, 3), "ubi", device=g.device).clone()
        lib = _lib_ready()
        res, n, m = _device_scalars(g.device)
        _lib.check(lib.gs_refine_assigned_dev(dptr(u), dptr(g), ng, dptr(lab), int(label), dptr(n),
                                              dptr(m), stream_ptr()), "refine_assigned")
        out = _read_scalars(res)
        if writable:
            ubi[...] = u.cpu().numpy()
        return out
    u = ubi if writable else _in_array(ubi, np.float64, (3, 3), "ubi")
    g = _in_array(gv, np.float64, (None, 3), "gv")
    ng = g.shape[0]
    lab = _in_array(labels, np.int32, (ng,), "labels")
    lib = _lib_ready()
    npk = ctypes.c_int(0)
    m = ctypes.c_double(0.0)
    lib.refine_assigned(hptr(u), hptr(g), hptr(lab), int(label), ctypes.byref(npk), ctypes.byref(m),
                        ng)
    _lib.check_host_call("refine_assigned")
    return npk.value, m.value


def _gv_call(what, width, xlylzl, omega, omegasign, wvln, wedge, chi, t, out, parity):
    """compute_gv (width 3) / compute_geometry (width 6) on host arrays or
    device tensors.  parity=True: sin/cos of omega, wedge and chi from the
    host's glibc, as the reference computes them, so the output is the
    reference's bit for bit (tth and eta of compute_geometry aside, which
    use the device atan2); parity=False: the device sincos."""
    tt = _in_array(np.asarray(t, dtype=np.float64).reshape(-1), np.float64, (3,), "t")
    on_dev = _is_dev(out) or _is_dev(xlylzl) or _is_dev(omega)
    if on_dev or not parity:
        host_out = None if on_dev else _inout_array(out, np.float64, (None, width), what)
        dev = next((a.device for a in (out, xlylzl, omega) if _is_dev(a)), torch.device("cuda"))
        x = _dev_in(xlylzl, torch.float64, (None, 3), "xlylzl", device=dev)
        n = x.shape[0]
        w = _dev_in(omega, torch.float64, (n,), "omega", device=dev)
        if on_dev:
            o = _dev_inout(out, torch.float64, (n, width), what)
        else:
            if host_out.shape[0] != n:
                raise ValueError("%s: expected shape %s, got %s" % (what, (n, width),
                                                                     host_out.shape))
            o = torch.empty((n, width), dtype=torch.float64, device=x.device)
        lib = _lib_ready()
        if parity:
            from . import batched
            so, co = batched.sincos_omega(w, omegasign, parity=True)
            mat = darr(batched.wedge_chi_matrix(wedge, chi).ravel())
            fn = lib.gs_compute_gv_trig_dev if width == 3 else lib.gs_compute_geometry_trig_dev
            _lib.check(fn(dptr(x), dptr(so), dptr(co), mat, float(wvln), darr(tt), dptr(o), n,
                          stream_ptr()), what)
        else:
            fn = lib.gs_compute_gv_dev if width == 3 else lib.gs_compute_geometry_dev
            _lib.check(fn(dptr(x), dptr(w), float(omegasign), float(wvln), float(wedge),
                          float(chi), darr(tt), dptr(o), n, stream_ptr()), what)
        torch.cuda.current_stream().synchronize()
        if host_out is not None:
            host_out[...] = o.cpu().numpy()
        return None
    x = _in_array(xlylzl, np.float64, (None, 3), "xlylzl")
    n = x.shape[0]
    w = _in_array(omega, np.float64, (n,), "omega")
    o = _inout_array(out, np.float64, (n, width), what)
    lib = _lib_ready()
    fn = lib.compute_gv if width == 3 else lib.compute_geometry
    fn(hptr(x), hptr(w), float(omegasign), float(wvln), float(wedge), float(chi), hptr(tt),
       hptr(o), n)
    _lib.check_host_call(what)
    return None


def compute_gv(xlylzl, omega, omegasign, wvln, wedge, chi, t, gv, parity=True):
    """g-vectors from lab-frame spot positions (``src/cdiffraction.c:112-168``),
    written into gv (n, 3).  parity=True (the default): the reference's
    values bit for bit (sin/cos from the host's glibc, as the reference
    computes them); parity=False: the device sincos, faster for device
    tensors, last-bit differences for a few tenths of a percent of peaks."""
    return _gv_call("compute_gv", 3, xlylzl, omega, omegasign, wvln, wedge, chi, t, gv, parity)


def compute_xlylzl(s, f, p, r, dist, xlylzl):
    """Lab-frame spot positions from detector pixel coordinates
    (``src/cdiffraction.c:190-224``), written into xlylzl (n, 3).
    p = [s_cen, f_cen, s_size, f_size], r = (tilt . flip).ravel(),
    dist = [dx, dy, dz]."""
    pp = _in_array(p, np.float64, (4,), "p")
    rr = _in_array(np.asarray(r, dtype=np.float64).reshape(-1), np.float64, (9,), "r")
    dd = _in_array(dist, np.float64, (3,), "dist")
    if _is_dev(xlylzl) or _is_dev(s) or _is_dev(f):
        ss = _dev_in(s, torch.float64, (None,), "s")
        n = ss.shape[0]
        ff = _dev_in(f, torch.float64, (n,), "f", device=ss.device)
        out = _dev_inout(xlylzl, torch.float64, (n, 3), "xlylzl")
        lib = _lib_ready()
        _lib.check(lib.gs_compute_xlylzl_dev(dptr(ss), dptr(ff), darr(pp), darr(rr), darr(dd),
                                             dptr(out), n, stream_ptr()), "compute_xlylzl")
        torch.cuda.current_stream().synchronize()
        return None
    ss = _in_array(s, np.float64, (None,), "s")
    n = ss.shape[0]
    ff = _in_array(f, np.float64, (n,), "f")
    out = _inout_array(xlylzl, np.float64, (n, 3), "xlylzl")
    lib = _lib_ready()
    lib.compute_xlylzl(hptr(ss), hptr(ff), hptr(pp), hptr(rr), hptr(dd), hptr(out), n)
    _lib.check_host_call("compute_xlylzl")
    return None


def closest(x, v):
    """(ibest, best): the index in x of the smallest |x[i] - v[j]| (first in
    x-major order) and that difference; (0, 99.0) when none is below 99
    (``src/closest.c:188-207``, f2py ``src/_cImageD11.pyf:88-103``).
    indexer.find calls it with x = cosines between peaks, v = the cosines
    the unit cell allows (``ImageD11/indexing.py:713``)."""
    if _is_dev(x) or _is_dev(v):
        raise ValueError("closest takes host arrays; the batched device search is "
                         "imaged11_amd.indexing.find_hits")
    xx = _in_array(x, np.float64, (None,), "x")
    vv = _in_array(v, np.float64, (None,), "v")
    lib = _lib_ready()
    ib = ctypes.c_int(0)
    best = ctypes.c_double(0.0)
    lib.closest(hptr(xx), hptr(vv), ctypes.byref(ib), ctypes.byref(best), xx.shape[0],
                vv.shape[0])
    _lib.check_host_call("closest")
    return ib.value, best.value


def quickorient(ubi, bt):
    """Orientation from two g-vectors held in rows 0 and 1 of ubi, written
    over ubi in place: ubi = bt . [t1; t2; t3] with t1 along g1, t3 along
    g1 x g2 (``src/cdiffraction.c:293-328``, f2py
    ``src/_cImageD11.pyf:603-613``; ubi intent(inout), bt intent(in))."""
    u = _inout_array(ubi, np.float64, (3, 3), "ubi")
    b = _in_array(bt, np.float64, (3, 3), "bt")
    lib = _lib_ready()
    lib.quickorient(hptr(u), hptr(b))
    _lib.check_host_call("quickorient")
    return None


def compute_geometry(xlylzl, omega, omegasign, wvln, wedge, chi, t, out, parity=True):
    """Per peak (tth, eta, ds, gx, gy, gz) from lab-frame positions and omega,
    written into out (n, 6) (``src/cdiffraction.c:31-94``; the
    ``updateGeometry`` pass of ``ImageD11/columnfile.py:469-504``).
    parity as for compute_gv (ds and g bit for bit in parity mode; tth and
    eta use the device's atan2, within a few ulps of glibc's)."""
    return _gv_call("compute_geometry", 6, xlylzl, omega, omegasign, wvln, wedge, chi, t, out,
                    parity)


def compute_xlylzl_xpos_variable(s, f, p, r, dist, xpos, xlylzl):
    """compute_xlylzl with a per-peak x offset xpos subtracted
    (``src/cdiffraction.c:245-277``; ``Ctransform.sf2gvlocal``)."""
    pp = _in_array(p, np.float64, (4,), "p")
    rr = _in_array(np.asarray(r, dtype=np.float64).reshape(-1), np.float64, (9,), "r")
    dd = _in_array(dist, np.float64, (3,), "dist")
    if _is_dev(xlylzl) or _is_dev(s) or _is_dev(f) or _is_dev(xpos):
        ss = _dev_in(s, torch.float64, (None,), "s")
        n = ss.shape[0]
        ff = _dev_in(f, torch.float64, (n,), "f", device=ss.device)
        xp = _dev_in(xpos, torch.float64, (n,), "xpos", device=ss.device)
        out = _dev_inout(xlylzl, torch.float64, (n, 3), "xlylzl")
        lib = _lib_ready()
        _lib.check(lib.gs_compute_xlylzl_xpos_dev(dptr(ss), dptr(ff), darr(pp), darr(rr),
                                                  darr(dd), dptr(xp), dptr(out), n,
                                                  stream_ptr()), "compute_xlylzl_xpos_variable")
        torch.cuda.current_stream().synchronize()
        return None
    ss = _in_array(s, np.float64, (None,), "s")
    n = ss.shape[0]
    ff = _in_array(f, np.float64, (n,), "f")
    xp = _in_array(xpos, np.float64, (n,), "xpos")
    out = _inout_array(xlylzl, np.float64, (n, 3), "xlylzl")
    lib = _lib_ready()
    lib.compute_xlylzl_xpos_variable(hptr(ss), hptr(ff), hptr(pp), hptr(rr), hptr(dd), hptr(xp),
                                     hptr(out), n)
    _lib.check_host_call("compute_xlylzl_xpos_variable")
    return None
