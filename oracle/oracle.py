"""ORACLE — test infrastructure only.

ctypes bindings for the CPU checker libraries built by ``oracle/Makefile``:

* ``liboracle.so`` — the clean-room C restatement (``gs_oracle.c``);
* ``_ref/libref.so`` — the reference C (``src/closest.c``,
  ``src/cdiffraction.c``) compiled from its own sources, present only where
  the reference tree was available when the checker was built.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  The wrappers take numpy arrays and follow
the f2py calling conventions of ``src/_cImageD11.pyf:106-170,551-583``
(in-place ``ubi``/``drlv2``/``labels``, tuples for ``intent(out)``).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def build(quiet=True):
    """Compile liboracle.so (and _ref/libref.so when the reference exists)."""
    out = subprocess.run(["make", "-B", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


class _Base:
    """Entry points shared by the restatement and the compiled reference."""

    prefix = ""

    def __init__(self, path):
        self.path = path
        self.lib = ctypes.CDLL(path)
        p = self.prefix
        L = self.lib
        self._verify = getattr(L, p + "verify_rounding")
        self._verify.argtypes = [ctypes.c_int]
        self._verify.restype = ctypes.c_int
        self._score = getattr(L, p + "score")
        self._score.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int]
        self._score.restype = ctypes.c_int
        self._sar = getattr(L, p + "score_and_refine")
        self._sar.argtypes = [_dp, _dp, ctypes.c_double, _ip, _dp, ctypes.c_int]
        self._sar.restype = None
        self._saa = getattr(L, p + "score_and_assign")
        self._saa.argtypes = [_dp, _dp, ctypes.c_double, _dp, _ip, ctypes.c_int, ctypes.c_int]
        self._saa.restype = ctypes.c_int
        self._ra = getattr(L, p + "refine_assigned")
        self._ra.argtypes = [_dp, _dp, _ip, ctypes.c_int, _ip, _dp, ctypes.c_int]
        self._ra.restype = None
        self._gv = getattr(L, p + "compute_gv")
        self._gv.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double, _dp, _dp, ctypes.c_int]
        self._gv.restype = None
        self._xl = getattr(L, p + "compute_xlylzl")
        self._xl.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        self._xl.restype = None
        self._geom = getattr(L, p + "compute_geometry")
        self._geom.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, _dp, _dp, ctypes.c_int]
        self._geom.restype = None
        self._xpos = getattr(L, p + "compute_xlylzl_xpos_variable")
        self._xpos.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        self._xpos.restype = None
        self._closest = getattr(L, p + "closest")
        self._closest.argtypes = [_dp, _dp, _ip, _dp, ctypes.c_int, ctypes.c_int]
        self._closest.restype = None
        self._qo = getattr(L, p + "quickorient")
        self._qo.argtypes = [_dp, _dp]
        self._qo.restype = None

    # f2py-style wrappers -------------------------------------------------
    def verify_rounding(self, n):
        return self._verify(int(n))

    def score(self, ubi, gv, tol):
        ubi = _f64(ubi, (3, 3))
        gv = _f64(gv)
        return self._score(_d(ubi), _d(gv), float(tol), gv.shape[0])

    def score_and_refine(self, ubi, gv, tol):
        """ubi is updated in place (must be C-contiguous float64 (3,3))."""
        assert ubi.dtype == np.float64 and ubi.flags.c_contiguous and ubi.shape == (3, 3)
        gv = _f64(gv)
        n = ctypes.c_int(0)
        m = ctypes.c_double(0)
        self._sar(_d(ubi), _d(gv), float(tol), ctypes.byref(n), ctypes.byref(m), gv.shape[0])
        return n.value, m.value

    def score_and_assign(self, ubi, gv, tol, drlv2, labels, label):
        ubi = _f64(ubi, (3, 3))
        gv = _f64(gv)
        assert drlv2.dtype == np.float64 and drlv2.flags.c_contiguous
        assert labels.dtype == np.int32 and labels.flags.c_contiguous
        return self._saa(_d(ubi), _d(gv), float(tol), _d(drlv2), _i(labels), int(label), gv.shape[0])

    def refine_assigned(self, ubi, gv, labels, label):
        """ubi is overwritten when it is C-contiguous float64 (f2py quirk)."""
        gv = _f64(gv)
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        npk = ctypes.c_int(0)
        m = ctypes.c_double(0)
        self._ra(_d(ubi), _d(gv), _i(labels), int(label), ctypes.byref(npk), ctypes.byref(m),
                 gv.shape[0])
        return npk.value, m.value

    def compute_gv(self, xlylzl, omega, omegasign, wvln, wedge, chi, t, gv):
        xlylzl = _f64(xlylzl)
        omega = _f64(omega)
        t = _f64(t, (3,))
        assert gv.dtype == np.float64 and gv.flags.c_contiguous
        self._gv(_d(xlylzl), _d(omega), float(omegasign), float(wvln), float(wedge), float(chi),
                 _d(t), _d(gv), xlylzl.shape[0])

    def compute_xlylzl(self, s, f, p, r, dist, xlylzl):
        s = _f64(s)
        f = _f64(f)
        p = _f64(p, (4,))
        r = _f64(r, (9,))
        dist = _f64(dist, (3,))
        assert xlylzl.dtype == np.float64 and xlylzl.flags.c_contiguous
        self._xl(_d(s), _d(f), _d(p), _d(r), _d(dist), _d(xlylzl), s.shape[0])

    def compute_geometry(self, xlylzl, omega, omegasign, wvln, wedge, chi, t, out):
        xlylzl = _f64(xlylzl)
        omega = _f64(omega)
        t = _f64(t, (3,))
        assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape == (len(omega), 6)
        self._geom(_d(xlylzl), _d(omega), float(omegasign), float(wvln), float(wedge),
                   float(chi), _d(t), _d(out), omega.shape[0])

    def compute_xlylzl_xpos_variable(self, s, f, p, r, dist, xpos, xlylzl):
        s, f, xpos = _f64(s), _f64(f), _f64(xpos)
        p, r, dist = _f64(p, (4,)), _f64(r, (9,)), _f64(dist, (3,))
        assert xlylzl.dtype == np.float64 and xlylzl.flags.c_contiguous
        self._xpos(_d(s), _d(f), _d(p), _d(r), _d(dist), _d(xpos), _d(xlylzl), s.shape[0])

    def closest(self, x, v):
        x = _f64(x)
        v = _f64(v)
        ib = ctypes.c_int(0)
        b = ctypes.c_double(0)
        self._closest(_d(x), _d(v), ctypes.byref(ib), ctypes.byref(b), x.shape[0], v.shape[0])
        return ib.value, b.value

    def quickorient(self, ubi, bt):
        """ubi (3,3) C-contiguous float64, overwritten."""
        assert ubi.dtype == np.float64 and ubi.flags.c_contiguous and ubi.shape == (3, 3)
        bt = _f64(bt, (3, 3))
        self._qo(_d(ubi), _d(bt))


class Reference(_Base):
    """The reference C built from its own sources into ``_ref/libref.so``."""

    prefix = ""

    def __init__(self, path=REF_SO):
        super().__init__(path)
        self._setthreads = self.lib.cimaged11_omp_set_num_threads
        self._setthreads.argtypes = [ctypes.c_int]

    def set_threads(self, n):
        self._setthreads(int(n))


class Oracle(_Base):
    """The clean-room restatement ``liboracle.so`` plus its batched drivers."""

    prefix = "orc_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.orc_inverse3x3.argtypes = [_dp]
        L.orc_inverse3x3.restype = ctypes.c_int
        L.orc_wedge_chi_matrix.argtypes = [ctypes.c_double, ctypes.c_double, _dp]
        L.orc_sincos_omega.argtypes = [_dp, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_compute_gv_trig.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp,
                                          ctypes.c_int]
        L.orc_score_many.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _ip,
                                     ctypes.c_int]
        L.orc_assign_many.argtypes = [_dp, _ip, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                      _dp, _ip, _ip, ctypes.c_int]
        L.orc_gv_assign_many.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _ip,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip,
                                         _ip, ctypes.c_int]
        L.orc_refine_lists.argtypes = [_dp, ctypes.c_int, _dp, _ip, _llp, ctypes.c_double,
                                       ctypes.c_int, _ip, _dp, ctypes.c_int]
        L.orc_joint_loss.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp,
                                     _dp, ctypes.c_int, _ip, _llp, _ip]
        L.orc_joint_loss.restype = ctypes.c_double
        L.orc_openmp_threads.restype = ctypes.c_int
        L.orc_unit_rows.argtypes = [_dp, _ip, ctypes.c_int, _dp]
        L.orc_dot3_rows.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_compute_geometry_trig.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp,
                                                ctypes.c_int]
        L.orc_find_hits.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                    _dp]
        L.orc_find_all.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                   ctypes.c_double, _llp, ctypes.c_longlong]
        L.orc_find_all.restype = ctypes.c_longlong
        L.orc_orient_hits.argtypes = [_dp, _ip, _ip, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                      _ip]

    def compute_geometry_trig(self, xlylzl, so, co, mat, wvln, t):
        xlylzl = _f64(xlylzl)
        so, co = _f64(so), _f64(co)
        mat, t = _f64(mat, (9,)), _f64(t, (3,))
        out = np.empty((len(so), 6))
        self.lib.orc_compute_geometry_trig(_d(xlylzl), _d(so), _d(co), _d(mat), float(wvln),
                                           _d(t), _d(out), len(so))
        return out

    def dot3_rows(self, x, y, mode):
        x = _f64(x, (-1, 3))
        y = _f64(y, (-1, 3))
        out = np.empty(len(x))
        self.lib.orc_dot3_rows(_d(x), _d(y), len(x), int(mode), _d(out))
        return out

    def unit_rows(self, gv, idx):
        gv = _f64(gv, (-1, 3))
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.empty((len(idx), 3))
        self.lib.orc_unit_rows(_d(gv), _i(idx), len(idx), _d(out))
        return out

    def find_hits(self, n1, n2, cs):
        n1 = _f64(n1, (-1, 3))
        n2 = _f64(n2, (-1, 3))
        cs = _f64(cs)
        ib = np.zeros(len(n1), np.int32)
        best = np.zeros(len(n1))
        self.lib.orc_find_hits(_d(n1), len(n1), _d(n2), len(n2), _d(cs), len(cs), _i(ib),
                               _d(best))
        return ib, best

    def find_all(self, n1, n2, cs, mtol):
        n1 = _f64(n1, (-1, 3))
        n2 = _f64(n2, (-1, 3))
        cs = _f64(cs)
        n = self.lib.orc_find_all(_d(n1), len(n1), _d(n2), len(n2), _d(cs), len(cs),
                                  float(mtol), None, 0)
        keys = np.zeros(max(n, 1), np.int64)
        self.lib.orc_find_all(_d(n1), len(n1), _d(n2), len(n2), _d(cs), len(cs), float(mtol),
                              keys.ctypes.data_as(_llp), n)
        return keys[:n]

    def orient_hits(self, gv, hi, hj, c2ab, bt):
        gv = _f64(gv, (-1, 3))
        hi = np.ascontiguousarray(hi, dtype=np.int32)
        hj = np.ascontiguousarray(hj, dtype=np.int32)
        c2ab = _f64(c2ab)
        bt = _f64(bt, (-1, 3, 3))
        ubis = np.zeros((len(hi), 3, 3))
        choice = np.zeros(len(hi), np.int32)
        self.lib.orc_orient_hits(_d(gv), _i(hi), _i(hj), len(hi), _d(c2ab), len(c2ab), _d(bt),
                                 _d(ubis), _i(choice))
        return ubis, choice

    def inverse3x3(self, a):
        a = _f64(a, (3, 3)).copy()
        ok = self.lib.orc_inverse3x3(_d(a))
        return ok, a

    def wedge_chi_matrix(self, wedge, chi):
        m = np.zeros(9)
        self.lib.orc_wedge_chi_matrix(float(wedge), float(chi), _d(m))
        return m

    def sincos_omega(self, omega, omegasign):
        omega = _f64(omega)
        so = np.empty_like(omega)
        co = np.empty_like(omega)
        self.lib.orc_sincos_omega(_d(omega), float(omegasign), _d(so), _d(co), omega.shape[0])
        return so, co

    def compute_gv_trig(self, xlylzl, so, co, mat, wvln, t):
        xlylzl = _f64(xlylzl)
        so = _f64(so)
        co = _f64(co)
        mat = _f64(mat, (9,))
        t = _f64(t, (3,))
        out = np.empty_like(xlylzl)
        self.lib.orc_compute_gv_trig(_d(xlylzl), _d(so), _d(co), _d(mat), float(wvln), _d(t),
                                     _d(out), xlylzl.shape[0])
        return out

    def score_many(self, ubis, gv, tol, nthreads=0):
        ubis = _f64(ubis, (-1, 3, 3))
        gv = _f64(gv)
        counts = np.zeros(ubis.shape[0], np.int32)
        self.lib.orc_score_many(_d(ubis), ubis.shape[0], _d(gv), gv.shape[0], float(tol),
                                _i(counts), int(nthreads))
        return counts

    def assign_many(self, ubis, grain_labels, gv, tol, drlv2, labels, nthreads=0):
        ubis = _f64(ubis, (-1, 3, 3))
        gl = np.ascontiguousarray(grain_labels, dtype=np.int32)
        gv = _f64(gv)
        assert drlv2.dtype == np.float64 and labels.dtype == np.int32
        counts = np.zeros(ubis.shape[0], np.int32)
        self.lib.orc_assign_many(_d(ubis), _i(gl), ubis.shape[0], _d(gv), gv.shape[0], float(tol),
                                 _d(drlv2), _i(labels), _i(counts), int(nthreads))
        return counts

    def gv_assign_many(self, xlylzl, so, co, mat, wvln, ubis, trans, grain_labels, tol, drlv2,
                       labels, nthreads=0):
        xlylzl = _f64(xlylzl)
        so = _f64(so)
        co = _f64(co)
        mat = _f64(mat, (9,))
        ubis = _f64(ubis, (-1, 3, 3))
        trans = _f64(trans, (-1, 3))
        gl = np.ascontiguousarray(grain_labels, dtype=np.int32)
        counts = np.zeros(ubis.shape[0], np.int32)
        self.lib.orc_gv_assign_many(_d(xlylzl), _d(so), _d(co), _d(mat), float(wvln), _d(ubis),
                                    _d(trans), _i(gl), ubis.shape[0], xlylzl.shape[0], float(tol),
                                    _d(drlv2), _i(labels), _i(counts), int(nthreads))
        return counts

    def refine_lists(self, ubis, gv, idx, off, tol, n_iter, nthreads=0):
        """Returns (ubis_out, n[n_iter,K], mean[n_iter,K]); input ubis untouched."""
        u = _f64(ubis, (-1, 3, 3)).copy()
        gv = _f64(gv)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.int64)
        K = u.shape[0]
        n = np.zeros((n_iter, K), np.int32)
        m = np.zeros((n_iter, K), np.float64)
        self.lib.orc_refine_lists(_d(u), K, _d(gv), _i(idx), off.ctypes.data_as(_llp), float(tol),
                                  int(n_iter), _i(n), _d(m), int(nthreads))
        return u, n, m

    def joint_loss(self, sc, fc, so, co, p, r, dist, wvln, mat, ubis, trans, idx, off, hkl):
        sc, fc, so, co = (_f64(a) for a in (sc, fc, so, co))
        p = _f64(p, (4,))
        r = _f64(r, (9,))
        dist = _f64(dist, (3,))
        mat = _f64(mat, (9,))
        ubis = _f64(ubis, (-1, 3, 3))
        trans = _f64(trans, (-1, 3))
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        off = np.ascontiguousarray(off, dtype=np.int64)
        hkl = np.ascontiguousarray(hkl, dtype=np.int32)
        return self.lib.orc_joint_loss(_d(sc), _d(fc), _d(so), _d(co), _d(p), _d(r), _d(dist),
                                       float(wvln), _d(mat), _d(ubis), _d(trans), ubis.shape[0],
                                       _i(idx), off.ctypes.data_as(_llp), _i(hkl))

    def openmp_threads(self):
        return self.lib.orc_openmp_threads()


def load_oracle():
    if not os.path.exists(ORACLE_SO):
        build()
    return Oracle()


def load_reference():
    """The compiled reference, or None where it was never built."""
    if os.path.exists(REF_SO):
        return Reference()
    return None
