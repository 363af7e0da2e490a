"""ctypes binding of libgscore.so.

torch is imported before the library is opened: torch ships the HIP
runtime the process uses (its ``libamdhip64.so`` carries the SONAME
``libamdhip64.so.7``), and libgscore's own ``libamdhip64.so.7`` dependency
then resolves to that same, already-loaded runtime, so device pointers and
stream handles from torch are valid in the library.  Loading fails loudly
if the library was never built; there is no CPU fallback.
"""
import ctypes
import os
import re

import torch  # noqa: F401  (must precede the CDLL, see above)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgscore.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "gscore.h")


class GscoreError(RuntimeError):
    """A failure reported by libgscore (HIP error, bad argument, no GPU)."""


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "gs_last_error": (ctypes.c_char_p, []),
    "gs_clear_error": (_i, []),
    "gs_version": (_i, []),
    "gs_device_count": (_i, []),
    "gs_init": (_i, []),
    "gs_selftest_sincos_host": (None, [_vp, _vp, _vp, _i]),
    "gs_selftest_sincos_series_host": (None, [_vp, _vp, _vp, _i, _vp]),
    "gs_selftest_certified_host": (None, [_vp, _vp, _i, _d, _d, _vp]),
    "gs_selftest_certified_fract_host": (None, [_vp, _vp, _i, _d, _d, _vp]),
    "gs_selftest_scan_host": (None, [_vp, _vp, _vp, _vp, _vp, _i, _vp, _d, _d, _vp]),
    "verify_rounding": (_i, [_i]),
    "score": (_i, [_vp, _vp, _d, _i]),
    "score_and_refine": (None, [_vp, _vp, _d, _ip, _dp, _i]),
    "score_and_assign": (_i, [_vp, _vp, _d, _vp, _vp, _i, _i]),
    "refine_assigned": (None, [_vp, _vp, _vp, _i, _ip, _dp, _i]),
    "compute_gv": (None, [_vp, _vp, _d, _d, _d, _d, _vp, _vp, _i]),
    "compute_xlylzl": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "gs_score_many": (_i, [_vp, _i, _vp, _i, _d, _vp, _i, _i, _vp]),
    "gs_assign_many": (_i, [_vp, _vp, _i, _vp, _i, _d, _vp, _vp, _vp, _i, _vp]),
    "gs_score_and_refine_dev": (_i, [_vp, _vp, _i, _d, _vp, _vp, _vp]),
    "gs_refine_assigned_dev": (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "gs_score_and_refine_dev_noalloc": (_i, [_vp, _vp, _i, _d, _vp, _vp, _vp]),
    "gs_score_and_refine_host_ubi": (_i, [_vp, _vp, _i, _d, _vp, _vp, _vp]),
    "gs_refine_assigned_host_ubi": (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "gs_compute_xlylzl_dev": (_i, [_vp, _vp, _dp, _dp, _dp, _vp, _i, _vp]),
    "gs_compute_gv_dev": (_i, [_vp, _vp, _d, _d, _d, _d, _dp, _vp, _i, _vp]),
    "gs_compute_gv_trig_dev": (_i, [_vp, _vp, _vp, _dp, _d, _dp, _vp, _i, _vp]),
    "gs_sincos_omega": (_i, [_vp, _d, _vp, _vp, _i, _vp]),
    "gs_sincos_omega_host": (_i, [_vp, _d, _vp, _vp, _i]),
    "gs_wedge_chi_matrix_host": (_i, [_d, _d, _vp]),
    "gs_sincos_omega_parity": (_i, [_vp, _d, _vp, _vp, _i, _vp]),
    "gs_sf2gv_trig": (_i, [_vp, _vp, _vp, _vp, _dp, _dp, _dp, _dp, _d, _dp, _vp, _vp, _i, _vp]),
    "gs_sf2gv": (_i, [_vp, _vp, _vp, _dp, _dp, _dp, _d, _d, _d, _d, _dp, _vp, _vp, _vp, _vp, _i,
                      _vp]),
    "gs_refine_lists": (_i, [_vp, _i, _vp, _i, _vp, ctypes.c_longlong, _vp, _vp, _d, _i, _vp, _vp,
                             _i, _i, _vp, _vp]),
    "gs_refine_lists_long": (_i, [_vp, _i, _vp, _i, _vp, ctypes.c_longlong, _vp, _vp, _d, _i,
                                  _vp, _vp, _i, _i, ctypes.c_longlong, ctypes.c_longlong, _vp,
                                  _vp]),
    "gs_refine_graph_create_long": (_i, [ctypes.POINTER(_vp), _vp, _i, _vp, _i, _vp,
                                         ctypes.c_longlong, _vp, _vp, _d, _i, _vp, _vp, _i, _i,
                                         ctypes.c_longlong, ctypes.c_longlong, _vp, _vp]),
    "gs_gather_rows": (_i, [_vp, _i, _vp, ctypes.c_longlong, _vp, _vp]),
    "gs_index_decide": (_i, [_i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _i, _d, _i, _i, _vp, _i,
                             _vp, _vp, _vp, _vp]),
    "gs_refine_graph_create": (_i, [ctypes.POINTER(_vp), _vp, _i, _vp, _i, _vp, ctypes.c_longlong,
                                    _vp, _vp, _d, _i, _vp, _vp, _i, _i, _vp, _vp]),
    "gs_graph_launch": (_i, [_vp, _vp]),
    "gs_gv_assign_many": (_i, [_vp, _vp, _vp, _i, _d, _d, _d, _dp, _vp, _vp, _vp, _i, _d, _vp, _vp,
                               _vp, _vp]),
    "gs_gv_lists": (_i, [_vp, _vp, _vp, _i, _d, _d, _d, _dp, _vp, _i, _vp, _vp, ctypes.c_longlong,
                         _vp, _vp]),
    "gs_joint_hkl": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _i, _vp, _vp, ctypes.c_longlong,
                          _vp, _vp]),
    "gs_joint_schedule": (_i, [_vp, _i, ctypes.c_longlong, _i, _vp, _vp, _vp, ctypes.c_longlong,
                               _vp]),
    "gs_joint_workspace_doubles": (ctypes.c_longlong, [ctypes.c_longlong]),
    "gs_joint_adjoint": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _i, _vp, ctypes.c_longlong,
                              _vp, _vp, _vp, _vp, _vp, ctypes.c_longlong, _i, _i, _vp, _vp, _vp,
                              _vp, _vp, _vp, _i, _vp, _vp]),
    "gs_gauss_newton_ubis": (_i, [_vp, _vp, _i, _vp, _vp]),
    "gs_selftest_gn_ubis_host": (None, [_vp, _vp, _i, _vp]),
    "gs_simplex_order": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _vp,
                              _vp]),
    "gs_simplex_reflected": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp, _vp,
                                  _vp, _vp]),
    "gs_simplex_trial": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gs_graph_destroy": (_i, [_vp]),
    "closest": (None, [_vp, _vp, _ip, _dp, _i, _i]),
    "quickorient": (None, [_vp, _vp]),
    "gs_gv_ds": (_i, [_vp, _i, _vp, _vp]),
    "gs_ring_assign": (_i, [_vp, _i, _vp, _i, _d, _vp, _vp]),
    "gs_unit_rows": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "gs_find_hits": (_i, [_vp, _i, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "gs_find_all": (_i, [_vp, _i, _vp, _i, _vp, _i, _d, _vp, _vp, ctypes.c_longlong, _vp]),
    "gs_orient_hits": (_i, [_vp, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gs_select_many": (_i, [_vp, _i, _vp, _i, _d, _vp, _vp, _vp, _vp]),
    "compute_geometry": (None, [_vp, _vp, _d, _d, _d, _d, _vp, _vp, _i]),
    "compute_xlylzl_xpos_variable": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "gs_compute_geometry_dev": (_i, [_vp, _vp, _d, _d, _d, _d, _dp, _vp, _i, _vp]),
    "gs_compute_geometry_trig_dev": (_i, [_vp, _vp, _vp, _dp, _d, _dp, _vp, _i, _vp]),
    "gs_compute_xlylzl_xpos_dev": (_i, [_vp, _vp, _dp, _dp, _dp, _vp, _vp, _i, _vp]),
    "gs_sf2geometry": (_i, [_vp, _vp, _vp, _dp, _dp, _dp, _d, _d, _d, _d, _dp, _vp, _vp, _i,
                            _vp]),
    "gs_grain_gv": (_i, [_vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _vp, _vp, _i, _dp, _dp, _d, _vp,
                         _vp, ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp]),
    "gs_pbp_assign": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _i, _d, _d, _d, _vp, _vp,
                           _vp, _vp, _vp, _vp]),
    "gs_pbp_merge_sums": (_i, [_vp, _vp, ctypes.c_longlong, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _i, _vp, _vp]),
    "gs_pbp_ubi_fit": (_i, [_vp, _vp, _vp, _vp, _dp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gs_pbp_strain_fit": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "gs_pbp_rows_dot": (_i, [_vp, _vp, ctypes.c_longlong, _vp, _d, _vp, _vp, _vp, _vp]),
    "gs_segment_sum": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "gs_segment_max": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "gs_grain_gv_host": (_i, [_vp, _vp, _vp, _vp, _i, _d, _d, _d, _d, _vp, _vp, _i, _dp, _dp, _d,
                              _vp, _vp, ctypes.c_longlong, _vp, _vp, _vp, _vp]),
}

SCORE_CERTIFIED = 0
SCORE_EXACT = 1
SCORE_SIEVE = 2
SIEVE_TOL_MAX = 0.1  # the one-row sieve pays while ~2 tol of the pairs pass row 0


def score_mode(tol):
    """The faster of the certified score kernels for this tolerance: the
    sieve (row 0 of every pair first) for |tol| <= SIEVE_TOL_MAX, else the
    all-rows certified kernel.  Both give the reference's counts."""
    return SCORE_SIEVE if abs(float(tol)) <= SIEVE_TOL_MAX else SCORE_CERTIFIED

_lib = None


def header_symbols(path=HEADER):
    """Function names declared in include/gscore.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def hip_runtimes_loaded():
    """Distinct HIP runtime files mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return sorted(paths)


def load(path=LIB_PATH):
    """Open libgscore.so once and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            "libgscore.so is not built (%s); run `python -m imaged11_amd.build` "
            "or __graft_entry__.build()" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    runtimes = hip_runtimes_loaded()
    if len(runtimes) > 1:
        raise ImportError("two HIP runtimes are mapped (%s); import torch before "
                          "imaged11_amd" % ", ".join(runtimes))
    _lib = lib
    return lib


def last_error():
    msg = load().gs_last_error()
    return msg.decode() if msg else ""


def check(status, what=""):
    """Raise GscoreError for a non-zero gs_status."""
    if status != 0:
        raise GscoreError("%s%s (status %d)" % (what + ": " if what else "", last_error(), status))


def check_host_call(what):
    """After a reference-prototype call (no status): raise if it recorded an error."""
    msg = last_error()
    if msg:
        load().gs_clear_error()
        raise GscoreError("%s: %s" % (what, msg))


def require_gpu():
    """The product path runs on the GPU only."""
    if not torch.cuda.is_available():
        raise GscoreError("imaged11_amd needs a HIP GPU (torch.cuda.is_available() is False)")
    n = load().gs_device_count()
    if n <= 0:
        raise GscoreError("libgscore sees no HIP device")


def stream_ptr(stream=None):
    """hipStream_t of `stream` (default: torch's current stream) as an int."""
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def dptr(t):
    """Device address of a tensor's first element."""
    return ctypes.c_void_p(t.data_ptr())


def hptr(a):
    """Host address of a numpy array's first element."""
    return ctypes.c_void_p(a.ctypes.data)


def darr(values):
    """A small host double array (parameters read at launch time)."""
    vals = [float(v) for v in values]
    return (ctypes.c_double * len(vals))(*vals)
