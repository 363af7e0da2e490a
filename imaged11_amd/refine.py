"""Batched least-squares refinement of many grains on the GPU.

``refinegrains.refine`` (``ImageD11/refinegrains.py:459-498``) calls
``score_and_refine`` a few times per grain from a Python loop, and
``refinegrains.gof`` repeats that for every grain at every simplex step.
Here all grains are refined together: each grain owns a CSR list of peak
indices into one HBM-resident g-vector table, and the n_iter successive
``score_and_refine`` calls per grain run either

* ``"fused"``: inside one persistent launch (each grain keeps its UBI in
  registers between iterations), or
* ``"graph"``: as n_iter single-iteration persistent launches captured
  once into a hipGraph and replayed (`RefineGraph`),

with launch parameters (lanes per grain, resident waves per CU, the list
lengths above which a grain gets 32 or 64 lanes) picked by `autotune` and
cached per problem shape.
"""
import ctypes
import json
import os
import time

import numpy as np
import torch

from . import _lib, batched
from ._lib import dptr, stream_ptr

LANES = (8, 16, 32, 64)
WAVES_PER_CU = (4, 8, 16)
MODES = ("fused", "graph")
CACHE_ENV = "IMAGED11_AMD_TUNE_CACHE"


class RefineGraph:
    """n_iter score_and_refine passes over all grains, captured as a
    hipGraph on a private stream.  ``ubis`` is refined in place by each
    replay; ``n`` and ``mean`` (n_iter, K) hold the per-call results of the
    last replay.  The tensors passed in must stay alive with the graph.

    With an index array (and n_iter >= 2) the graph runs on the list rows
    packed once by `batched.gather_rows`; each replay packs them again
    first, so a replay sees the current contents of gv.  idx None: row q
    of gv is list entry q."""

    def __init__(self, ubis, gv, idx, off, tol, n_iter, order=None, lanes_per_grain=0,
                 waves_per_cu=0, long_min=0, mid_min=0):
        batched._need(ubis, torch.float64, "ubis", (None, 3, 3))
        batched._need(gv, torch.float64, "gv", (None, 3))
        K = ubis.shape[0]
        if idx is not None:
            batched._need(idx, torch.int32, "idx", (None,))
        batched._need(off, torch.int64, "off", (K + 1,))
        if order is None:
            order = batched.list_order(off)
        batched._need(order, torch.int32, "order", (K,))
        self.ubis, self.gv, self.idx, self.off, self.order = ubis, gv, idx, off, order
        self.rows = None
        rows, ridx = gv, idx
        if idx is not None and n_iter >= 2:
            self.rows = rows = batched.gather_rows(gv, idx)
            ridx = None
        n_idx = rows.shape[0] if ridx is None else ridx.numel()
        dev = gv.device
        self.n = torch.zeros((n_iter, K), dtype=torch.int32, device=dev)
        self.mean = torch.zeros((n_iter, K), dtype=torch.float64, device=dev)
        self.queue = torch.zeros(n_iter + 2, dtype=torch.int32, device=dev)
        self.stream = torch.cuda.Stream(device=dev)
        lib = _lib.load()
        lib.gs_clear_error()
        handle = ctypes.c_void_p()
        torch.cuda.current_stream(dev).synchronize()
        _lib.check(lib.gs_refine_graph_create_long(
            ctypes.byref(handle), dptr(ubis), K, dptr(rows), rows.shape[0],
            None if ridx is None else dptr(ridx), n_idx,
            dptr(off), dptr(order), float(tol), int(n_iter), dptr(self.n), dptr(self.mean),
            int(lanes_per_grain), int(waves_per_cu), int(long_min), int(mid_min),
            dptr(self.queue), stream_ptr(self.stream)), "refine graph capture")
        self.handle = handle

    def replay(self, stream=None):
        """Enqueue one replay on `stream` (default: torch's current stream)."""
        lib = _lib.load()
        if self.rows is not None:
            _lib.check(lib.gs_gather_rows(dptr(self.gv), self.gv.shape[0], dptr(self.idx),
                                          self.idx.numel(), dptr(self.rows),
                                          stream_ptr(stream)), "gather_rows")
        _lib.check(lib.gs_graph_launch(self.handle, stream_ptr(stream)), "graph launch")

    def close(self):
        if self.handle:
            _lib.load().gs_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _cache_path():
    return os.environ.get(CACHE_ENV) or os.path.join(
        os.path.expanduser("~"), ".cache", "imaged11_amd", "refine_tune.json")


def problem_key(off, n_iter):
    """Shape of a refinement problem: arch, grain count and list-length
    quantiles, each rounded to a power-of-two bucket."""
    lengths = (off[1:] - off[:-1]).double()
    q = torch.quantile(lengths, torch.tensor([0.5, 0.9, 1.0], dtype=torch.float64,
                                              device=lengths.device)).cpu().numpy()
    b = lambda v: int(2 ** round(np.log2(max(float(v), 1.0))))
    props = torch.cuda.get_device_properties(off.device)
    arch = getattr(props, "gcnArchName", props.name)
    return "%s|K%d|L%d-%d-%d|it%d|t3" % (arch, b(off.numel() - 1), b(q[0]), b(q[1]), b(q[2]),
                                         n_iter)


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune(ubis, gv, idx, off, tol, n_iter, reps=5, use_cache=True, verbose=False):
    """Pick (mode, lanes_per_grain, waves_per_cu, long_min, mid_min) for
    this problem by timing every candidate with HIP events (long_min: 0, or
    the 90th / 99th percentile of the list lengths: the grains with lists
    at least that long get a whole wave; mid_min: 0, or the 50th / 75th
    percentile: those get 32 lanes); results are cached on disk by
    `problem_key`.  Returns the best config dict (with its time in ms)."""
    key = problem_key(off, n_iter)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            cache = json.load(open(path))
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]
    order = batched.list_order(off)
    work = ubis.clone()
    lengths = (off[1:] - off[:-1]).double()
    q = torch.quantile(lengths, torch.tensor([0.5, 0.75, 0.9, 0.99], dtype=torch.float64,
                                             device=lengths.device)).cpu().numpy()
    longs = sorted({0} | {int(round(v)) for v in q[2:] if v >= 1})
    mids = sorted({0} | {int(round(v)) for v in q[:2] if v >= 1})
    results = []
    for lanes in LANES:
        tiers = [(lm, mm) for lm in (longs if lanes < 64 else [0])
                 for mm in (mids if lanes < 32 else [0]) if not (lm and mm >= lm)]
        for wpc in WAVES_PER_CU:
            for lm, mm in tiers:
                cfg = {"lanes_per_grain": lanes, "waves_per_cu": wpc, "long_min": lm,
                       "mid_min": mm}

                def fused():
                    work.copy_(ubis)
                    batched.refine_lists(work, gv, idx, off, tol, n_iter, order=order,
                                         lanes_per_grain=lanes, waves_per_cu=wpc,
                                         inplace=True, long_min=lm, mid_min=mm)
                results.append((dict(cfg, mode="fused"), _time(fused, reps)))
                g = RefineGraph(work, gv, idx, off, tol, n_iter, order=order,
                                lanes_per_grain=lanes, waves_per_cu=wpc, long_min=lm,
                                mid_min=mm)

                def graph():
                    work.copy_(ubis)
                    g.replay()
                results.append((dict(cfg, mode="graph"), _time(graph, reps)))
                g.close()
    results.sort(key=lambda r: r[1])
    if verbose:
        for cfg, ms in results:
            print("%-6s lanes %2d waves/CU %2d long_min %5d mid_min %5d : %.3f ms" % (
                cfg["mode"], cfg["lanes_per_grain"], cfg["waves_per_cu"], cfg["long_min"],
                cfg["mid_min"], ms))
    best = dict(results[0][0], ms=results[0][1], key=key, tuned_at=time.time())
    cache[key] = best
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp%d" % os.getpid()
            json.dump(cache, open(tmp, "w"), indent=1)
            os.replace(tmp, path)
        except OSError:
            pass
    return best


def refine(ubis, gv, idx, off, tol, n_iter=2, config=None):
    """Refine every grain with the given (or autotuned) configuration.
    Returns (ubis_out, n (n_iter,K), mean (n_iter,K))."""
    if config is None:
        config = autotune(ubis, gv, idx, off, tol, n_iter)
    if config.get("mode", "fused") == "graph":
        u = ubis.clone()
        g = RefineGraph(u, gv, idx, off, tol, n_iter, lanes_per_grain=config["lanes_per_grain"],
                        waves_per_cu=config["waves_per_cu"],
                        long_min=config.get("long_min", 0), mid_min=config.get("mid_min", 0))
        g.replay()
        n, m = g.n.clone(), g.mean.clone()
        torch.cuda.current_stream().synchronize()
        g.close()
        return u, n, m
    return batched.refine_lists(ubis, gv, idx, off, tol, n_iter,
                                lanes_per_grain=config["lanes_per_grain"],
                                waves_per_cu=config["waves_per_cu"],
                                long_min=config.get("long_min", 0),
                                mid_min=config.get("mid_min", 0))
