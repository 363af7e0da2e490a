"""Benchmark: grain x peak pairs scored per second on one MI355X.

Workload (one "step"): raw peaks (sc, fc, omega) -> compute_xlylzl ->
compute_gv -> score for K = 1000 grains x N = 1e6 peaks in float64, tol
0.05; peaks forward-simulated from 1000 seeded grains (half cubic
a~U[3,6] A, half orthorhombic a,b,c~U[3,9] A) on a tilted 2048^2 detector
at 50 um / 150 mm, wavelength 0.2 A, omega 0-180 deg in 0.25 deg frames,
0.5 px jitter, 10% spurious peaks.  Inputs are resident in HBM before the
timed region.  value = K*N*steps / elapsed.

Printed: ONE JSON line, with a roofline object for the dominant kernel
(k_score_many, FP64 vector ALU bound) timed live with HIP events on its
stream, and a cpu_baseline object: the reference C (oracle/_ref, compiled
from its sources) running the same pipeline on the host cores, the grain
loop spread over threads (ctypes releases the GIL), with its per-grain
counts compared to the GPU's.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 23      # h = UBI.g: 9 mul + 6 add; 3 sub; sumsq: 3 mul + 2 add
VALU_OPS_PER_PAIR = 30   # as issued: + 6 rounding adds + 1 compare (no FMA)
PEAK_FP64_VALU = 78.6e12  # MI355X FP64 vector, FMA counted as 2
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grains", type=int, default=1000)
    ap.add_argument("--peaks", type=int, default=1_000_000)
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(wl, tol, threads, gpu_counts):
    """The same pipeline with the reference C on the host (or the oracle
    port where the reference was not built)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from imaged11_amd import geometry
    from oracle import oracle

    ref = oracle.load_reference()
    kind = "reference"
    if ref is None:
        ref = oracle.load_oracle()
        kind = "port"
    else:
        ref.set_threads(threads)
    pk, pars = wl["peaks"], wl["pars"]
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    n = len(pk["sc"])
    ubis = wl["ubi"]
    times = []
    counts = None
    with ThreadPoolExecutor(max_workers=threads) as pool:
        for rep in range(3):
            t0 = time.perf_counter()
            xl = np.zeros((n, 3))
            ref.compute_xlylzl(pk["sc"], pk["fc"], p, r, dist, xl)
            gv = np.zeros((n, 3))
            ref.compute_gv(xl, pk["omega"], sign, wvln, wedge, chi, np.zeros(3), gv)
            counts = list(pool.map(lambda u: ref.score(u, gv, tol), ubis))
            times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    pairs = len(ubis) * n
    return {
        "value": pairs / t, "unit": "pairs/s", "cores": threads, "kind": kind,
        "sample": "full workload (%d grains x %d peaks), median of 3; compute_xlylzl serial, "
                  "compute_gv OpenMP, score() per grain over %d threads" % (len(ubis), n, threads),
        "seconds": t,
        "counts_bit_exact_vs_gpu": bool(np.array_equal(np.array(counts), gpu_counts)),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every GPU box here has one, and the "
              "multi-GPU path is design-only (peak sharding + an int32 all-reduce "
              "of the K counts). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from imaged11_amd import _lib, batched, simulate

    _lib.load()
    _lib.require_gpu()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    t0 = time.perf_counter()
    wl = simulate.scoring_workload(K=args.grains, N=args.peaks, seed=args.seed)
    gen_s = time.perf_counter() - t0
    pk, pars = wl["peaks"], wl["pars"]

    def T(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)

    sc, fc, om, ubis = T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), T(wl["ubi"])
    counts = torch.empty(args.grains, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        g = batched.sf2gv(sc, fc, om, pars)
        if ev is not None:
            ev[0].record(stream)
        batched.score_many(ubis, g, args.tol, out=counts)
        if ev is not None:
            ev[1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for i in range(args.steps):
        step(evs[i])
    end.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(end)
    ms_per_step = total_ms / args.steps
    score_ms = statistics.mean(a.elapsed_time(b) for a, b in evs)
    pairs = args.grains * args.peaks
    value = pairs * args.steps / (total_ms * 1e-3)
    achieved = FLOPS_PER_PAIR * pairs / (score_ms * 1e-3)
    gpu_counts = counts.cpu().numpy()

    out = {
        "metric": "grain x peak pairs/s scored (compute_xlylzl -> compute_gv -> score)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded forward simulation, generated in %.1f s)" % gen_s,
        "config": {"workload": "raw peaks -> compute_xlylzl -> compute_gv -> score, "
                               "%d grains x %d peaks, tol %g" % (args.grains, args.peaks, args.tol),
                   "grains": args.grains, "peaks": args.peaks, "tol": args.tol},
        "roofline": {
            "kernel": "k_score_many",
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64_VALU / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_VALU,
            "traffic": None,
            "kernel_ms": score_ms,
            "flops_per_pair": FLOPS_PER_PAIR,
            "valu_ops_per_pair": VALU_OPS_PER_PAIR,
            "issued_fp64_ops_frac_of_nonfma_peak": VALU_OPS_PER_PAIR * pairs / (score_ms * 1e-3)
            / (PEAK_FP64_VALU / 2),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(wl, args.tol, threads, gpu_counts)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
