"""Benchmark: grain x peak pairs scored per second on one MI355X.

Workload (one "step"): raw peaks (sc, fc, omega) -> compute_xlylzl ->
compute_gv -> score for K = 1000 grains x N = 1e6 peaks in float64, tol
0.05; peaks forward-simulated from 1000 seeded grains (half cubic
a~U[3,6] A, half orthorhombic a,b,c~U[3,9] A) on a tilted 2048^2 detector
at 50 um / 150 mm, wavelength 0.2 A, omega 0-180 deg in 0.25 deg frames,
0.5 px jitter, 10% spurious peaks.  Inputs are resident in HBM before the
timed region.  value = K*N*steps / elapsed.

Printed: ONE JSON line, with a roofline object for the dominant kernel
(k_score_many_sieve at tol <= 0.1, else k_score_many_fast; FP64 vector ALU
bound) timed live with HIP events on its
stream, and a cpu_baseline object: the reference C (oracle/_ref, compiled
from its sources) running the same pipeline on the host cores, the grain
loop spread over threads (ctypes releases the GIL), with its per-grain
counts compared to the GPU's.

    python bench.py [--gpus 1] [--steps 100] [--warmup 50] [--warmup-s 1.5]
                    [--no-cpu] [--no-aux]

Warmup: the --warmup steps, then more until --warmup-s seconds of steps
have run, whatever --warmup says: the GPU clock ramps up over the first
launches after idle (they run 15-25% slower), and the timed region should
see the steady clock.  The JSON line states both (warmup, warmup_s,
warmup_steps_run).
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
# FP64 VALU instructions issued per pair by the certified kernel: rows
# h + 1/2 with 3 FMA each (9), fract per row (3), sum f (f - 1) in 5, two
# compares (2)
VALU_OPS_PER_PAIR = 19
# the sieved score (k_score_many_sieve<4>, the default for tol <= 0.1),
# counted in its gfx950 assembly: every pair 8.5 VALU (row 0 + r by three
# FMAs, v_fract_f64, one compare; two mbcnt and the queue address; one
# entry per grain step for all four slots), every pair passing row 0
# another 26.5 (17 FP64 of the certified test, the entry decode, addresses
# and compares).  The SQ counters count 12.3 per pair in all, the tile
# setup and the queue drains included.
SIEVE_STAGE1_VALU = 8.5
SIEVE_SETTLE_VALU = 26.5
# FP64 VALU instructions per pair of the fused compute_gv + score_and_assign
# kernel's certified loop, counted in its gfx950 assembly: 70 per grain step
# of two peaks per lane (22 v_fmac, 16 v_fma, 8 v_mul, 12 v_add, 6
# v_fract, 2 v_rsq, 4 v_cmp; one Newton step after each v_rsq_f64), plus
# 2.5 integer VALU per step (the label test of each peak, one LDS address)
SCAN_VALU_OPS_PER_PAIR = 35.0
# FP64 flops per list entry of the adjoint pass (hkl-space loss with the
# Gauss-Newton sums), FMA = 2, counted from its source: forward
# compute_xlylzl + compute_gv 75, residual and loss 24, dL/dG 18, the
# Gauss-Newton sums 30, reverse sweep 128
ADJ_FLOPS_PER_ENTRY = 275
PEAK_FP64_VALU = 78.6e12  # MI355X FP64 vector, FMA counted as 2
PEAK_HBM = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--warmup-s", type=float, default=1.5,
                    help="keep warming up until this many seconds of steps have run")
    ap.add_argument("--grains", type=int, default=1000)
    ap.add_argument("--peaks", type=int, default=1_000_000)
    ap.add_argument("--tol", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the batched-refinement and adjoint measurements")
    ap.add_argument("--profile", choices=("refine", "adjoint", "scan"), default=None,
                    help="only run one workload a few times for a profiler (no timing "
                         "line): the 5000-grain batched refinement or its adjoint, or "
                         "the 20000-grain x 8e6-peak scan")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic that calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def host_cpus(requested=0):
    """(threads, facts) for the host baselines: the CPUs this process may
    run on (its affinity mask), capped by OMP_NUM_THREADS when that is set
    (the CPU share given to one GPU), unless --cpu-threads says
    otherwise.  facts records nproc, the affinity count and the cap."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = nproc
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else None
    threads = requested or (min(affinity, cap) if cap else affinity)
    return threads, {"nproc": nproc, "affinity_cpus": affinity, "omp_num_threads": cap,
                     "threads_used": threads,
                     "rule": "--cpu-threads if given, else the affinity CPUs capped by "
                             "OMP_NUM_THREADS"}


def measured_traffic(kernel):
    """HBM bytes per launch of the score kernel on the bench workload, from
    the committed rocprofv3 PMC summary (null when absent)."""
    path = os.path.join(ROOT, "profiles", kernel + "__traffic_1000x1e6.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


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
        for rep in range(6):  # one warm-up, then the median of five
            t0 = time.perf_counter()
            xl = np.zeros((n, 3))
            ref.compute_xlylzl(pk["sc"], pk["fc"], p, r, dist, xl)
            gv = np.zeros((n, 3))
            ref.compute_gv(xl, pk["omega"], sign, wvln, wedge, chi, np.zeros(3), gv)
            counts = list(pool.map(lambda u: ref.score(u, gv, tol), ubis))
            times.append(time.perf_counter() - t0)
    t = statistics.median(times[1:])
    pairs = len(ubis) * n
    # one host thread on a 100-grain sample (score is serial per grain as
    # shipped; the OpenMP parts on one thread too)
    if kind == "reference":
        ref.set_threads(1)
    t0 = time.perf_counter()
    xl1 = np.zeros((n, 3))
    ref.compute_xlylzl(pk["sc"], pk["fc"], p, r, dist, xl1)
    gv1 = np.zeros((n, 3))
    ref.compute_gv(xl1, pk["omega"], sign, wvln, wedge, chi, np.zeros(3), gv1)
    for u in ubis[:100]:
        ref.score(u, gv1, tol)
    t1 = time.perf_counter() - t0
    if kind == "reference":
        ref.set_threads(threads)
    return {
        "value": pairs / t, "unit": "pairs/s", "cores": threads, "kind": kind,
        "sample": "full workload (%d grains x %d peaks), median of 5 after one warm-up; "
                  "compute_xlylzl serial, compute_gv OpenMP, score() per grain over %d threads"
                  % (len(ubis), n, threads),
        "seconds": t,
        "single_thread": {"value": min(100, len(ubis)) * n / t1, "unit": "pairs/s", "cores": 1,
                          "sample": "the same pipeline for the first 100 grains on one thread"},
        "counts_bit_exact_vs_gpu": bool(np.array_equal(np.array(counts), gpu_counts)),
    }


def _events_ms(fn, reps=3, warm_s=0.2):
    """Mean ms per call of fn over `reps` back-to-back calls, bracketed by
    HIP events on the current stream, after calls for at least warm_s
    seconds (a leg that follows CPU-only work starts on a lowered clock)."""
    import torch
    t0 = time.perf_counter()
    while True:
        fn()
        torch.cuda.synchronize()
        if time.perf_counter() - t0 >= warm_s:
            break
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def aux_refine_adjoint(threads, with_cpu):
    """Batched refinement of 5000 grains with lognormal lists (median 300,
    clipped [20, 2000]) from a 2e6-peak table, 10 score_and_refine
    iterations: autotuned batched launch vs the per-grain drop-in loop on
    the same GPU (device-resident lists) vs the reference C per grain on the
    host; then one adjoint evaluation of the joint objective on the same
    lists and 50 joint refinement steps."""
    import numpy as np
    import torch
    from imaged11_amd import adjoint, batched, geometry, refine, simulate
    from imaged11_amd import cImageD11 as gpu_c

    dev = torch.device("cuda:0")
    K, N, n_iter, tol = 5000, 2_000_000, 10, 0.05
    t0 = time.perf_counter()
    wl = simulate.scoring_workload(K=K, N=N, seed=3)
    idx, off = simulate.refine_lists(wl["peaks"]["grain"], K, seed=3)
    gen_s = time.perf_counter() - t0
    pk, pars = wl["peaks"], wl["pars"]
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    gv = batched.sf2gv(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), pars)
    rng = np.random.default_rng(4)
    start = wl["ubi"] @ (np.eye(3) + rng.normal(scale=1e-3, size=(K, 3, 3)))
    ubis0 = T(start)
    idx_d, off_d = T(idx, torch.int32), T(off, torch.int64)
    M = int(off[-1])
    best = refine.autotune(ubis0, gv, idx_d, off_d, tol, n_iter, reps=3)
    order = batched.list_order(off_d)  # a property of the lists: computed once
    work = ubis0.clone()
    if best["mode"] == "graph":
        g = refine.RefineGraph(work, gv, idx_d, off_d, tol, n_iter, order=order,
                               lanes_per_grain=best["lanes_per_grain"],
                               waves_per_cu=best["waves_per_cu"],
                               long_min=best.get("long_min", 0),
                               mid_min=best.get("mid_min", 0))
        run = lambda: (work.copy_(ubis0), g.replay())
    else:
        run = lambda: batched.refine_lists(ubis0, gv, idx_d, off_d, tol, n_iter, order=order,
                                           lanes_per_grain=best["lanes_per_grain"],
                                           waves_per_cu=best["waves_per_cu"],
                                           long_min=best.get("long_min", 0),
                                           mid_min=best.get("mid_min", 0))
    batched_ms = _events_ms(run, reps=20)
    # the same iterations captured as one hipGraph of per-iteration
    # persistent launches (the graph mode), in the tuned shape
    gwork = ubis0.clone()
    graph = refine.RefineGraph(gwork, gv, idx_d, off_d, tol, n_iter, order=order,
                               lanes_per_grain=best["lanes_per_grain"],
                               waves_per_cu=best["waves_per_cu"],
                               long_min=best.get("long_min", 0),
                               mid_min=best.get("mid_min", 0))
    graph_ms = _events_ms(lambda: (gwork.copy_(ubis0), graph.replay()), reps=20)
    graph.close()
    # per-grain drop-in loop on the same GPU: lists gathered once, then the
    # reference's call pattern (n_iter score_and_refine calls per grain)
    lists = [gv[idx_d[off[k]:off[k + 1]].long()].contiguous() for k in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u_loop = start.copy()
    for k in range(K):
        u = u_loop[k]
        for _ in range(n_iter):
            gpu_c.score_and_refine(u, lists[k], tol)
    loop_ms = (time.perf_counter() - t0) * 1e3
    out = {"grains": K, "peaks": N, "list_entries": M, "n_iter": n_iter, "gen_s": gen_s,
           "config": best, "batched_ms": batched_ms, "graph_mode_ms": graph_ms,
           "per_grain_gpu_loop_ms": loop_ms,
           "speedup_vs_gpu_loop": loop_ms / batched_ms,
           "refine_iterations_per_s": n_iter / (batched_ms * 1e-3),
           "pairs_per_s": M * n_iter / (batched_ms * 1e-3),
           # gathered refine: 28 B per list entry (int32 index + 24 B g) and
           # 156 B per grain (UBI in and out, n, mean) per iteration
           "algorithmic_hbm_gb_s": (28 * M + 156 * K) * n_iter / (batched_ms * 1e-3) / 1e9}
    # roofline of the call (gather and tier count included): per list entry
    # and iteration the reference's drlv2 (23 flops), per kept entry the
    # n and sum-drlv2 adds and the 15 R and H terms by FMA (32 flops); the
    # kept counts from one more call outside the timed region
    n_kept = batched.refine_lists(ubis0, gv, idx_d, off_d, tol, n_iter, order=order,
                                  lanes_per_grain=best["lanes_per_grain"],
                                  waves_per_cu=best["waves_per_cu"],
                                  long_min=best.get("long_min", 0),
                                  mid_min=best.get("mid_min", 0))[1]
    flops = 23.0 * M * n_iter + 32.0 * float(n_kept.sum())
    out["roofline"] = {"bound": "valu", "achieved": flops / (batched_ms * 1e-3) / 1e12,
                       "peak": PEAK_FP64_VALU / 1e12, "unit": "TFLOP/s",
                       "frac": flops / (batched_ms * 1e-3) / PEAK_FP64_VALU,
                       "kept_fraction": float(n_kept.sum()) / (M * n_iter),
                       "frac_is": "the whole call (gather and tier count included) over the FP64 "
                                  "peak; the kernel runs one resident wave per SIMD"}
    if with_cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle
        ref = oracle.load_reference() or oracle.load_oracle()
        gvh = gv.cpu().numpy()
        sub = [np.ascontiguousarray(gvh[idx[off[k]:off[k + 1]]]) for k in range(K)]

        def one(k):
            u = start[k].copy()
            nm = [ref.score_and_refine(u, sub[k], tol) for _ in range(n_iter)]
            return u, nm
        with ThreadPoolExecutor(max_workers=threads) as pool:
            t0 = time.perf_counter()
            res = list(pool.map(one, range(K)))
            cpu_ms = (time.perf_counter() - t0) * 1e3
        out["cpu_reference_ms"] = cpu_ms
        out["cpu_threads"] = threads
        out["speedup_vs_cpu"] = cpu_ms / batched_ms
        # parity against the reference's per-grain loop (outside the timed
        # region): every iteration's n bit-exact, UBIs within 1e-10 of the
        # largest entry, the first call's mean within 1e-12 relative (later
        # calls start from UBIs that already differ by rounding)
        u_g, n_g, m_g = batched.refine_lists(ubis0, gv, idx_d, off_d, tol, n_iter, order=order,
                                             lanes_per_grain=best["lanes_per_grain"],
                                             waves_per_cu=best["waves_per_cu"],
                                             long_min=best.get("long_min", 0),
                                             mid_min=best.get("mid_min", 0))
        u_g, n_g, m_g = u_g.cpu().numpy(), n_g.cpu().numpy(), m_g.cpu().numpy()
        u_c = np.array([r[0] for r in res])
        n_c = np.array([[nm[0] for nm in r[1]] for r in res]).T
        m_c = np.array([[nm[1] for nm in r[1]] for r in res]).T
        rel = np.abs(m_g - m_c) / np.maximum(np.abs(m_c), 1e-300)
        ubi_err = float(np.max(np.abs(u_g - u_c).max(axis=(1, 2)) /
                               np.abs(u_c).max(axis=(1, 2))))
        out["parity_vs_reference"] = {
            "n_bit_exact_all_iterations": bool(np.array_equal(n_g, n_c)),
            "max_ubi_rel_err": ubi_err, "ubi_within_1e-10": ubi_err <= 1e-10,
            "max_mean_rel_err_first_call": float(rel[0].max()),
            "max_mean_rel_err_later_calls": float(rel[1:].max()) if n_iter > 1 else None,
            "sample": "all %d grains x %d iterations against the reference C's "
                      "score_and_refine loop" % (K, n_iter)}
    # adjoint on the same lists (translations zero), then 50 joint steps
    prob = adjoint.JointProblem(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), pars["omegasign"],
                                idx_d, off_d)
    params = {k: torch.tensor(float(pars[k]), dtype=torch.float64, device=dev)
              for k in adjoint.FREE[:8]}
    trans = torch.zeros((K, 3), dtype=torch.float64, device=dev)
    geo = adjoint.pack_geometry(params, pars)
    prob.assign_hkl(geo, ubis0, trans)
    adj_ms = _events_ms(lambda: prob.evaluate(geo, ubis0, trans, want_sums=True), reps=50)
    p2 = dict(params, distance=params["distance"] + 100.0)
    scales = {k: 1.0 for k in p2} | {"distance": 100.0, "tilt_x": 1e-3, "tilt_y": 1e-3,
                                     "tilt_z": 1e-3, "wavelength": 1e-5, "wedge": 0.01}
    # warm the code paths (the optimiser and the host autograd start cold)
    adjoint.refine_joint(prob, ubis0, trans, p2, pars, steps=3, lr=0.05, scales=scales)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, p_fin, hist = adjoint.refine_joint(prob, ubis0, trans, p2, pars, steps=50, lr=0.05,
                                          scales=scales)
    torch.cuda.synchronize()
    joint_ms = (time.perf_counter() - t0) * 1e3
    # what the joint steps should reach (outside the timed region): the
    # loss at the simulation's true geometry after the same Gauss-Newton
    # UBI updates from the same start, i.e. the floor the peak noise sets
    u_true = ubis0
    for _ in range(50):
        loss_true, _, _, _, sums = prob.evaluate(geo, u_true, trans, want_sums=True)
        u_true = adjoint.gauss_newton_ubis(sums, u_true)
    loss_true = float(loss_true)
    # roofline of the evaluation (the pass over pieces and the per-grain
    # finish, bracketed together by the events): per list entry the pass
    # streams 44 algorithmic bytes (sc, fc, sin, cos: 32 B, hkl: 12 B) and
    # does ADJ_FLOPS_PER_ENTRY FP64 flops (FMA = 2)
    adj_s = adj_ms * 1e-3
    out["adjoint"] = {"eval_ms": adj_ms, "pairs": M,
                      "pairs_per_s": M / adj_s,
                      "pieces": prob.n_pieces, "piece_len": prob.piece_len,
                      "roofline": {"bound": "hbm", "achieved": 44 * M / adj_s / 1e9,
                                   "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                   "frac": 44 * M / adj_s / PEAK_HBM,
                                   "fp64_tflops": ADJ_FLOPS_PER_ENTRY * M / adj_s / 1e12,
                                   "fp64_frac": ADJ_FLOPS_PER_ENTRY * M / adj_s / PEAK_FP64_VALU},
                      "joint_steps": 50, "joint_refine_ms": joint_ms,
                      "joint_start": "UBIs perturbed by 1e-3, distance 100 um off",
                      "loss_first": hist[0], "loss_last": hist[-1],
                      "loss_at_true_geometry": loss_true,
                      "loss_last_over_true": hist[-1] / loss_true,
                      "distance_error_um": {
                          "start": 100.0,
                          "final": float(p_fin["distance"]) - float(pars["distance"])}}
    out["batched_refine_plus_adjoint_ms"] = batched_ms + adj_ms
    out["speedup_refine_plus_adjoint_vs_gpu_loop"] = loop_ms / (batched_ms + adj_ms)
    out["speedup_graph_mode_plus_adjoint_vs_gpu_loop"] = loop_ms / (graph_ms + adj_ms)
    if "cpu_reference_ms" in out:
        out["speedup_refine_plus_adjoint_vs_cpu"] = out["cpu_reference_ms"] / (batched_ms + adj_ms)
        out["speedup_graph_mode_plus_adjoint_vs_cpu"] = out["cpu_reference_ms"] / (graph_ms +
                                                                                  adj_ms)
    return out


def aux_index(with_cpu, K=3000, seed=11):
    """Ring-pair indexing (indexer.assigntorings + find + scorethem) of K
    fcc grains' g-vectors on the GPU, against the CPU checker's sequential
    restatement of the same search (one thread) on the same input."""
    import numpy as np
    import torch
    from imaged11_amd import indexing, simulate, unitcell
    t0 = time.perf_counter()
    w = simulate.indexing_workload(K=K, seed=seed)
    gen_s = time.perf_counter() - t0
    pars = dict(ring_1=0, ring_2=1, minpks=20, hkl_tol=0.02, max_grains=2 * K)
    uc = unitcell.unitcell(list(w["cell"]), w["symmetry"])

    def run():
        ind = indexing.indexer(uc, w["gv"], **pars)
        marks = [time.perf_counter()]
        nh = 0
        for step in (ind.assigntorings, ind.find, ind.scorethem):
            step()
            torch.cuda.synchronize()
            marks.append(time.perf_counter())
            nh = max(nh, len(ind.hits))
        return ind, nh, np.diff(marks) * 1e3

    run()  # the ring tables and cosine caches are built once per unit cell
    ind, nh, ms = run()
    out = {"grains": K, "peaks": int(len(w["gv"])), "gen_s": gen_s, "hits": nh,
           "grains_found": len(ind.ubis), "assigntorings_ms": ms[0], "find_ms": ms[1],
           "scorethem_ms": ms[2], "total_ms": float(ms.sum()),
           "trial_pairs_scored": nh * int(len(w["gv"]))}
    if with_cpu:
        from oracle import oracle as O, indexer_oracle as IO
        orc = O.load_oracle()
        ucc = unitcell.unitcell(list(w["cell"]), w["symmetry"], ext=orc)
        t0 = time.perf_counter()
        res = IO.index(orc, ucc, w["gv"], w["ds"], 0, 1, minpks=20, hkl_tol=0.02,
                       max_grains=2 * K)
        cpu_s = time.perf_counter() - t0
        out["cpu_hits"] = len(res["hits"])
        out["cpu"] = {"kind": "port", "threads": 1, "seconds": cpu_s,
                      "sample": "the whole search, sequential C restatement + numpy"}
        out["speedup_vs_cpu"] = cpu_s * 1e3 / out["total_ms"]
        out["same_grains_and_scores_as_cpu"] = bool(
            len(res["ubis"]) == len(ind.ubis) and res["scores"] == ind.scores
            and np.array_equal(res["ga"], ind.ga))
    return out


def aux_refinegrains(with_cpu, K=2000, maxiters=100, cpu_grains=10, seed=21):
    """refinegrains.refinepositions for K grains (translations 15 microns
    off) on a binned-omega peak table: all K simplexes stepped together on
    the GPU (omega-float on), against the oracle's grain-by-grain loop (the
    reference's call pattern, one thread) on the first `cpu_grains` grains,
    whose per-grain time is scaled to K."""
    import numpy as np
    import torch
    from imaged11_amd import columnfile, simulate
    from imaged11_amd import refinegrains as RG
    t0 = time.perf_counter()
    w = simulate.refinegrains_workload(K=K, N=600 * K, seed=seed)
    gen_s = time.perf_counter() - t0
    pk = w["peaks"]
    rng = np.random.default_rng(seed)
    start = w["trans"] + rng.uniform(-15, 15, size=w["trans"].shape)

    def make():
        rg = RG.refinegrains(tolerance=0.05, OmFloat=True, OmSlop=0.25)
        rg.parameterobj.parameters.update(w["pars"])
        for i in range(K):
            rg.grainnames.append(i)
            rg.ubisread[i] = w["ubi"][i]
            rg.translationsread[i] = start[i]
        rg.loadfiltered(columnfile.colfile_from_dict(
            {"sc": pk["sc"], "fc": pk["fc"], "omega": pk["omega"]}))
        rg.generate_grains()
        return rg
    make().refinepositions(maxiters=3)  # warm the code paths
    rg = make()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = rg.refinepositions(maxiters=maxiters)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    s = rg.scannames[0]
    err = np.array([rg.grains[(i, s)].translation for i in range(K)]) - w["trans"]
    iters = np.array([r[2] for r in res.values()])
    out = {"grains": K, "peaks": int(len(pk["sc"])), "gen_s": gen_s, "maxiters": maxiters,
           "gpu_s": gpu_s, "grains_per_s": K / gpu_s, "evaluation_rounds": rg.evaluation_rounds,
           "ms_per_round": gpu_s * 1e3 / max(rg.evaluation_rounds, 1),
           "median_iterations": float(np.median(iters)),
           "median_translation_error_um": float(np.median(np.abs(err).max(1)))}
    if with_cpu:
        from oracle import oracle, refinegrains_oracle as R
        orc = oracle.load_oracle()
        sub = slice(0, cpu_grains)
        o = R.RefineGrains(orc, dict(rg.parameterobj.parameters), pk["sc"], pk["fc"],
                           pk["omega"], w["ubi"], start, tolerance=0.05, OmFloat=True,
                           OmSlop=0.25)
        # the sample grains' peak lists as assignlabels fixed them
        for key in range(K)[sub]:
            gr, ind = o.grains[key], rg.grains[(key, s)].ind
            gr.ind = ind
            gr.sc, gr.fc, gr.om = pk["sc"][ind], pk["fc"][ind], pk["omega"][ind]
            gr.peaks_xyz = R.compute_xyz_lab([gr.sc, gr.fc], **o.parameters).T.copy()
        steps = {k: rg.parameterobj.stepsizes[k] for k in ("t_x", "t_y", "t_z")}
        o.tolerance = 1.0
        o.varylist = ["t_x", "t_y", "t_z"]
        worst = 0.0
        t0 = time.perf_counter()
        for key in range(K)[sub]:
            # refinegrains.refinepositions' loop body for one grain
            o.grains_to_refine = [key]
            o.set_translation(key)
            guess = [o.parameters[n] for n in o.varylist]
            R.Simplex(o.gof, guess, [steps[n] for n in o.varylist]).minimize(maxiters=maxiters)
            last = np.array([o.parameters["t_x"], o.parameters["t_y"], o.parameters["t_z"]])
            worst = max(worst, float(np.abs(last - rg.grains[(key, s)].translation).max()))
        cpu_s = time.perf_counter() - t0
        per_grain = cpu_s / cpu_grains
        same = worst < 1e-6
        out["cpu"] = {"kind": "port", "threads": 1, "grains": cpu_grains, "seconds": cpu_s,
                      "seconds_per_grain": per_grain,
                      "sample": "first %d grains, one simplex each, sequential" % cpu_grains}
        out["cpu_s_scaled_to_all_grains"] = per_grain * K
        out["speedup_vs_cpu"] = per_grain * K / gpu_s
        out["sample_translations_equal_gpu_within_1e-6_um"] = bool(same)
    return out


def aux_pbp(with_cpu, NY=101, grains=24, cpu_points=40):
    """point_by_point.refine_map on an NY x NY map (a disc of voxels in
    `grains` Voronoi grains, one detector peak per grain, reflection and
    stage position): the whole map on
    the GPU (selection + first scoring of every (voxel, UBI) task in one
    kernel, the rest vectorised over tasks), against the oracle's voxel loop
    (the reference's per-voxel code, one thread) on the first `cpu_points`
    voxels, scaled to the map."""
    import numpy as np
    import torch
    from imaged11_amd import point_by_point as PP, simulate
    t0 = time.perf_counter()
    w = simulate.pbp_workload(NY=NY, grains=grains, gmax=1.0, seed=41)
    gen_s = time.perf_counter() - t0
    keys = ("refine_points", "all_pbpmap_ubis", "ri_col", "rj_col", "sx_grid", "sy_grid", "mask",
            "sc", "fc", "eta", "sum_intensity", "sinomega", "cosomega", "omega", "dty", "dtyi",
            "xpos", "ystep", "y0", "B0")
    args = [w[k] for k in keys]
    PP.refine_map(*args, **w["geometry"])  # warm
    runs = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        U, E, N, Q = PP.refine_map(*args, **w["geometry"])
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - t0)
    gpu_s = statistics.median(runs)
    M = w["all_pbpmap_ubis"].shape[2]
    voxels = int(w["mask"].sum())
    npeaks = len(w["sc"])
    # the selection kernel alone (both passes), timed with events
    dev = torch.device("cuda")
    T_ = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                     device=dev)
    pts, ptask, ubis = [], [0], []
    order = np.lexsort((w["rj_col"], w["ri_col"]))
    last = None
    for m in order:
        u = w["all_pbpmap_ubis"][:, :, m]
        if np.isnan(u[0, 0]):
            continue
        key = (w["ri_col"][m], w["rj_col"][m])
        if key != last:
            pts.append((w["sx_grid"][key], w["sy_grid"][key]))
            ptask.append(ptask[-1])
            last = key
        ubis.append(u)
        ptask[-1] += 1
    gv = PP._gve_device(T_(w["sc"]), T_(w["fc"]), T_(w["omega"]), T_(w["xpos"]), w["geometry"])
    dargs = (T_(w["sinomega"]), T_(w["cosomega"]), T_(w["dty"]), gv, T_(np.array(pts)),
             T_(np.array(ptask, np.int32), torch.int32), T_(np.array(ubis)), w["y0"],
             w["ystep"], 0.1)
    assign_ms = _events_ms(lambda: PP.assign_tasks(*dargs), reps=3)
    out = {"map": "%dx%d" % (NY, NY), "grains": grains, "voxels": voxels, "ubis": M,
           "peaks": npeaks,
           "gen_s": gen_s, "gpu_s": gpu_s, "gpu_s_runs": runs, "voxels_per_s": voxels / gpu_s,
           "selection_ms": assign_ms,
           "selection_voxel_peak_pairs_per_s": len(pts) * npeaks / (assign_ms * 1e-3),
           "refined": int((~np.isnan(E[0, 0])).sum())}
    if with_cpu:
        from oracle import pbp_oracle as P
        inmask = [tuple(p) for p in w["refine_points"] if w["mask"][tuple(p)]]
        sub_points = np.array(inmask[:cpu_points])
        sargs = [sub_points if k == "refine_points" else w[k] for k in keys]
        t0 = time.perf_counter()
        Uc, Ec, Nc, Qc = P.refine_map(*sargs, **w["geometry"])
        cpu_s = time.perf_counter() - t0
        done = ~np.isnan(Nc)
        same = bool(np.array_equal(Nc[done], N[done]) and
                    np.allclose(Uc[:, :, done], U[:, :, done], atol=1e-9))
        out["cpu"] = {"kind": "port", "threads": 1, "voxels": len(sub_points), "seconds": cpu_s,
                      "sample": "first %d voxels of the map, the oracle's voxel loop" %
                                len(sub_points)}
        out["cpu_s_scaled_to_map"] = cpu_s / len(sub_points) * voxels
        out["speedup_vs_cpu"] = out["cpu_s_scaled_to_map"] / gpu_s
        out["sample_results_equal_gpu"] = same
    return out


def aux_geometry(threads, with_cpu, N=8_000_000, seed=7):
    """The updateGeometry pass of a column file (compute_xlylzl then
    compute_geometry: xl, yl, zl, tth, eta, ds, gx, gy, gz of every peak)
    over N raw peaks in one fused kernel (k_sf2geometry), against the
    reference C compiled here (compute_xlylzl serial as shipped,
    compute_geometry with OpenMP) on the same peaks.  Algorithmic HBM bytes
    per peak: 24 in (sc, fc, omega) + 72 out."""
    import numpy as np
    import torch
    from imaged11_amd import _lib, geometry, simulate
    from imaged11_amd._lib import darr, dptr, stream_ptr
    rng = np.random.default_rng(seed)
    pars = simulate.detector_parameters(rng, wedge=0.3, chi=-0.2)
    sc, fc = rng.uniform(0, 2048, N), rng.uniform(0, 2048, N)
    om = rng.uniform(-180, 180, N)
    t = np.array([12.0, -7.0, 3.0])
    dev = torch.device("cuda")
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    s_d, f_d, w_d = T(sc), T(fc), T(om)
    xl = torch.empty((N, 3), dtype=torch.float64, device=dev)
    out = torch.empty((N, 6), dtype=torch.float64, device=dev)
    p, r, dist = geometry.pack_xlylzl(pars)
    sign, wvln, wedge, chi = geometry.pack_gv(pars)
    lib = _lib.load()
    args = (darr(p), darr(r), darr(dist), float(sign), float(wvln), float(wedge), float(chi),
            darr(t))

    def run():
        _lib.check(lib.gs_sf2geometry(dptr(s_d), dptr(f_d), dptr(w_d), *args, dptr(xl),
                                      dptr(out), N, stream_ptr()), "sf2geometry")
    ms = _events_ms(run, reps=10)
    res = {"peaks": N, "gpu_ms": ms, "peaks_per_s": N / (ms * 1e-3),
           # 24 B in, 72 B out per peak
           "algorithmic_hbm_gb_s": 96.0 * N / (ms * 1e-3) / 1e9}
    if with_cpu:
        from oracle import oracle
        ref = oracle.load_reference()
        kind = "reference"
        if ref is None:
            ref, kind = oracle.load_oracle(), "port"
        xl_h = np.zeros((N, 3))
        out_h = np.zeros((N, 6))
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ref.compute_xlylzl(sc, fc, p, r, dist, xl_h)
            ref.compute_geometry(xl_h, om, sign, wvln, wedge, chi, t, out_h)
            runs.append(time.perf_counter() - t0)
        cpu_s = statistics.median(runs)
        got_xl, got = xl.cpu().numpy(), out.cpu().numpy()
        res["cpu"] = {"kind": kind, "threads": threads, "seconds": cpu_s,
                      "sample": "all %d peaks, median of 3" % N}
        res["speedup_vs_cpu"] = cpu_s / (ms * 1e-3)
        # xl bit-exact; ds and g within device-vs-glibc sin/cos rounding; tth,
        # eta within the two libraries' atan2
        res["xl_bit_exact"] = bool(np.array_equal(got_xl, xl_h))
        res["max_rel_diff_ds_g"] = float(np.max(np.abs(got[:, 2:] - out_h[:, 2:]) /
                                                np.maximum(np.abs(out_h[:, 2:]), 1e-300)))
        d = np.abs(got[:, :2] - out_h[:, :2])
        d[:, 1] = np.minimum(d[:, 1], 360.0 - d[:, 1])  # eta: across the +-180 cut
        res["max_abs_diff_tth_eta_deg"] = float(d.max())
    return res


def aux_scan(threads, with_cpu, K=20000, N=8_000_000, n_iter=20, cpu_peaks=50_000):
    """Large scan: compute_gv per grain translation + score_and_assign for
    K grains x N peaks (per-peak dty, 15% spurious) in one fused launch,
    then n_iter captured score_and_refine iterations on the assigned peaks.
    Timed with the parity mode's sin/cos; the fast mode's assignment is run
    beside it (untimed) and compared.  The CPU leg runs the reference's
    compute_gv + score_and_assign for every grain over a seeded sample of
    `cpu_peaks` peaks (its own OpenMP loops, as assignlabels calls them),
    reports the per-pair rate and checks the GPU's labels and drlv2 on
    those peaks."""
    import numpy as np
    import torch
    from imaged11_amd import drivers, geometry, refine, simulate

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    wl = simulate.scan_workload(K=K, N=N, seed=5, device="cuda")
    gen_s = time.perf_counter() - t0
    pk = wl["peaks"]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    setup = {}
    scans = {}
    for mode in ("parity", "fast"):  # parity: sin/cos of omega from the host's glibc
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scans[mode] = drivers.Scan(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), wl["pars"],
                                   dty=T(pk["dty"]), parity=mode == "parity")
        torch.cuda.synchronize()
        setup[mode] = (time.perf_counter() - t0) * 1e3
    scan = scans["parity"]
    ubis, trans = T(wl["ubi"]), T(wl["trans"])
    drivers.assignlabels(scan, ubis[:64].contiguous(), trans[:64].contiguous(), 0.05)  # warm
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    labels, drlv2, counts = drivers.assignlabels(scan, ubis, trans, 0.05)
    b.record()
    torch.cuda.synchronize()
    assign_ms = a.elapsed_time(b)
    # the same assignment from the fast mode's sin/cos (outside the timed region)
    lab_f, d2_f, cnt_f = drivers.assignlabels(scans["fast"], ubis, trans, 0.05)
    lab_p_h, d2_p_h = labels.cpu().numpy(), drlv2.cpu().numpy()
    lab_f_h, d2_f_h = lab_f.cpu().numpy(), d2_f.cpu().numpy()
    relf = np.abs(d2_f_h - d2_p_h) / np.maximum(np.abs(d2_p_h), 1e-300)
    fast_vs_parity = {
        "labels_equal": bool(np.array_equal(lab_f_h, lab_p_h)),
        "counts_equal": bool(torch.equal(cnt_f, counts)),
        "drlv2_bit_equal_fraction": float(np.mean(d2_f_h == d2_p_h)),
        "drlv2_max_rel_diff": float(relf.max()),
        "drlv2_peaks_rel_diff_above_1e-12": int((relf > 1e-12).sum()),
        "sincos_bit_equal_fraction": float(
            (torch.eq(scans["fast"].so, scan.so) & torch.eq(scans["fast"].co, scan.co))
            .double().mean())}
    del scans["fast"], lab_f, d2_f
    gvl, idx, off = drivers.grain_gvectors(scan, trans, labels, K)
    # the captured iterations in the autotuner's launch shape for these lists
    best = refine.autotune(ubis, gvl, None, off, 0.05, n_iter, reps=2)
    work = ubis.clone()
    g = refine.RefineGraph(work, gvl, None, off, 0.05, n_iter,
                           lanes_per_grain=best["lanes_per_grain"],
                           waves_per_cu=best["waves_per_cu"],
                           long_min=best.get("long_min", 0),
                           mid_min=best.get("mid_min", 0))
    refine_ms = _events_ms(lambda: (work.copy_(ubis), g.replay()), reps=3)
    own = pk["grain"]
    lab = lab_p_h
    # per peak: xl 24 B + sin/cos 16 B + (drlv2, label) read 12 B and
    # written 12 B; per grain: UBI 72 B + translation 24 B + label 4 B +
    # count 4 B
    algo_bytes = 64 * N + 104 * K
    out = {"grains": K, "peaks": N, "gen_s": gen_s, "assign_ms": assign_ms,
           "trig_mode_timed": "parity (sin/cos of omega from the host's glibc)",
           "scan_setup_ms": setup,
           "assign_pairs_per_s": K * N / (assign_ms * 1e-3),
           "assign_algorithmic_hbm_gb_s": algo_bytes / (assign_ms * 1e-3) / 1e9,
           "assign_valu_ops_per_pair": SCAN_VALU_OPS_PER_PAIR,
           "assign_issued_fp64_ops_frac_of_nonfma_peak":
               SCAN_VALU_OPS_PER_PAIR * K * N / (assign_ms * 1e-3) / 39.3e12,
           "assigned": int((lab >= 0).sum()),
           "own_label_fraction": float(np.mean(lab[own >= 0] == own[own >= 0])),
           "fast_mode_vs_parity_mode": fast_vs_parity,
           "refine_iterations": n_iter, "refine_list_entries": int(gvl.shape[0]),
           "refine_graph_ms": refine_ms,
           "refine_graph_shape": {"lanes_per_grain": best["lanes_per_grain"],
                                  "waves_per_cu": best["waves_per_cu"],
                                  "long_min": best.get("long_min", 0),
                                  "mid_min": best.get("mid_min", 0)}}
    if with_cpu:
        from oracle import oracle
        ref = oracle.load_reference()
        kind = "reference"
        if ref is None:
            ref, kind = oracle.load_oracle(), "port"
        else:
            ref.set_threads(threads)
        # every grain over a seeded sample of peaks: the reference's
        # assignlabels loop (compute_gv with the grain's translation, then
        # score_and_assign) from labels -1 / drlv2 1.  A peak's label and
        # drlv2 depend on that peak alone, so the sample's final state must
        # equal the GPU's on the same peaks, bit for bit.
        rs = np.random.default_rng(55)
        sel = np.sort(rs.choice(N, size=min(cpu_peaks, N), replace=False))
        xl = scan.xl.cpu().numpy()[sel]
        om = np.ascontiguousarray(pk["omega"][sel])
        sign, wvln, wedge, chi = geometry.pack_gv(wl["pars"])
        d = np.ones(len(sel))
        lb = np.full(len(sel), -1, np.int32)
        gv = np.empty((len(sel), 3))
        t0 = time.perf_counter()
        for k in range(K):
            ref.compute_gv(xl, om, sign, wvln, wedge, chi, wl["trans"][k], gv)
            ref.score_and_assign(wl["ubi"][k], gv, 0.05, d, lb, k)
        cpu_s = time.perf_counter() - t0
        out["cpu"] = {"kind": kind, "threads": threads, "grains": K, "peaks": len(sel),
                      "seconds": cpu_s, "pairs_per_s": K * len(sel) / cpu_s,
                      "sample": "all %d grains x a seeded sample of %d peaks" % (K, len(sel))}
        out["speedup_vs_cpu"] = out["assign_pairs_per_s"] / out["cpu"]["pairs_per_s"]
        out["parity_vs_reference"] = {
            "labels_bit_exact": bool(np.array_equal(lab_p_h[sel], lb)),
            "drlv2_bit_exact": bool(np.array_equal(d2_p_h[sel].view(np.int64), d.view(np.int64))),
            "fast_mode_labels_equal": bool(np.array_equal(lab_f_h[sel], lb)),
            "sample": "the CPU leg's %d peaks, final state after all %d grains" % (len(sel), K)}
    g.close()
    return out


def profile_workload(which, copy_mb, launches=5):
    """The refinement or adjoint workload of aux_refine_adjoint (5000 grains,
    lognormal lists from a 2e6-peak table), launched a few times with a
    fixed shape, for rocprofv3: the batched refinement is 10 fused
    score_and_refine iterations at 16 lanes per grain, 4 waves per CU,
    lists of 853 entries or more on 64 lanes and of 524 or more on 32 (the
    shape the autotuner picks on MI355X); the adjoint is one evaluation
    with the Gauss-Newton sums at the default pieces and lanes.  "scan":
    aux_scan's parity-mode assignlabels of 20000 grains over 8e6 peaks,
    twice."""
    import numpy as np
    import torch
    from imaged11_amd import adjoint, batched, drivers, simulate

    dev = torch.device("cuda:0")
    if copy_mb:
        src = torch.ones(copy_mb * 131072, dtype=torch.float64, device=dev)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        del src
    if which == "scan":
        K, N = 20000, 8_000_000
        wl = simulate.scan_workload(K=K, N=N, seed=5, device="cuda")
        pk = wl["peaks"]
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
        scan = drivers.Scan(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), wl["pars"],
                            dty=T(pk["dty"]), parity=True)
        ubis, trans = T(wl["ubi"]), T(wl["trans"])
        for _ in range(2):
            drivers.assignlabels(scan, ubis, trans, 0.05)
        torch.cuda.synchronize()
        print(json.dumps({"profiled": which, "launches": 2, "grains": K, "peaks": N}),
              flush=True)
        return
    K, N = 5000, 2_000_000
    wl = simulate.scoring_workload(K=K, N=N, seed=3)
    idx, off = simulate.refine_lists(wl["peaks"]["grain"], K, seed=3)
    pk, pars = wl["peaks"], wl["pars"]
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    rng = np.random.default_rng(4)
    ubis0 = T(wl["ubi"] @ (np.eye(3) + rng.normal(scale=1e-3, size=(K, 3, 3))))
    idx_d, off_d = T(idx, torch.int32), T(off, torch.int64)
    if which == "refine":
        gv = batched.sf2gv(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), pars)
        order = batched.list_order(off_d)
        for _ in range(launches):
            batched.refine_lists(ubis0, gv, idx_d, off_d, 0.05, 10, order=order,
                                 lanes_per_grain=16, waves_per_cu=4, long_min=853,
                                 mid_min=524)
    else:
        prob = adjoint.JointProblem(T(pk["sc"]), T(pk["fc"]), T(pk["omega"]), pars["omegasign"],
                                    idx_d, off_d)
        params = {k: torch.tensor(float(pars[k]), dtype=torch.float64, device=dev)
                  for k in adjoint.FREE[:8]}
        trans = torch.zeros((K, 3), dtype=torch.float64, device=dev)
        geo = adjoint.pack_geometry(params, pars)
        prob.assign_hkl(geo, ubis0, trans)
        for _ in range(launches):
            prob.evaluate(geo, ubis0, trans, want_sums=True)
    torch.cuda.synchronize()
    print(json.dumps({"profiled": which, "launches": launches, "grains": K, "peaks": N,
                      "list_entries": int(off[-1])}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU; the multi-GPU path is design-only (peak "
              "sharding + an int32 all-reduce of the K counts). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from imaged11_amd import _lib, batched, simulate

    _lib.load()
    _lib.require_gpu()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.profile:
        profile_workload(args.profile, args.copy_mb)
        return

    t0 = time.perf_counter()
    wl = simulate.scoring_workload(K=args.grains, N=args.peaks, seed=args.seed)
    gen_s = time.perf_counter() - t0
    pk, pars = wl["peaks"], wl["pars"]

    def T(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)

    if args.copy_mb:
        src = torch.ones(args.copy_mb * 131072, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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

    t_warm = time.perf_counter()
    warm_steps = 0
    while warm_steps < args.warmup or time.perf_counter() - t_warm < args.warmup_s:
        for _ in range(10):
            step()
        warm_steps += 10
        torch.cuda.synchronize()
    warm_s = time.perf_counter() - t_warm
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
    if batched.score_mode(args.tol) == _lib.SCORE_SIEVE:
        kernel = "k_score_many_sieve"
        g = batched.sf2gv(sc, fc, om, pars)
        passed = 0.0
        for k0 in range(0, min(args.grains, 64), 16):  # pairs passing row 0, 64 grains
            h0 = ubis[k0:k0 + 16, 0, :] @ g.T
            passed += float(((h0 - torch.round(h0)).abs() < args.tol).sum())
            del h0
        pass_frac = passed / (min(args.grains, 64) * args.peaks)
        valu_ops = SIEVE_STAGE1_VALU + SIEVE_SETTLE_VALU * pass_frac
        del g
    else:
        kernel, pass_frac, valu_ops = "k_score_many_fast", None, VALU_OPS_PER_PAIR
    gpu_counts = counts.cpu().numpy()
    # boundary-peak audit of the device-sincos g-vectors (outside the timed
    # region): pairs within 1e-12 relative of tol^2, where a last-bit
    # difference from the host libm's sin/cos could flip a count
    boundary = int(batched.boundary_pairs(ubis, batched.sf2gv(sc, fc, om, pars),
                                          args.tol).sum())
    # the same counts from the parity mode's g-vectors (sin/cos of omega from
    # the host's glibc: the reference's compute_gv bit for bit), untimed
    parity_counts = batched.score_many(ubis, batched.sf2gv(sc, fc, om, pars, parity=True),
                                       args.tol).cpu().numpy()

    out = {
        "metric": "grain\u00d7peak pairs/s scored; batched+adjoint refine vs per-grain loop, "
                  "same GPU",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_s": warm_s,
        "warmup_steps_run": warm_steps,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded forward simulation, generated in %.1f s)" % gen_s,
        "config": {"workload": "raw peaks -> compute_xlylzl -> compute_gv -> score, "
                               "%d grains x %d peaks, tol %g" % (args.grains, args.peaks, args.tol),
                   "grains": args.grains, "peaks": args.peaks, "tol": args.tol,
                   "boundary_pairs_rel_1e-12": boundary,
                   "trig_mode_timed": "fast (device sincos)",
                   "counts_equal_parity_mode": bool(np.array_equal(parity_counts, gpu_counts)),
                   "value_is": "the pairs/s of this workload; the batched+adjoint refine "
                               "against the per-grain GPU loop is "
                               "aux.refine_5000_grains.speedup_refine_plus_adjoint_vs_gpu_loop"},
        "roofline": {
            "kernel": kernel,
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64_VALU / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_VALU,
            "frac_is": "algorithmic: flops_per_pair for every pair decided, over the FP64 "
                       "peak; the kernel issues valu_ops_per_pair instructions per pair, "
                       "and issued_valu_ops_frac_of_nonfma_peak is their issue rate",
            "traffic": measured_traffic(kernel),
            "kernel_ms": score_ms,
            "flops_per_pair": FLOPS_PER_PAIR,
            "valu_ops_per_pair": valu_ops,
            "row0_pass_fraction": pass_frac,
            "issued_valu_ops_frac_of_nonfma_peak": valu_ops * pairs / (score_ms * 1e-3)
            / (PEAK_FP64_VALU / 2),
        },
        "cpu_baseline": None,
    }
    threads, cpu_facts = host_cpus(args.cpu_threads)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.tol, threads, gpu_counts)
        out["cpu_baseline"]["host"] = cpu_facts
    if not args.no_aux:
        with_cpu = not args.no_cpu
        legs = (("refine_5000_grains", lambda: aux_refine_adjoint(threads, with_cpu)),
                ("scan_20000_grains", lambda: aux_scan(threads, with_cpu)),
                ("index_3000_grains", lambda: aux_index(with_cpu)),
                ("refinegrains_2000_grains", lambda: aux_refinegrains(with_cpu)),
                ("pbp_refine_map_101x101", lambda: aux_pbp(with_cpu)),
                ("geometry_8e6_peaks", lambda: aux_geometry(threads, with_cpu)))
        out["aux"] = {}
        for name, leg in legs:
            t0 = time.perf_counter()
            out["aux"][name] = leg()
            # progress on stderr; the JSON line stays the only stdout
            print("bench: %s done in %.1f s" % (name, time.perf_counter() - t0),
                  file=sys.stderr, flush=True)
        out["refine_plus_adjoint_vs_per_grain_gpu_loop"] = \
            out["aux"]["refine_5000_grains"]["speedup_refine_plus_adjoint_vs_gpu_loop"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
