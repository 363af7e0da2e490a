"""Several host threads calling the library at once, each on its own
stream: the reference functions are re-entrant, and the drop-in must be
too (host-pointer staging is per thread, device entry points take the
caller's stream and, for the persistent kernels, the caller's work-queue
counter)."""
import threading

import numpy as np
import pytest
import torch

from imaged11_amd import batched, simulate
from imaged11_amd import cImageD11 as gpu

pytestmark = pytest.mark.gpu


def _run_threads(fn, n):
    errors, out = [], [None] * n

    def body(k):
        try:
            out[k] = fn(k)
        except BaseException as e:  # noqa: BLE001 - reported below
            errors.append(e)
    ts = [threading.Thread(target=body, args=(k,)) for k in range(n)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    if errors:
        raise errors[0]
    return out


def test_dropin_calls_from_threads(orc):
    """score and score_and_refine on host arrays from 6 threads, 10 calls
    each, interleaved: every result equals the oracle's."""
    cases = []
    for k in range(6):
        ubi, gv, _ = simulate.lattice_gvectors(seed=40 + k, n_lattice=3000, n_random=2000)
        u = ubi.copy()
        want_n = orc.score(ubi, gv, 0.05)
        want_r = orc.score_and_refine(u, gv, 0.05)
        cases.append((ubi, gv, want_n, want_r, u))

    def work(k):
        ubi, gv, want_n, want_r, u_ref = cases[k]
        for _ in range(10):
            assert gpu.score(ubi, gv, 0.05) == want_n
            u = ubi.copy()
            n, m = gpu.score_and_refine(u, gv, 0.05)
            assert n == want_r[0]
            assert abs(m - want_r[1]) <= 1e-12 * abs(want_r[1])
            np.testing.assert_allclose(u, u_ref, rtol=0, atol=1e-10)
        return True
    assert all(_run_threads(work, 6))


def test_batched_calls_on_thread_streams(dev, orc):
    """score_many and refine_lists from 4 threads, each on its own stream
    (default priority), on shared read-only device inputs."""
    wl = simulate.scoring_workload(K=64, N=20000, seed=12)
    pk = wl["peaks"]
    gv = batched.sf2gv(torch.as_tensor(pk["sc"], device=dev), torch.as_tensor(pk["fc"], device=dev),
                       torch.as_tensor(pk["omega"], device=dev), wl["pars"])
    torch.cuda.synchronize()
    gvh = gv.cpu().numpy()
    ubis = torch.as_tensor(wl["ubi"], device=dev)
    want = orc.score_many(wl["ubi"], gvh, 0.05)
    idx, off = simulate.refine_lists(pk["grain"], 64, seed=3, median=100, lo=20, hi=400)
    idx_d = torch.as_tensor(idx, dtype=torch.int32, device=dev)
    off_d = torch.as_tensor(off, dtype=torch.int64, device=dev)
    base_u, _, base_m = batched.refine_lists(ubis, gv, idx_d, off_d, 0.05, 3)
    torch.cuda.synchronize()
    base_u, base_m = base_u.cpu().numpy(), base_m.cpu().numpy()

    def work(k):
        s = torch.cuda.Stream(device=dev)
        with torch.cuda.stream(s):
            for _ in range(5):
                got = batched.score_many(ubis, gv, 0.05)
                u, _, m = batched.refine_lists(ubis, gv, idx_d, off_d, 0.05, 3)
            s.synchronize()
            assert np.array_equal(got.cpu().numpy(), want)
            # the same arithmetic in the same order: bit for bit
            assert np.array_equal(u.cpu().numpy(), base_u)
            assert np.array_equal(m.cpu().numpy(), base_m)
        return True
    assert all(_run_threads(work, 4))
