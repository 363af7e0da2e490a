"""Nelder-Mead minimisation as ImageD11's simplex (``ImageD11/simplex.py``),
for one problem or for many problems stepped together.

`SimplexBatch` runs B independent simplexes in lockstep: every round asks
the objective for at most one point per problem, in the order the
sequential algorithm would ask for it (initial vertices 0..n; then per
iteration the reflected point, the expanded or contracted point, and after
a shrink the non-lowest vertices in vertex order).  A batched objective,
one GPU launch per round over all problems, therefore sees each problem's
calls in the same order, with the same state, as the sequential loop —
which matters to objectives with side effects, such as refinegrains'
gof, whose omega-float step reads the UBI the previous call left.

The algorithm is the reference's: n+1 vertices plus the centroid and the
reflected point, reflection -1, expansion 2, contraction 1/2, shrink
towards the lowest vertex, convergence when the standard deviation of the
vertex errors (divided by n) is <= epsilon; ties and NaN compare as the
reference's `<` / `>` tests do (a NaN never wins a comparison).  Results,
the returned vertex (the lowest vertex of the last iteration's ordering)
and the iteration count follow ``simplex.py:115-227`` step by step;
``tests/test_refinegrains_cpu.py`` checks them bit for bit against the
reference's own evaluation sequences.  The only arithmetic difference is
(e - mean)**2 written as a product: the reference's float ** 2 goes through
libm's pow, which can differ from the correctly rounded product by one
rounding in rare cases, and only the convergence test reads it.
"""
import ctypes

import numpy as np

from . import _lib


class SimplexBatch:
    """B problems of n variables.  ``evaluate(points, which)`` gets an
    (P, n) float64 array of points and the (P,) problem indices they belong
    to (each problem at most once) and returns P objective values."""

    def __init__(self, evaluate, guesses, increments, kR=-1, kE=2, kC=0.5):
        g = np.array(guesses, dtype=np.float64, ndmin=2)
        inc = np.broadcast_to(np.asarray(increments, dtype=np.float64), g.shape)
        self.evaluate = evaluate
        self.B, self.n = g.shape
        self.kR, self.kE, self.kC = float(kR), float(kE), float(kC)
        n = self.n
        self.vertices = np.repeat(g[:, None, :], n + 3, axis=1)
        for v in range(1, n + 1):
            self.vertices[:, v, v - 1] = g[:, v - 1] + inc[:, v - 1]
        self.errors = np.zeros((self.B, n + 1))
        self.last = g.copy()          # the point of each problem's latest call
        self.calls = np.zeros(self.B, dtype=np.int64)
        self.rounds = 0
        everyone = np.ones(self.B, dtype=bool)
        for v in range(n + 1):
            self.errors[everyone, v] = self._call(self.vertices[everyone, v], everyone)

    def _call(self, points, mask):
        which = np.nonzero(mask)[0]
        if len(which) == 0:
            return np.zeros(0)
        pts = np.ascontiguousarray(points, dtype=np.float64)
        vals = np.asarray(self.evaluate(pts, which), dtype=np.float64).reshape(len(which))
        self.last[which] = pts
        self.calls[which] += 1
        self.rounds += 1
        return vals

    def minimize(self, epsilon=0.0001, maxiters=250):
        """Returns (points (B,n), errors (B,), iterations (B,)).  Each
        iteration's bookkeeping between the objective calls runs in native
        code over all problems (gs_simplex_order, gs_simplex_reflected,
        gs_simplex_trial); the calls themselves go to the batched
        objective as before."""
        lib = _lib.load()
        B, n = self.B, self.n
        S, E = self.vertices, self.errors
        assert S.flags.c_contiguous and E.flags.c_contiguous
        ar = np.arange(B)
        done = np.zeros(B, dtype=np.uint8)
        lowest = np.full(B, -1, dtype=np.int64)
        iters = np.zeros(B, dtype=np.int64)
        hi, lo, second = (np.zeros(B, dtype=np.int64) for _ in range(3))
        a, expand, step, shrink = (np.zeros(B, dtype=np.uint8) for _ in range(4))
        fr = np.full(B, np.nan)
        fb = np.full(B, np.nan)
        trial = np.zeros((B, n))
        cnt = ctypes.c_int()
        p = lambda x: x.ctypes.data
        for it in range(maxiters):
            if done.all():
                break
            _lib.check(lib.gs_simplex_order(B, n, p(S), p(E), p(done), p(lowest), p(iters), it,
                                            float(epsilon), self.kR, p(hi), p(lo), p(second),
                                            p(a), ctypes.byref(cnt)), "simplex")
            if cnt.value == 0:
                continue
            am = a.view(bool)
            fr[am] = self._call(S[am, n + 2], am)
            _lib.check(lib.gs_simplex_reflected(B, n, p(S), p(E), p(a), p(hi), p(lo), p(second),
                                                p(fr), self.kE, self.kC, p(trial), p(expand),
                                                p(step), ctypes.byref(cnt)), "simplex")
            sm = step.view(bool)
            if cnt.value:
                fb[sm] = self._call(trial[sm], sm)
            _lib.check(lib.gs_simplex_trial(B, n, p(S), p(E), p(step), p(expand), p(hi), p(lo),
                                            p(fb), p(trial), p(shrink), ctypes.byref(cnt)),
                       "simplex")
            if cnt.value:
                km = shrink.view(bool)
                for v in range(n + 1):
                    m = km & (lo != v)
                    if m.any():
                        E[m, v] = self._call(S[m, v], m)
        # the lowest vertex of the last ordering; before any iteration the
        # reference's index -1 (the reflected-point slot, the last error)
        pick = np.where(lowest < 0, n + 2, lowest)
        epick = np.where(lowest < 0, n, lowest)
        return S[ar, pick].copy(), E[ar, epick].copy(), iters


class Simplex:
    """The reference's one-problem interface (``simplex.py:51-95``):
    ``Simplex(testfunc, guess, increments).minimize(epsilon, maxiters,
    monitor)`` -> (guess, error, iterations).  ``testfunc`` is called with
    the ``guess`` list itself, updated in place before each call."""

    def __init__(self, testfunc, guess, increments, kR=-1, kE=2, kC=0.5):
        self.testfunc = testfunc
        self.guess = guess

        def evaluate(points, which):
            for x in range(len(self.guess)):
                self.guess[x] = float(points[0, x])
            return [self.testfunc(self.guess)]
        self._batch = SimplexBatch(evaluate, [list(map(float, guess))], [increments], kR, kE, kC)

    def minimize(self, epsilon=0.0001, maxiters=250, monitor=0):
        pts, err, it = self._batch.minimize(epsilon, maxiters)
        for x in range(len(self.guess)):
            self.guess[x] = float(pts[0, x])
        if monitor:
            print("Iterations = %d   Best = %f" % (it[0], err[0]))
        return self.guess, float(err[0]), int(it[0])

