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
import numpy as np


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
        """Returns (points (B,n), errors (B,), iterations (B,))."""
        B, n = self.B, self.n
        S, E = self.vertices, self.errors
        ar = np.arange(B)
        done = np.zeros(B, dtype=bool)
        lowest = np.full(B, -1)
        iters = np.zeros(B, dtype=np.int64)
        for it in range(maxiters):
            act = ~done
            if not act.any():
                break
            hi = np.zeros(B, dtype=np.int64)
            lo = np.zeros(B, dtype=np.int64)
            for v in range(1, n + 1):
                hi = np.where(E[:, v] > E[ar, hi], v, hi)
                lo = np.where(E[:, v] < E[ar, lo], v, lo)
            second = lo.copy()
            for v in range(n + 1):
                upd = (v != hi) & (v != second) & (E[:, v] > E[ar, second])
                second = np.where(upd, v, second)
            tot = np.zeros(B)
            for v in range(n + 1):
                tot = tot + E[:, v]
            mean = tot / (n + 1)
            ss = np.zeros(B)
            for v in range(n + 1):
                d = E[:, v] - mean
                ss = ss + d * d
            conv = act & (np.sqrt(ss / n) <= epsilon)
            lowest[act] = lo[act]
            iters[act] = it
            done |= conv
            a = act & ~conv
            if not a.any():
                continue
            # centroid of all vertices but the highest, in vertex order
            csum = np.zeros((B, n))
            for v in range(n + 1):
                csum = np.where((v != hi)[:, None], csum + S[:, v, :], csum)
            cen = csum / n
            S[a, n + 1] = cen[a]
            refl = self.kR * S[ar, hi] + (1 - self.kR) * cen
            S[a, n + 2] = refl[a]
            fr = np.full(B, np.nan)
            fr[a] = self._call(refl[a], a)
            acc = a & (fr < E[ar, hi])
            E[acc, hi[acc]] = fr[acc]
            S[acc, hi[acc]] = S[acc, n + 2]
            expand = a & (fr <= E[ar, lo])
            contract = a & ~expand & (fr >= E[ar, second])
            trial = np.where(expand[:, None], self.kE * refl + (1 - self.kE) * cen,
                             self.kC * S[ar, hi] + (1 - self.kC) * cen)
            step = expand | contract
            fb = np.full(B, np.nan)
            fb[step] = self._call(trial[step], step)
            take = step & (fb < E[ar, hi])
            E[take, hi[take]] = fb[take]
            S[take, hi[take]] = trial[take]
            shrink = contract & ~take
            if shrink.any():
                base = S[ar, lo]
                for v in range(n + 1):
                    m = shrink & (lo != v)
                    S[m, v] = 0.5 * (S[m, v] + base[m])
                for v in range(n + 1):
                    m = shrink & (lo != v)
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

