"""ORACLE — test infrastructure only.

Sequential CPU restatement of the refinegrains driver
(``ImageD11/refinegrains.py``) and the numpy geometry it calls
(``ImageD11/transform.py``, ``ImageD11/gv_general.py``), plus the
Nelder-Mead simplex it minimises with (``ImageD11/simplex.py``).  The C
routines on the path (compute_gv, score_and_assign, score_and_refine) come
from the oracle's C restatement (``oracle.load_oracle()``), called grain by
grain exactly as the reference's Python loops call them.

Only ``tests/`` and the CPU legs of ``bench.py`` import this module.

Pinning (what a plausible mistake would fail):
* the transform chain: ``compute_g_vectors(compute_tth_eta_from_xyz(...))``
  equals the compiled reference's ``compute_gv`` with the same translation
  to 1e-13 (an independent formula; a wrong sign, rotation order or
  translation frame fails it); ``uncompute_g_vectors`` inverts
  ``compute_g_vectors`` for the cases of the reference's own test
  (``test/test_uncomputegv.py``: four wedge/chi pairs, six peaks);
* ``Simplex``: the evaluation sequence and results of the reference's
  simplex on five problems, recorded by ``tests/golden/make_simplex_golden.py``
  (``tests/golden/simplex_golden.json``), bit for bit;
* the lattice-symmetry projections: cells of the projected UBIs have the
  symmetry's constraints, the orientation is unchanged, and triclinic is the
  identity;
* the driver (``RefineGrains``): gof = 0 at the true geometry of
  noise-free simulated data, and the C calls are the reference's per-grain
  calls in the reference's order.
``ubi_to_cell`` / ``ubi_to_u`` / ``u_to_ubi`` stand in for the absent
``xfab.tools`` functions of the same names; they are pinned by closed forms
(a UBI built as (U B(cell))^-1 from known cells, cubic to triclinic, returns
that cell and U, and u_to_ubi rebuilds it) and by the properties above;
bit-level agreement with xfab itself cannot be checked here (xfab absent).
"""
import math

import numpy as np

# ---------------------------------------------------------------------------
# transform.py


def detector_rotation_matrix(tilt_x, tilt_y, tilt_z):
    """transform.py:51-82: r3(x) . r2(y) . r1(z)."""
    r1 = np.array([[np.cos(tilt_z), -np.sin(tilt_z), 0], [np.sin(tilt_z), np.cos(tilt_z), 0],
                   [0, 0, 1]], float)
    r2 = np.array([[np.cos(tilt_y), 0, np.sin(tilt_y)], [0, 1, 0],
                   [-np.sin(tilt_y), 0, np.cos(tilt_y)]], float)
    r3 = np.array([[1, 0, 0], [0, np.cos(tilt_x), -np.sin(tilt_x)],
                   [0, np.sin(tilt_x), np.cos(tilt_x)]], float)
    return np.dot(np.dot(r3, r2), r1)


def compute_xyz_lab(peaks, y_center=0.0, y_size=0.0, tilt_y=0.0, z_center=0.0, z_size=0.0,
                    tilt_z=0.0, tilt_x=0.0, distance=0.0, o11=1.0, o12=0.0, o21=0.0, o22=-1.0,
                    **kwds):
    """transform.py:85-147: (2, n) detector (sc, fc) -> (3, n) lab."""
    r2r1 = detector_rotation_matrix(tilt_x, tilt_y, tilt_z)
    pd = np.array(peaks, float)
    pd[0, :] = (pd[0, :] - z_center) * z_size
    pd[1, :] = (pd[1, :] - y_center) * y_size
    flipped = np.dot(np.array([[o11, o12], [o21, o22]], float), pd)
    vec = np.array([np.zeros(flipped.shape[1]), flipped[1, :], flipped[0, :]], float)
    rotvec = np.dot(r2r1, vec)
    rotvec[0, :] = rotvec[0, :] + distance
    return rotvec


def compute_grain_origins(omega, wedge=0.0, chi=0.0, t_x=0.0, t_y=0.0, t_z=0.0):
    """transform.py:359-416: lab position of the grain at each omega."""
    t = np.zeros((3, omega.shape[0]), float)
    om_r = np.radians(omega)
    t[0, :] = np.cos(om_r) * t_x - np.sin(om_r) * t_y
    t[1, :] = np.sin(om_r) * t_x + np.cos(om_r) * t_y
    t[2, :] = t_z
    if chi != 0.0:
        c = np.cos(np.radians(chi))
        s = np.sin(np.radians(chi))
        u = np.zeros(t.shape, float)
        u[0, :] = t[0, :]
        u[1, :] = c * t[1, :] + -s * t[2, :]
        u[2, :] = s * t[1, :] + c * t[2, :]
        t = u
    if wedge != 0.0:
        c = np.cos(np.radians(wedge))
        s = np.sin(np.radians(wedge))
        u = np.zeros(t.shape, float)
        u[0, :] = c * t[0, :] + -s * t[2, :]
        u[1, :] = t[1, :]
        u[2, :] = s * t[0, :] + c * t[2, :]
        t = u
    return t


def compute_tth_eta_from_xyz(peaks_xyz, omega, t_x=0.0, t_y=0.0, t_z=0.0, wedge=0.0, chi=0.0,
                             **kwds):
    """transform.py:203-239 (degrees)."""
    if omega is None or (t_x == 0.0 and t_y == 0 and t_z == 0):
        s1 = peaks_xyz
    else:
        s1 = peaks_xyz - compute_grain_origins(omega, wedge, chi, t_x, t_y, t_z)
    eta = np.degrees(np.arctan2(-s1[1, :], s1[2, :]))
    s1_perp_x = np.sqrt(s1[1, :] * s1[1, :] + s1[2, :] * s1[2, :])
    tth = np.degrees(np.arctan2(s1_perp_x, s1[0, :]))
    return tth, eta


def compute_k_vectors(tth, eta, wvln):
    """transform.py:456-472."""
    tth = np.radians(tth)
    eta = np.radians(eta)
    c = np.cos(tth / 2)
    s = np.sin(tth / 2)
    ds = 2 * s / wvln
    k = np.zeros((3, tth.shape[0]), float)
    k[0, :] = -ds * s
    k[1, :] = -ds * c * np.sin(eta)
    k[2, :] = ds * c * np.cos(eta)
    return k


def compute_g_from_k(k, omega, wedge=0, chi=0):
    """transform.py:486-526: g = Omega . chi . wedge . k."""
    om = np.radians(omega)
    g = np.zeros((3, k.shape[1]), float)
    t = np.zeros((3, k.shape[1]), float)
    if wedge != 0.0:
        c = np.cos(np.radians(wedge))
        s = np.sin(np.radians(wedge))
        t[0, :] = c * k[0, :] + s * k[2, :]
        t[1, :] = k[1, :]
        t[2, :] = -s * k[0, :] + c * k[2, :]
        k = t.copy()
    if chi != 0.0:
        c = np.cos(np.radians(chi))
        s = np.sin(np.radians(chi))
        t[0, :] = k[0, :]
        t[1, :] = c * k[1, :] + s * k[2, :]
        t[2, :] = -s * k[1, :] + c * k[2, :]
        k = t.copy()
    g[0, :] = np.cos(om) * k[0, :] + np.sin(om) * k[1, :]
    g[1, :] = -np.sin(om) * k[0, :] + np.cos(om) * k[1, :]
    g[2, :] = k[2, :]
    return g


def compute_g_vectors(tth, eta, omega, wvln, wedge=0.0, chi=0.0):
    """transform.py:475-483."""
    return compute_g_from_k(compute_k_vectors(tth, eta, wvln), omega, wedge, chi)


# ---------------------------------------------------------------------------
# gv_general.py


def _angmod(a):
    """gv_general.py:37-38."""
    return np.arctan2(np.sin(a), np.cos(a))


def _rotate_vectors(direction, vectors, angles):
    """gv_general.py:63-91 (per-vector angles in degrees)."""
    p = np.array(vectors, float)
    q = np.radians(angles)
    rp = p * np.cos(q)
    a_dot_p = np.dot(direction, p)
    apa = np.outer(direction, a_dot_p)
    rp += apa * (1 - np.cos(q))
    rp += np.sin(q) * np.transpose(np.cross(direction, np.transpose(p)))
    return rp


def k_to_g(k, angles, axis=(0, 0, 1), pre=None, post=None):
    """gv_general.py:158-189: pre . rot(axis, angle) . post . k."""
    direction = np.asarray(axis, float)
    pk = np.dot(post, k) if post is not None else k
    rpk = _rotate_vectors(direction, pk, angles)
    return np.dot(pre, rpk) if pre is not None else rpk


def g_to_k(g, wavelength, axis=(0, 0, 1), pre=None, post=None):
    """gv_general.py:192-280: the two diffraction angles (degrees) and
    whether they exist."""
    axis = np.asarray(axis, float)
    rg = np.dot(pre, g) if pre is not None else g
    beam = np.array([-1.0 / wavelength, 0, 0], float)
    rb = np.dot(post.T, beam) if post is not None else beam
    a1 = np.transpose(np.cross(axis, rg.T))
    a2 = np.transpose(np.cross(a1.T, axis))
    a0 = rg - a2
    rbda0 = np.sum(rb[:, np.newaxis] * a0, 0)
    rbda1 = np.sum(rb[:, np.newaxis] * a1, 0)
    rbda2 = np.sum(rb[:, np.newaxis] * a2, 0)
    modg = np.sqrt(np.sum(g * g, 0))
    kdotbeam = -modg * modg / 2.
    phi = np.arctan2(rbda2, rbda1)
    den = np.sqrt(rbda1 * rbda1 + rbda2 * rbda2)
    msk = (den <= 0)
    quot = (kdotbeam - rbda0) / (den + msk)
    valid = (~msk) & (quot >= -1) & (quot <= 1)
    quot = np.where(valid, quot, 0)
    x_plus_p = np.arcsin(quot)
    sol1 = x_plus_p + phi
    sol2 = math.pi - x_plus_p + phi
    return np.degrees(_angmod(sol1)), np.degrees(_angmod(sol2)), valid


def wedgemat(w):
    """gv_general.py:283-289."""
    cw, sw = np.cos(np.radians(w)), np.sin(np.radians(w))
    return np.array([[cw, 0, sw], [0, 1, 0], [-sw, 0, cw]])


def chimat(c):
    """gv_general.py:291-298."""
    cc, sc = np.cos(np.radians(c)), np.sin(np.radians(c))
    return np.array([[1, 0, 0], [0, cc, sc], [0, -sc, cc]])


def wedgechi(wedge=0., chi=0.):
    """gv_general.py:300-305."""
    return np.dot(wedgemat(wedge), chimat(chi))


def chiwedge(chi=0., wedge=0.):
    """gv_general.py:307-312."""
    return np.dot(chimat(chi), wedgemat(wedge))


def uncompute_g_vectors(g, wavelength, wedge=0.0, chi=0.0):
    """transform.py:529-564: tth, [eta1, eta2], [omega1, omega2]."""
    post = None if wedge == chi == 0 else wedgechi(wedge=wedge, chi=chi)
    omega1, omega2, valid = g_to_k(g, wavelength, axis=[0, 0, -1], pre=None, post=post)
    pre = None if post is None else chiwedge(wedge=wedge, chi=chi).T
    k_one = k_to_g(g, omega1, axis=[0, 0, 1], pre=pre, post=None)
    k_two = k_to_g(g, omega2, axis=[0, 0, 1], pre=pre, post=None)
    eta_one = np.arctan2(-k_one[1, :], k_one[2, :])
    eta_two = np.arctan2(-k_two[1, :], k_two[2, :])
    ds = np.sqrt(np.sum(g * g, 0))
    s = ds * wavelength / 2.0
    tth = np.degrees(np.arcsin(s) * 2.0) * valid
    eta1 = np.degrees(eta_one) * valid
    eta2 = np.degrees(eta_two) * valid
    return tth, [eta1, eta2], [omega1 * valid, omega2 * valid]


def uncompute_vectors(wavelength, wedge, chi):
    """The fixed rb = post^T.beam and pre of uncompute_g_vectors (None when
    wedge == chi == 0), as the device kernel takes them."""
    beam = np.array([-1.0 / wavelength, 0, 0], float)
    if wedge == chi == 0:
        return beam, None
    post = wedgechi(wedge=wedge, chi=chi)
    return np.dot(post.T, beam), chiwedge(wedge=wedge, chi=chi).T.copy()


# ---------------------------------------------------------------------------
# lattice symmetry of refinegrains.refine (refinegrains.py:36-79) and the
# xfab.tools conversions it uses


def triclinic(cp):
    return cp


def monoclinic_a(cp):
    a, b, c, al, be, ga = cp
    return [a, b, c, al, 90., 90.]


def monoclinic_b(cp):
    a, b, c, al, be, ga = cp
    return [a, b, c, 90., be, 90.]


def monoclinic_c(cp):
    a, b, c, al, be, ga = cp
    return [a, b, c, 90., 90., ga]


def orthorhombic(cp):
    a, b, c, al, be, ga = cp
    return [a, b, c, 90., 90., 90.]


def tetragonal(cp):
    a, b, c, al, be, ga = cp
    return [(a + b) / 2., (a + b) / 2., c, 90., 90., 90.]


def trigonalP(cp):
    a, b, c, al, be, ga = cp
    anew = (a + b + c) / 3.
    alnew = (al + be + ga) / 3.
    return [anew, anew, anew, alnew, alnew, alnew]


def trigonalH(cp):
    a, b, c, al, be, ga = cp
    anew = (a + b) / 2.
    return [anew, anew, c, 90., 90., 120.]


hexagonal = trigonalH


def cubic(cp):
    anew = (cp[0] + cp[1] + cp[2]) / 3.
    return [anew, anew, anew, 90., 90., 90.]


def ubi_to_cell(ubi):
    """Real-space cell (a, b, c, alpha, beta, gamma in degrees) of a UBI:
    its rows are the direct lattice vectors."""
    a, b, c = (np.sqrt(np.dot(r, r)) for r in ubi)
    ang = lambda u, v: np.degrees(np.arccos(np.dot(u, v) / np.sqrt(np.dot(u, u) * np.dot(v, v))))
    return [a, b, c, ang(ubi[1], ubi[2]), ang(ubi[0], ubi[2]), ang(ubi[0], ubi[1])]


def b_matrix(cell):
    """Reciprocal-space B (upper triangular, no 2 pi): B.h = g in the
    crystal's Cartesian frame, B^T B the reciprocal metric."""
    a, b, c = cell[:3]
    ca, cb, cg = (np.cos(np.radians(x)) for x in cell[3:])
    g = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca],
                  [a * c * cb, b * c * ca, c * c]])
    gi = np.linalg.inv(g)
    # the upper-triangular factor with positive diagonal of the reciprocal
    # metric (gi = L L^T, B = L^T): unique, so any upper-triangular
    # convention (Busing-Levy's included) gives this B up to a scale
    return np.linalg.cholesky(gi).T


def ubi_to_u(ubi):
    """Orientation U with UB = U.B(cell of ubi)."""
    B = b_matrix(ubi_to_cell(ubi))
    return np.dot(np.linalg.inv(ubi), np.linalg.inv(B))


def u_to_ubi(u, cell):
    """UBI = (U.B(cell))^-1."""
    return np.linalg.inv(np.dot(u, b_matrix(cell)))


def symmetrize(mat, latticesymmetry):
    """refinegrains.py:470-474."""
    if latticesymmetry is triclinic:
        return mat
    cp = ubi_to_cell(mat)
    U = ubi_to_u(mat)
    return u_to_ubi(U, latticesymmetry(cp)).copy()


# ---------------------------------------------------------------------------
# simplex.py


class Simplex:
    """Nelder-Mead as ``ImageD11/simplex.py:51-259``: n+3 vertices (the
    last two hold the centroid and the reflected point), the guess list
    mutated in place, the same order of every comparison and update."""

    def __init__(self, testfunc, guess, increments, kR=-1, kE=2, kC=0.5):
        self.testfunc = testfunc
        self.guess = guess
        self.increments = increments
        self.kR, self.kE, self.kC = kR, kE, kC
        n = self.numvars = len(guess)
        self.lowest = self.highest = self.secondhighest = -1
        self.errors = []
        self.currenterror = 0
        self.simplex = [list(guess) for _ in range(n + 3)]
        for v in range(n + 1):
            for x in range(n):
                if x == v - 1:
                    self.simplex[v][x] = guess[x] + increments[x]
            self.errors.append(0)
        self._errors_at_vertices()

    def _errors_at_vertices(self):
        for v in range(self.numvars + 1):
            if v == self.lowest:
                continue
            for x in range(self.numvars):
                self.guess[x] = self.simplex[v][x]
            self.currenterror = self.testfunc(self.guess)
            self.errors[v] = self.currenterror

    def minimize(self, epsilon=0.0001, maxiters=250, monitor=0):
        n = self.numvars
        e, s = self.errors, self.simplex
        it = 0
        for it in range(maxiters):
            self.highest = self.lowest = 0
            for v in range(1, n + 1):
                if e[v] > e[self.highest]:
                    self.highest = v
                if e[v] < e[self.lowest]:
                    self.lowest = v
            self.secondhighest = self.lowest
            for v in range(n + 1):
                if v == self.highest or v == self.secondhighest:
                    continue
                if e[v] > e[self.secondhighest]:
                    self.secondhighest = v
            S = 0.0
            for v in range(n + 1):
                S = S + e[v]
            F2 = S / (n + 1)
            S1 = 0.0
            for v in range(n + 1):
                S1 = S1 + (e[v] - F2) ** 2
            if math.sqrt(S1 / n) <= epsilon:
                break
            h = self.highest
            for x in range(n):
                S = 0.0
                for v in range(n + 1):
                    if v != h:
                        S = S + s[v][x]
                s[n + 1][x] = S / n
            for x in range(n):
                self.guess[x] = self.kR * s[h][x] + (1 - self.kR) * s[n + 1][x]
                s[n + 2][x] = self.guess[x]
            self.currenterror = self.testfunc(self.guess)
            if self.currenterror < e[h]:
                e[h] = self.currenterror
                for x in range(n):
                    s[h][x] = s[n + 2][x]
            if self.currenterror <= e[self.lowest]:
                for x in range(n):
                    self.guess[x] = self.kE * self.guess[x] + (1 - self.kE) * s[n + 1][x]
                self.currenterror = self.testfunc(self.guess)
                if self.currenterror < e[h]:
                    e[h] = self.currenterror
                    for x in range(n):
                        s[h][x] = self.guess[x]
            elif self.currenterror >= e[self.secondhighest]:
                for x in range(n):
                    self.guess[x] = self.kC * s[h][x] + (1 - self.kC) * s[n + 1][x]
                self.currenterror = self.testfunc(self.guess)
                if self.currenterror < e[h]:
                    e[h] = self.currenterror
                    for x in range(n):
                        s[h][x] = self.guess[x]
                else:
                    lo = self.lowest
                    for v in range(n + 1):
                        if v == lo:
                            continue
                        for x in range(n):
                            s[v][x] = 0.5 * (s[v][x] + s[lo][x])
                    self._errors_at_vertices()
        for x in range(n):
            self.guess[x] = s[self.lowest][x]
        self.currenterror = e[self.lowest]
        return self.guess, self.currenterror, it


# ---------------------------------------------------------------------------
# refinegrains.py


class Grain:
    def __init__(self, ubi, translation):
        self.ubi = np.array(ubi, float)
        self.translation = np.array(translation, float)


class RefineGrains:
    """The reference's refinegrains for one scan, grain by grain.

    ``pars``: parameter dict; ``sc``, ``fc``, ``omega``: peak columns;
    ``ubis``/``translations``: the grains read (``readubis``).  ``orc`` is
    the oracle's C library (``oracle.load_oracle()``)."""

    def __init__(self, orc, pars, sc, fc, omega, ubis, translations, tolerance=0.01,
                 latticesymmetry=triclinic, OmFloat=True, OmSlop=0.25):
        self.orc = orc
        self.parameters = dict(pars)
        self.sc, self.fc = np.asarray(sc, float), np.asarray(fc, float)
        self.omega = np.asarray(omega, float)
        self.tolerance = tolerance
        self.latticesymmetry = latticesymmetry
        self.OMEGA_FLOAT = OmFloat
        self.slop = OmSlop
        self.grainnames = list(range(len(ubis)))
        self.ubisread = {i: np.array(u, float) for i, u in enumerate(ubis)}
        self.grains = {i: Grain(ubis[i], translations[i]) for i in self.grainnames}
        self.recompute_xlylzl = False
        self.varylist = []
        self.grains_to_refine = []
        self.evaluations = 0

    def set_translation(self, g):
        """refinegrains.py:630-633."""
        t = self.grains[g].translation
        self.parameters["t_x"], self.parameters["t_y"], self.parameters["t_z"] = t[0], t[1], t[2]

    def assignlabels(self):
        """refinegrains.py:684-800 (the parts the refinement reads)."""
        p = self.parameters
        nr = len(self.sc)
        peaks_xyz = compute_xyz_lab([self.sc, self.fc], **p).T.copy()
        gv = np.zeros((nr, 3), float)
        drlv2 = np.zeros(nr, float) + 1
        labels = np.zeros(nr, np.int32) - 1
        for g in self.grainnames:
            self.set_translation(g)
            self.orc.compute_gv(peaks_xyz, self.omega, p["omegasign"], p["wavelength"],
                                p["wedge"], p["chi"], self.grains[g].translation, gv)
            self.orc.score_and_assign(self.grains[g].ubi, gv, self.tolerance, drlv2, labels,
                                      int(g))
        self.labels, self.drlv2 = labels, drlv2
        for g in self.grainnames:
            gr = self.grains[g]
            gr.ind = np.compress(labels == g, np.arange(nr))
            gr.peaks_xyz = np.take(peaks_xyz, gr.ind, axis=0)
            gr.sc = np.take(self.sc, gr.ind)
            gr.fc = np.take(self.fc, gr.ind)
            gr.om = np.take(self.omega, gr.ind)
            gr.npks = len(gr.ind)
            self.set_translation(g)

    def compute_gv(self, thisgrain):
        """refinegrains.py:366-456 (update_columns off)."""
        p = self.parameters
        sign = p["omegasign"]
        om = thisgrain.om
        tth, eta = compute_tth_eta_from_xyz(thisgrain.peaks_xyz.T, omega=om * sign, **p)
        gv = compute_g_vectors(tth, eta, om * sign, float(p["wavelength"]), p["wedge"], p["chi"])
        self.tth, self.eta = tth, eta
        thisgrain.omega_calc = None
        if self.OMEGA_FLOAT:
            mat = thisgrain.ubi.copy()
            gvT = np.ascontiguousarray(gv.T)
            self.orc.score_and_refine(mat, gvT, self.tolerance)
            hklf = np.dot(mat, gv)
            hkli = np.round(hklf)
            gcalc = np.dot(np.linalg.inv(mat), hkli)
            _, (eta1, eta2), (omega1, omega2) = uncompute_g_vectors(
                gcalc, float(p["wavelength"]), p["wedge"], p["chi"])
            eta_err = np.array([np.abs(eta1 - eta), np.abs(eta2 - eta)])
            best = np.argmin(eta_err, axis=0)
            omega_calc = best * omega2 + (1 - best) * omega1
            omerr = om * sign - omega_calc
            omerr = omerr - (360 * np.round(omerr / 360.0))
            omega_calc = om * sign - np.clip(omerr, -self.slop, self.slop)
            thisgrain.omega_calc = omega_calc
            gv = compute_g_vectors(tth, eta, omega_calc, float(p["wavelength"]), p["wedge"],
                                   p["chi"])
        self.gv = np.ascontiguousarray(gv.T)

    def refine(self, ubi):
        """refinegrains.py:459-498: two score_and_refine calls."""
        mat = ubi.copy()
        self.npks, self.avg_drlv2 = self.orc.score_and_refine(mat, self.gv, self.tolerance)
        mat = symmetrize(mat, self.latticesymmetry)
        self.npks, self.avg_drlv2 = self.orc.score_and_refine(mat, self.gv, self.tolerance)
        return symmetrize(mat, self.latticesymmetry)

    def applyargs(self, args):
        for name, value in zip(self.varylist, args):
            self.parameters[name] = value

    def gof(self, args):
        """refinegrains.py:517-560."""
        self.evaluations += 1
        self.applyargs(args)
        diffs = 0.
        contribs = 0.
        for key in self.grains_to_refine:
            g = self.grains[key]
            if self.recompute_xlylzl:
                g.peaks_xyz = compute_xyz_lab([g.sc, g.fc], **self.parameters).T
            self.compute_gv(g)
            g.ubi = self.refine(self.ubisread[key])
            diffs += self.npks * self.avg_drlv2
            contribs += self.npks
        if contribs > 0:
            return 1e6 * diffs / contribs
        return 1e6

    def estimate_steps(self, guess, steps):
        """refinegrains.py:563-585."""
        cen = self.gof(guess)
        deriv = []
        for i in range(len(steps)):
            newguess = [g for g in guess]
            newguess[i] = newguess[i] + steps[i]
            deriv.append(self.gof(newguess) - cen)
        j = int(np.argmax(np.absolute(deriv)))
        return [s * abs(deriv[j]) / abs(d) for d, s in zip(deriv, steps)]

    def fit(self, stepsizes, maxiters=100):
        """refinegrains.py:588-627."""
        self.assignlabels()
        guess = [self.parameters[n] for n in self.varylist]
        inc = [stepsizes[n] for n in self.varylist]
        self.grains_to_refine = list(self.grainnames)
        self.recompute_xlylzl = any(n not in ("t_x", "t_y", "t_z") for n in self.varylist)
        inc = self.estimate_steps(guess, inc)
        s = Simplex(self.gof, guess, inc)
        newguess, error, it = s.minimize(maxiters=maxiters)
        for name, v in zip(self.varylist, newguess):
            self.parameters[name] = v
        for i, t in enumerate(("t_x", "t_y", "t_z")):
            if t in self.varylist:
                for g in self.grainnames:
                    self.grains[g].translation[i] = newguess[self.varylist.index(t)]
        return newguess, error, it

    def refinepositions(self, stepsizes, maxiters=100):
        """refinegrains.py:636-660: one simplex over (t_x, t_y, t_z) per
        grain, tolerance 1.0, assignments fixed; the translation kept is
        the parameters' value after the last objective call."""
        self.assignlabels()
        tolcache = self.tolerance
        self.tolerance = 1.0
        self.recompute_xlylzl = False
        results = {}
        for key in sorted(self.grainnames):
            self.grains_to_refine = [key]
            self.varylist = ["t_x", "t_y", "t_z"]
            self.set_translation(key)
            guess = [self.parameters[n] for n in self.varylist]
            inc = [stepsizes[n] for n in self.varylist]
            s = Simplex(self.gof, guess, inc)
            newguess, error, it = s.minimize(maxiters=maxiters)
            tr = self.grains[key].translation
            tr[0], tr[1], tr[2] = self.parameters["t_x"], self.parameters["t_y"], self.parameters["t_z"]
            results[key] = (list(newguess), error, it)
        self.tolerance = tolcache
        return results

    def refineubis(self, scoreonly=False):
        """refinegrains.py:663-681: per grain (npks, avg_drlv2)."""
        out = {}
        for key in sorted(self.grainnames):
            g = self.grains[key]
            self.set_translation(key)
            self.compute_gv(g)
            res = self.refine(g.ubi)
            out[key] = (self.npks, self.avg_drlv2)
            if not scoreonly:
                g.ubi = res
        return out
