"""Unit-cell tables for the indexer (host side, numpy).

The parts of ``ImageD11/unitcell.py`` that ``indexer.find`` / ``scorethem``
consume: the reciprocal metric and Busing-Levy B matrix, the powder rings
(``makerings``), the cosine/orientation tables of a ring pair
(``getanglehkls`` = ``cosangles_many`` + ``filter_pairs``) and ``orient``.
The reference's indexer takes its own ``unitcell`` object; an instance of
this class carries the same attributes (``B``, ``gi``, ``ringds``,
``ringhkls``, ``UBI``, ``UBIlist``) and methods, so either can be passed to
:class:`imaged11_amd.indexing.indexer`.

The arithmetic follows the reference expression by expression (same numpy
calls, same evaluation order) because the tables feed exact comparisons:
ring grouping, the sorted cosine table and its tie blocks, and the BT
matrices that fix every trial orientation.  The two calls into the compiled
extension (``quickorient``, ``score``) go to the GPU drop-ins of
:mod:`imaged11_amd.cImageD11`.
"""
import math

import numpy as np

from . import cImageD11


def radians(x):
    """ImageD11/unitcell.py:40-41 (x * pi first, then / 180)."""
    return x * math.pi / 180.0


def degrees(x):
    return x * 180.0 / math.pi


def unit(a):
    """a / sqrt(a . a) (ImageD11/unitcell.py:57-72)."""
    return a / np.sqrt(np.dot(a, a))


def _never(h, k, l):
    return False


def _odd_sum_kl(h, k, l):
    return (k + l) % 2 != 0


def _odd_sum_hl(h, k, l):
    return (h + l) % 2 != 0


def _odd_sum_hk(h, k, l):
    return (h + k) % 2 != 0


def _odd_sum_hkl(h, k, l):
    return (h + k + l) % 2 != 0


def _mixed_parity(h, k, l):
    return (h + k) % 2 != 0 or (h + l) % 2 != 0 or (k + l) % 2 != 0


def _not_obverse(h, k, l):
    return (-h + k + l) % 3 != 0


# Lattice-centring absences by symbol (ImageD11/unitcell.py:77-112).  The
# reference maps "A" to the body-centred rule; that mapping is kept so the
# rings are the reference's.
ABSENT = {"P": _never, "A": _odd_sum_hkl, "B": _odd_sum_hl, "C": _odd_sum_hk,
          "I": _odd_sum_hkl, "F": _mixed_parity, "R": _not_obverse}

# Test reflections for deciding whether two orientations are the same
# (ImageD11/unitcell.py:739-741): columns are hkl.
HKL0 = np.array([[0, 0, 1, 1, -1, 1, -1, 0, 0, 1, -1, 1, 1, 3, 11],
                 [0, 1, 0, 1, 1, 0, 0, 1, -1, 1, 1, -1, 1, 2, 12],
                 [1, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, -1, 1, 13]], float)


def BTmat(h1, h2, B, BI):
    """BI . [u1 u2 u3] for the hkl pair (h1, h2): u1 along B.h1, u3 along
    B.h1 x B.h2, u2 = u1 x u3 (ImageD11/unitcell.py:726-736); quickorient
    turns it into an orientation."""
    g1 = np.dot(B, h1)
    g2 = np.dot(B, h2)
    u1 = unit(g1)
    u3 = unit(np.cross(g1, g2))
    u2 = np.cross(u1, u3)
    return np.dot(BI, np.transpose((u1, u2, u3)))


def cosangles_many(ha, hb, gi):
    """Cosines between every hkl of ha and of hb in the reciprocal metric
    gi, shape (len(ha), len(hb)) (ImageD11/unitcell.py:136-149)."""
    hag = np.dot(ha, gi)
    hbg = np.dot(hb, gi)
    na = np.sqrt((hag * ha).sum(axis=1))
    nb = np.sqrt((hbg * hb).sum(axis=1))
    return np.dot(ha, hbg.T) / np.outer(na, nb)


def filter_pairs(h1, h2, c2a, B, BI, tol=1e-5, ext=cImageD11):
    """Distinct orientations from hkl pairs of two rings
    (ImageD11/unitcell.py:744-810).  Pairs are sorted by cosine and cut
    into blocks where consecutive cosines differ by <= 1e-8; a block whose
    first cosine has |c| >= 0.98 is dropped; within a block the first pair
    is kept and a later pair only when its orientation (built from the
    first pair's ideal g-vectors) fails to index the HKL0 test peaks of
    every pair kept so far.  Returns (pairs, cosines, BT matrices).
    ext provides quickorient and score (the GPU drop-ins by default)."""
    order = np.argsort(c2a.ravel())
    cs = c2a.flat[order]
    hi, hj = np.mgrid[0:len(h1), 0:len(h2)]
    hi = hi.ravel()[order]
    hj = hj.ravel()[order]
    pairs, cangs, matrs = [], [], []
    step = (cs[1:] - cs[:-1]) > 1e-8
    ends = list(np.arange(1, len(step) + 1, dtype=int)[step]) + [len(cs) - 1]
    p = 0
    for e in ends:
        block = cs[p:e]
        if not abs(cs[p]) < 0.98:
            p = e
            continue
        ha, hb = h1[hi[p]], h2[hj[p]]
        bt = BTmat(ha, hb, B, BI)
        pairs.append((ha, hb))
        cangs.append(cs[p])
        matrs.append(bt)
        if len(block) == 1:
            p = e
            continue
        assert (block.max() - block.min()) < 2.1e-8, "cosine block spread in filter_pairs"
        ga, gb = np.dot(B, ha), np.dot(B, hb)
        assert abs(np.dot(ga, gb) / np.sqrt(np.dot(ga, ga) * np.dot(gb, gb)) - cs[p]) < 2e-8
        gobs = np.array((ga, gb, (0, 0, 0)), float)
        u = gobs.copy()
        ext.quickorient(u, bt)
        tests = [np.dot(np.linalg.inv(u), HKL0).T.copy()]
        for q in range(p + 1, e):
            ha, hb = h1[hi[q]], h2[hj[q]]
            bt = BTmat(ha, hb, B, BI)
            new = True
            for gt in tests:
                u = gobs.copy()
                ext.quickorient(u, bt)
                if ext.score(u, gt, 1e-6) == HKL0.shape[1]:
                    new = False
                    break
            if new:
                pairs.append((ha, hb))
                cangs.append(cs[q])
                matrs.append(bt)
                tests.append(np.dot(np.linalg.inv(u), HKL0).T.copy())
        p = e
    return pairs, cangs, matrs


def ubi_equiv(ubilist, ublist, tol=1e-8):
    """Drop orientations that index another's HKL0 peaks exactly
    (ImageD11/unitcell.py:813-832): visited from the largest trace down,
    each kept one contributes its ideal HKL0 g-vectors."""
    if len(ubilist) < 2:
        return ubilist
    order = np.argsort([np.trace(u) for u in ubilist])
    keep = [ubilist[order[-1]]]
    refs = [np.dot(ublist[order[-1]], HKL0)]
    for i in order[:-1][::-1]:
        u = ubilist[i]
        worst = 1
        for pks in refs:
            h = np.dot(u, pks)
            worst = min(worst, np.abs(np.round(h) - h).sum())
            if worst <= tol:
                break
        if worst > tol:
            keep.append(u)
            refs.append(np.dot(np.linalg.inv(u), HKL0))
    return keep


class unitcell:
    """Lattice a, b, c, alpha, beta, gamma (degrees) with a centring symbol
    (ImageD11/unitcell.py:192-283).  Integer space groups need xfab's
    reflection generator, which this package does not carry.  ext: the
    module providing quickorient and score (default: the GPU drop-ins of
    imaged11_amd.cImageD11)."""

    def __init__(self, lattice_parameters, symmetry="P", verbose=0, name=None, ext=None):
        self.ext = cImageD11 if ext is None else ext
        self.lattice_parameters = np.array(lattice_parameters)
        if self.lattice_parameters.shape[0] != 6:
            raise ValueError("You must supply 6 lattice parameters a,b,c,alpha,beta,gamma")
        if isinstance(symmetry, str) and symmetry.isdigit():
            symmetry = int(symmetry)
        if isinstance(symmetry, int):
            raise NotImplementedError("space-group symmetry needs xfab (not available)")
        if symmetry not in ABSENT:
            raise ValueError("Your symmetry %s was not recognised" % symmetry)
        self.symmetry = symmetry
        self.absent = ABSENT[symmetry]
        if name is not None:
            self.name = name
        a, b, c = self.lattice_parameters[:3]
        self.alpha = radians(self.lattice_parameters[3])
        ca = math.cos(radians(self.lattice_parameters[3]))
        cb = math.cos(radians(self.lattice_parameters[4]))
        cg = math.cos(radians(self.lattice_parameters[5]))
        # real-space metric and its inverse, the reciprocal metric
        self.g = np.array([[a * a, a * b * cg, a * c * cb],
                           [a * b * cg, b * b, b * c * ca],
                           [a * c * cb, b * c * ca, c * c]], float)
        self.gi = np.linalg.inv(self.g)
        self.astar = np.sqrt(self.gi[0, 0])
        self.bstar = np.sqrt(self.gi[1, 1])
        self.cstar = np.sqrt(self.gi[2, 2])
        self.alphas = degrees(math.acos(self.gi[1, 2] / self.bstar / self.cstar))
        self.betas = degrees(math.acos(self.gi[0, 2] / self.astar / self.cstar))
        self.gammas = degrees(math.acos(self.gi[0, 1] / self.astar / self.bstar))
        # Busing & Levy (1967) eq. 3
        self.B = np.array(
            [[self.astar, self.bstar * math.cos(radians(self.gammas)),
              self.cstar * math.cos(radians(self.betas))],
             [0, self.bstar * math.sin(radians(self.gammas)),
              -self.cstar * math.sin(radians(self.betas)) * ca],
             [0, 0, 1.0 / c]], float)
        self.hkls = None
        self.peaks = None
        self.limit = 0
        self.ringtol = 0.001
        self.ringds = []
        self.ringhkls = {}
        self.anglehkl_cache = {"ringtol": self.ringtol, "B": self.B, "BI": np.linalg.inv(self.B)}
        self.UBI = None
        self.UB = None
        self.UBIlist = []

    def tostring(self):
        return ("%f " * 6) % tuple(self.lattice_parameters) + str(self.symmetry)

    def ds(self, h):
        """1/d of reflection h: sqrt(h . gi . h)."""
        return math.sqrt(np.dot(h, np.dot(self.gi, h)))

    def gethkls(self, dsmax):
        """[ds, (h, k, l)] for every allowed reflection with ds < dsmax,
        sorted (ImageD11/unitcell.py:461-523).

        The reference walks outwards from the origin along l, then k, then
        h, turning back or moving to the next line when ds first reaches
        dsmax; the set it returns is the one that walk visits, so the same
        walk is made here (|h|, |k|, |l| < 200)."""
        if dsmax == self.limit and self.peaks is not None:
            return self.peaks
        found = []
        h, k, l = 0, 0, 1  # (0, 0, 0) is never listed
        sh = sk = sl = 1
        misses = 0
        while abs(h) < 200:
            while abs(k) < 200:
                while abs(l) < 200:
                    d = self.ds([h, k, l])
                    if d < dsmax:
                        if not self.absent(h, k, l):
                            found.append([d, (h, k, l)])
                        misses = 0
                    elif sl == 1:  # end of the +l half-line: walk -l from 0
                        sl, l = -1, 0
                    else:  # end of the line: next k
                        sl, l = 1, 0
                        misses += 1
                        break
                    l += sl
                k += sk
                if misses > 1:  # two empty lines: turn k round, or next h
                    if sk == 1:
                        sk, k = -1, -1
                    else:
                        sk, k = 1, 0
                        misses += 1
                        break
            h += sh
            if misses > 3:  # empty planes: turn h round, or stop
                if sh == 1:
                    sh, h = -1, -1
                else:
                    break
        found.sort()
        self.peaks = found
        self.limit = dsmax
        return found

    def makerings(self, limit, tol=0.001):
        """Powder rings: reflections sorted by ds, grouped while each
        differs from the ring's first ds by < tol
        (ImageD11/unitcell.py:529-548)."""
        self.peaks = self.gethkls(limit + tol)
        self.ringds = []
        self.ringhkls = {}
        first = self.peaks[0]
        self.ringds.append(first[0])
        self.ringhkls[first[0]] = [first[1]]
        for d, hkl in self.peaks[1:]:
            if abs(d - self.ringds[-1]) < tol:
                self.ringhkls[self.ringds[-1]].append(hkl)
            else:
                self.ringds.append(d)
                self.ringhkls[d] = [hkl]
        self.ringtol = tol

    def ringtths(self, wavelength):
        if len(self.ringds) == 0:
            raise ValueError("No rings found. Call `makerings` before calling this function.")
        return [2 * np.degrees(np.arcsin(d * wavelength / 2)) for d in self.ringds]

    def anglehkls(self, h1, h2):
        """(angle in degrees, cosine) between reciprocal vectors h1, h2
        (ImageD11/unitcell.py:563-580)."""
        g1 = np.dot(h1, np.dot(self.gi, h1))
        g2 = np.dot(h2, np.dot(self.gi, h2))
        g12 = np.dot(h1, np.dot(self.gi, h2))
        c = g12 / math.sqrt(g1 * g2)
        try:
            return degrees(math.acos(c)), c
        except ValueError:
            if abs(c - 1) < 1e-6:
                return 0.0, 1.0
            if abs(c + 1) < 1e-6:
                return 180.0, -1.0
            raise

    def getanglehkls(self, ring1, ring2):
        """(pairs, sorted cosines, BT matrices) for two rings, cached per
        (ring1, ring2) while B and the ring tolerance stay the same
        (ImageD11/unitcell.py:582-603)."""
        if self.ringtol != self.anglehkl_cache["ringtol"] or \
                (self.B != self.anglehkl_cache["B"]).any():
            self.anglehkl_cache = {"ringtol": self.ringtol, "B": self.B,
                                   "BI": np.linalg.inv(self.B)}
        key = (ring1, ring2)
        if key not in self.anglehkl_cache:
            h1 = self.ringhkls[self.ringds[ring1]]
            h2 = self.ringhkls[self.ringds[ring2]]
            c = cosangles_many(h1, h2, self.gi)
            self.anglehkl_cache[key] = filter_pairs(h1, h2, c, self.anglehkl_cache["B"],
                                                    self.anglehkl_cache["BI"], ext=self.ext)
        return self.anglehkl_cache[key]

    def orient(self, ring1, g1, ring2, g2, verbose=0, crange=-1.0):
        """UBI from two observed g-vectors assigned to rings ring1, ring2
        (ImageD11/unitcell.py:605-666).  crange <= 0: the one table entry
        nearest the observed cosine; crange > 0: every entry within crange,
        reduced to distinct orientations.  Sets UBI, UB and UBIlist."""
        ct = np.dot(g1, g2) / np.sqrt((g1 * g1).sum() * (g2 * g2).sum())
        hab, c2ab, matrs = self.getanglehkls(ring1, ring2)
        if crange > 0:
            best = np.arange(len(c2ab), dtype=int)[abs(c2ab - ct) < crange]
        else:
            i = np.searchsorted(c2ab, ct, side="left")
            if i > 0 and (i == len(c2ab) or abs(ct - c2ab[i - 1]) < abs(ct - c2ab[i])):
                best = [i - 1]
            else:
                best = [i]
        self.UBIlist = []
        ubs = []
        for b in best:
            u = np.empty((3, 3), float)
            u[0] = g1
            u[1] = g2
            u[2] = 0.0
            self.ext.quickorient(u, matrs[b])
            self.UBI = u
            self.UB = np.linalg.inv(u)
            self.UBIlist.append(u)
            ubs.append(self.UB)
        self.UBIlist = ubi_equiv(self.UBIlist, ubs)


def cellfromstring(s, ext=None):
    """'a b c alpha beta gamma [symmetry]' (ImageD11/unitcell.py:152-160)."""
    items = s.split()
    latt = [float(x) for x in items[0:6]]
    symm = items[6] if len(items) > 6 else "P"
    return unitcell(latt, symm, ext=ext)


def unitcell_from_parameters(pars):
    """The unit cell named by a parameters object or dict: cell__a,
    cell__b, cell__c, cell_alpha, cell_beta, cell_gamma and the lattice
    symbol cell_lattice_[P,A,B,C,I,F,R] (ImageD11/unitcell.py:835-840)."""
    get = pars.get
    names = ("_a", "_b", "_c", "alpha", "beta", "gamma")
    return unitcell([float(get("cell_%s" % n)) for n in names],
                    get("cell_lattice_[P,A,B,C,I,F,R]"))
