"""The refinegrains driver (``ImageD11/refinegrains.py``) with its
per-grain loops batched on the GPU.

The reference refines grain orientations, grain positions and detector
parameters with Python loops over grains: ``gof`` (:517-560) calls
``compute_gv`` (:366-456, numpy geometry plus the omega-float step) and
``refine`` (:459-498, two ``score_and_refine`` calls) once per grain per
objective evaluation, ``fit`` (:588-627) minimises ``gof`` with a
Nelder-Mead simplex, and ``refinepositions`` (:636-660) runs one simplex
over (t_x, t_y, t_z) per grain, one grain after the other.

Here one objective evaluation for any set of (grain, translation) requests
is a handful of launches: ``gs_grain_gv`` (the per-grain geometry of every
list entry, including the omega-float step), ``gs_refine_lists`` (the
score_and_refine calls of every grain), and one host copy of the per-grain
counts and means.  ``refinepositions`` steps the simplexes of all grains
together (``simplex.SimplexBatch``): each round evaluates one point for
every grain still running, in the order each grain's own simplex asks for
it, so every grain follows the trajectory the reference's sequential loop
gives it.

Semantics kept from the reference, deliberately:
* ``gof`` computes every grain's g-vectors with the translation held in the
  parameters (t_x, t_y, t_z), not the grains' own (the reference passes the
  whole parameter dict to ``compute_tth_eta_from_xyz``);
* every objective call overwrites the grain's UBI with its refined one, and
  the omega-float step starts from that UBI;
* ``refinepositions`` keeps, per grain, the translation of the LAST point
  the simplex evaluated (the reference reads the parameters after
  ``minimize``), and runs with tolerance 1.0;
* labels and assignments come from ``assignlabels`` and stay fixed.
Lattice symmetry is applied between and after the two score_and_refine
calls with ``ubi_to_cell`` / ``ubi_to_u`` / ``u_to_ubi`` (the xfab.tools
functions of the same names, restated: U from UB = U.B with B the
upper-triangular factor of the reciprocal metric).
"""
import math
import re

import numpy as np
import torch

from . import _lib, batched, columnfile, geometry, grain, parameters, simplex

# ---------------------------------------------------------------------------
# lattice symmetries (refinegrains.py:36-79): cell -> constrained cell


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


LATTICES = {f.__name__: f for f in (triclinic, monoclinic_a, monoclinic_b, monoclinic_c,
                                    orthorhombic, tetragonal, trigonalP, trigonalH, cubic)}
LATTICES["hexagonal"] = hexagonal


def ubi_to_cell(ubi):
    """(..., 3, 3) UBIs -> (..., 6) cells (a, b, c, alpha, beta, gamma in
    degrees): the rows of a UBI are the direct lattice vectors."""
    u = np.asarray(ubi, float)
    n = np.sqrt(np.sum(u * u, -1))
    ang = lambda i, j: np.degrees(np.arccos(np.sum(u[..., i, :] * u[..., j, :], -1) /
                                            (n[..., i] * n[..., j])))
    return np.stack([n[..., 0], n[..., 1], n[..., 2], ang(1, 2), ang(0, 2), ang(0, 1)], -1)


def b_matrix(cells):
    """(..., 6) cells -> (..., 3, 3) upper-triangular B with B^T B the
    reciprocal metric (B.h is g in the crystal's Cartesian frame)."""
    c = np.asarray(cells, float)
    a, b, cc = c[..., 0], c[..., 1], c[..., 2]
    ca, cb, cg = (np.cos(np.radians(c[..., i])) for i in (3, 4, 5))
    g = np.stack([np.stack([a * a, a * b * cg, a * cc * cb], -1),
                  np.stack([a * b * cg, b * b, b * cc * ca], -1),
                  np.stack([a * cc * cb, b * cc * ca, cc * cc], -1)], -2)
    return np.swapaxes(np.linalg.cholesky(np.linalg.inv(g)), -1, -2)


def ubi_to_u(ubi):
    """Orientation U with inv(ubi) = U.B(cell of ubi)."""
    u = np.asarray(ubi, float)
    return np.matmul(np.linalg.inv(u), np.linalg.inv(b_matrix(ubi_to_cell(u))))


def u_to_ubi(u, cells):
    """UBI = inv(U.B(cell))."""
    return np.linalg.inv(np.matmul(u, b_matrix(cells)))


# ---- symmetry-equivalent orientation choice (ImageD11/sym_u.py) -------------
# The rotation groups makeuniq chooses within, as generator operators in the
# "x,y,z" notation of ImageD11/sym_u.py:175-225 (cubic P432, hexagonal
# P6/mmm, trigonal P321, rhombohedral R3, tetragonal P4, orthorhombic P222,
# monoclinic with the 2-fold along c, a or b, triclinic).  The generators
# fix the order in which the group's elements are generated, which decides
# between equal traces.
GROUP_GENERATORS = {
    "cubic": ("z,x,y", "-y,x,z"),
    "hexagonal": ("-y,x-y,z", "-x,-y,z", "y,x,-z"),
    "trigonal": ("y,-x-y,z", "y,x,-z"),
    "rhombohedralP": ("z,x,y", "-z,-y,-x"),
    "tetragonal": ("-y,x,z", "-x,y,-z"),
    "orthorhombic": ("-x,-y,z", "-x,y,-z"),
    "monoclinic_c": ("-x,-y,z",),
    "monoclinic_a": ("x,-y,-z",),
    "monoclinic_b": ("-x,y,-z",),
    "triclinic": ("x,y,z",),
}
GROUP_GENERATORS["trigonalP"] = GROUP_GENERATORS["rhombohedralP"]


def symmetry_operator(spec):
    """The 3x3 matrix of an operator such as "-y,x-y,z": row i is the image
    of the i-th unit vector (the convention of sym_u.m_from_string,
    ImageD11/sym_u.py:13-22).  Parsed, not evaluated: each component is a
    signed sum of x, y and z."""
    parts = [p.replace(" ", "") for p in spec.split(",")]
    if len(parts) != 3:
        raise ValueError("operator %r needs three components" % spec)
    m = np.zeros((3, 3))
    for j, part in enumerate(parts):
        pos = 0
        for sign, var in re.findall(r"([+-]?)([xyz])", part):
            m["xyz".index(var), j] += -1.0 if sign == "-" else 1.0
            pos += len(sign) + 1
        if pos != len(part):
            raise ValueError("cannot read component %r of %r" % (part, spec))
    return m


_GROUPS = {}


def point_group(name):
    """The elements of a rotation group, in the order the reference's
    group.additem / makegroup build them (ImageD11/sym_u.py:64-116): the
    identity, then each generator not yet present followed by the closure
    (every product a.b of the list so far, in list order, appended when new;
    equality within 1e-5)."""
    if name not in GROUP_GENERATORS:
        raise ValueError("%s not in %s" % (name, sorted(GROUP_GENERATORS)))
    if name in _GROUPS:
        return _GROUPS[name]
    elems = [np.identity(3)]
    member = lambda m: any(np.allclose(e, m, rtol=1e-5, atol=1e-5) for e in elems)
    for spec in GROUP_GENERATORS[name]:
        m = symmetry_operator(spec)
        if not member(m):
            elems.append(m)
        new = True
        while new:
            i = 0
            while i < len(elems):
                j = 0
                while j < len(elems):
                    c = np.dot(elems[i], elems[j])
                    new = not member(c)
                    if new:
                        elems.append(c)
                    j += 1
                i += 1
    _GROUPS[name] = elems
    return elems


def find_uniq_u(u, ops):
    """Among o.u for the group's operators o, the one with the largest
    trace, keeping u itself unless some o.u is strictly larger and the
    first of equal ones (ImageD11/sym_u.py:229-239)."""
    best = np.asarray(u, float)
    tmax = np.trace(best)
    for o in ops:
        cand = np.dot(o, u)
        t = np.trace(cand)
        if t > tmax:
            best, tmax = cand, t
    return np.array(best)


def symmetrize(ubis, latticesymmetry):
    """The projection refinegrains.refine applies (refinegrains.py:470-480)
    to each of the (K, 3, 3) UBIs."""
    if latticesymmetry is triclinic:
        return ubis
    u = np.asarray(ubis, float).reshape(-1, 3, 3)
    cells = ubi_to_cell(u)
    U = ubi_to_u(u)
    new = np.array([latticesymmetry(list(cp)) for cp in cells], float)
    return u_to_ubi(U, new).reshape(np.shape(ubis))


# ---------------------------------------------------------------------------
# intensity bookkeeping of assignlabels (refinegrains.py:823-995), host-side


def lf(tth, eta):
    """Lorentz factor sin(2 theta) |sin(eta)| (refinegrains.py:872-897)."""
    return np.sin(np.radians(tth)) * np.abs(np.sin(np.radians(eta)))


def compute_lp_factor(colfile, **kwds):
    """Lorentz and Lorentz_per_grain columns (refinegrains.py:823-869)."""
    if "tth" in colfile.titles and "eta" in colfile.titles:
        colfile.addcolumn(lf(colfile.tth, colfile.eta), "Lorentz")
    if "tth_per_grain" in colfile.titles and "eta_per_grain" in colfile.titles:
        colfile.addcolumn(lf(colfile.tth_per_grain, colfile.eta_per_grain), "Lorentz_per_grain")


def _weighted_intensity(colfile):
    t = colfile.titles
    if "sum_intensity" in t:
        raw = colfile.sum_intensity
    elif "Number_of_pixels" in t and "avg_intensity" in t:
        raw = colfile.Number_of_pixels * colfile.avg_intensity
    elif "npixels" in t and "avg_intensity" in t:
        raw = colfile.npixels * colfile.avg_intensity
    else:
        raw = np.ones(colfile.nrows)
    if "Lorentz_per_grain" in t:
        lor = colfile.Lorentz_per_grain
    elif "Lorentz" in t:
        lor = colfile.Lorentz
    else:
        lor = np.ones(colfile.nrows, np.float32)
    return raw * lor


def _tth_column(colfile):
    t = colfile.titles
    if "tth_per_grain" in t:
        return colfile.tth_per_grain
    if "tth" in t:
        return colfile.tth
    return np.ones(colfile.nrows)


def _intensity_summary(weighted, tth, indices, tth_range, ntrim, quiet):
    return _summary_of(np.take(weighted, indices), np.take(tth, indices), tth_range, ntrim,
                       quiet)


def _summary_of(intensities, tv, tth_range, ntrim, quiet):
    """The summary string of one grain's intensities (weighted, in peak
    order) and their tth values."""
    sigma_i = np.sum(intensities)
    if tth_range is not None:
        intensities = np.compress((tv > min(tth_range)) & (tv < max(tth_range)), intensities)
    if len(intensities) < 1:
        return "no peaks"
    intensities = np.sort(intensities)
    if len(intensities) > ntrim * 2 + 1:
        intensities = intensities[ntrim:-ntrim]
    ret = ("sum_of_all = %f , middle %d from %f to %f in tth: median = %f , min = %f , "
           "max = %f , mean = %f , std = %f , n = %d" % (
               sigma_i, len(intensities), min(tth_range), max(tth_range),
               intensities[len(intensities) // 2], intensities.min(), intensities.max(),
               intensities.mean(), intensities.std(), intensities.shape[0]))
    if not quiet:
        print(ret)
    return ret


def compute_total_intensity(colfile, indices, tth_range, ntrim=2, quiet=False):
    """The summary string of a grain's peak intensities (intensity times the
    Lorentz factor, trimmed, within tth_range) (refinegrains.py:919-995)."""
    return _intensity_summary(_weighted_intensity(colfile), _tth_column(colfile), indices,
                              tth_range, ntrim, quiet)


# ---------------------------------------------------------------------------


class _ScanState:
    """Device-resident peaks of one scan and the per-grain lists that
    assignlabels fixed."""

    def __init__(self, col, device, parity=False):
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
        self.device = device
        self.parity = parity
        self.sc, self.fc, self.omega = T(col.sc), T(col.fc), T(col.omega)
        self.n = int(col.nrows)
        self.xl = None
        self.sign = None
        self.so = self.co = None
        self.idx = self.off = None
        self.off_host = None
        self.requests = None  # (positions, lists) of the last _requests call
        # per grain position (assignlabels' order): the UBIs read and the
        # current ones, on the device
        self.names = []
        self.ubi_read = self.ubi_cur = None

    def lab(self, pars):
        """Lab positions for the current detector parameters."""
        p, r, dist = geometry.pack_xlylzl(pars)
        self.xl = batched.xlylzl(self.sc, self.fc, p, r, dist)

    def trig(self, omegasign):
        if self.sign != float(omegasign):
            self.so, self.co = batched.sincos_omega(self.omega, float(omegasign),
                                                    parity=self.parity)
            self.sign = float(omegasign)


class _RoundGraph:
    """One round of the lockstep position simplex for every grain of a scan
    at once, captured as a hipGraph: compute_gv (with the omega-float step
    when on) and two score_and_refine iterations for all grains at the
    translations in ``trans``, then the refined UBIs stored for the grains
    whose ``mask`` entry is set (the others are computed and dropped: a
    grain's result depends on its own request only, so the selected grains
    get the bits the eager round gives them).  ``out`` holds (npks, mean
    drlv2) of every grain.  Triclinic lattices only (the other symmetries
    project the UBIs on the host between the two iterations)."""

    def __init__(self, rg, sname, pos):
        st = rg._state[sname]
        dev = st.device
        self.rg, self.sname = rg, sname
        self.pos = np.asarray(pos, np.int64)
        R = len(self.pos)
        self.pos_d = torch.as_tensor(self.pos, device=dev)
        self.trans = torch.zeros((R, 3), dtype=torch.float64, device=dev)
        self.mask = torch.zeros(R, dtype=torch.bool, device=dev)
        self.trans_h = torch.zeros((R, 3), dtype=torch.float64, pin_memory=True)
        self.mask_h = torch.zeros(R, dtype=torch.bool, pin_memory=True)
        self.out_h = torch.zeros((2, R), dtype=torch.float64, pin_memory=True)
        # every lazily built input exists before capture: the lists of
        # these positions, their order, sin/cos of omega
        st.trig(rg.parameterobj.parameters["omegasign"])
        _, _, _, off_d = rg._requests(sname, self.pos)
        rg._order_of(sname, off_d)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):  # one eager pass (mask clear: no update)
            self._body()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self._body()

    def _body(self):
        rg, st = self.rg, self.rg._state[self.sname]
        idx, owner, off, off_d = rg._requests(self.sname, self.pos)
        gv, _, _, _ = rg._gv_device(self.sname, idx, owner, off_d, self.trans, None, False,
                                    ubis=st.ubi_cur[self.pos_d])
        u = st.ubi_read[self.pos_d].contiguous()
        u, n, m = batched.refine_lists(u, gv, None, off_d, rg.tolerance, 2,
                                       order=rg._order_of(self.sname, off_d))
        st.ubi_cur[self.pos_d] = torch.where(self.mask[:, None, None], u,
                                             st.ubi_cur[self.pos_d])
        return torch.stack([n[1].double(), m[1]])

    def __call__(self, trans, which):
        """(npks, mean drlv2) of the grains `which` (indices into pos) at
        their rows of trans (R,3); their UBIs updated as the eager round
        updates them."""
        self.trans_h.numpy()[:] = trans
        mh = self.mask_h.numpy()
        mh[:] = False
        mh[which] = True
        self.trans.copy_(self.trans_h, non_blocking=True)
        self.mask.copy_(self.mask_h, non_blocking=True)
        self.graph.replay()
        self.out_h.copy_(self.out, non_blocking=True)
        torch.cuda.current_stream(self.trans.device).synchronize()
        nm = self.out_h.numpy()
        return nm[0][which].astype(np.int64), nm[1][which].copy()


class refinegrains:
    """Refine grain orientations, positions and detector parameters
    (``ImageD11/refinegrains.py:137-820``), batched over grains on the GPU.
    Method names, arguments, defaults and side effects follow the
    reference's."""

    # refinegrains.py:143-169
    pars = {"cell__a": 4.1569162, "cell__b": 4.1569162, "cell__c": 4.1569162,
            "cell_alpha": 90.0, "cell_beta": 90.0, "cell_gamma": 90.0,
            "cell_lattice_[P,A,B,C,I,F,R]": 'P', "chi": 0.0, "distance": 7367.8452302,
            "fit_tolerance": 0.5, "o11": 1, "o12": 0, "o21": 0, "o22": 1, "omegasign": 1,
            "t_x": 33.0198146824, "t_y": 14.6384893741, "t_z": 0.0, "tilt_x": 0.0,
            "tilt_y": 0.0623952920101, "tilt_z": 0.00995011461696, "wavelength": 0.2646,
            "wedge": 0.0, "y_center": 732.950204632, "y_size": 4.6,
            "z_center": 517.007049626, "z_size": 4.6}

    # refinegrains.py:172-187
    stepsizes = {"wavelength": 0.001, "y_center": 0.2, "z_center": 0.2, "distance": 200.,
                 "tilt_y": math.radians(0.1), "tilt_z": math.radians(0.1),
                 "tilt_x": math.radians(0.1), "wedge": math.radians(0.1),
                 "chi": math.radians(0.1), "t_x": 0.2, "t_y": 0.2, "t_z": 0.2,
                 "y_size": 0.2, "z_size": 0.2}

    def __init__(self, tolerance=0.01, intensity_tth_range=(6.1, 6.3),
                 latticesymmetry=triclinic, OmFloat=True, OmSlop=0.25, device=None,
                 parity=False):
        _lib.load()
        _lib.require_gpu()
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.OMEGA_FLOAT = OmFloat
        # parity=True: sin/cos of omega from the host's glibc (the values
        # the reference's compute_gv uses, bit for bit); else the device's
        self.parity = bool(parity)
        self.slop = OmSlop
        self.tolerance = tolerance
        self.grainnames = []
        self.ubisread = {}
        self.translationsread = {}
        self.scannames = []
        self.scantitles = {}
        self.scandata = {}
        self.grains = {}
        self.grains_to_refine = []
        self.latticesymmetry = latticesymmetry
        self.drlv = None
        self.parameterobj = parameters.parameters(**self.pars)
        self.intensity_tth_range = intensity_tth_range
        self.recompute_xlylzl = False
        for k, s in self.stepsizes.items():
            self.parameterobj.stepsizes[k] = s
        self._state = {}
        self._pos = {}
        self.npks = 0
        self.avg_drlv2 = 0.0
        # refinepositions replays one captured evaluation per simplex round
        # (triclinic lattices); False runs each round's launches eagerly
        self.round_graphs = True

    # -- files -------------------------------------------------------------
    def loadparameters(self, filename):
        self.parameterobj.loadparameters(filename)

    def saveparameters(self, filename):
        self.parameterobj.saveparameters(filename)

    def readubis(self, filename):
        """Grains of a .ubi file, named 0, 1, ... (refinegrains.py:224-240)."""
        for i, g in enumerate(grain.read_grain_file(filename)):
            self.grainnames.append(i)
            self.ubisread[i] = g.ubi
            self.translationsread[i] = g.translation

    def loadfiltered(self, filename):
        """A peak file, with drlv2, labels, sc and fc columns added when
        absent (refinegrains.py:297-316)."""
        col = filename if isinstance(filename, columnfile.columnfile) else \
            columnfile.columnfile(filename)
        name = filename if isinstance(filename, str) else "scan%d" % len(self.scannames)
        self.scannames.append(name)
        self.scantitles[name] = col.titles
        if "drlv2" not in col.titles:
            col.addcolumn(np.ones(col.nrows, float), "drlv2")
        if "labels" not in col.titles:
            col.addcolumn(np.ones(col.nrows, float) - 2, "labels")
        if "sc" not in col.titles:
            col.addcolumn(col.xc.copy(), "sc")
        if "fc" not in col.titles:
            col.addcolumn(col.yc.copy(), "fc")
        self.scandata[name] = col
        self._state[name] = _ScanState(col, self.device, self.parity)
        return name

    def generate_grains(self):
        """One grain per (grain, scan) (refinegrains.py:319-341)."""
        t = np.array([self.parameterobj.parameters[s] for s in ("t_x", "t_y", "t_z")], float)
        for gname in self.grainnames:
            for sname in self.scannames:
                if (gname, sname) in self.grains:
                    continue
                tr = self.translationsread[gname]
                g = grain.grain(self.ubisread[gname], translation=t if tr is None else tr)
                g.name = (str(gname) + ":" + sname).replace(" ", "_")
                self.grains[(gname, sname)] = g
        for sname in self.scannames:
            self.reset_labels(sname)

    def reset_labels(self, scanname):
        col = self.scandata[scanname]
        x = col.xc if "xc" in col.titles else col.sc
        y = col.yc if "yc" in col.titles else col.fc
        col.labels = col.labels * 0 - 2
        col.drlv2 = col.drlv2 * 0 + 1
        for g in self.grainnames:
            gr = self.grains[(g, scanname)]
            gr.x, gr.y, gr.om = x, y, col.omega

    def savegrains(self, filename, sort_npks=True):
        """Write the grains, after filling the per-grain columns (tth, eta,
        omega_calc, g, hkl) of their peaks (refinegrains.py:242-283)."""
        ks = list(self.grains.keys())
        if sort_npks:
            order = np.argsort([self.grains[k].npks for k in ks])
            ks = [ks[i] for i in order[::-1]]
        else:
            ks.sort()
        if len(self.scandata) != 1:
            raise ValueError("savegrains handles one scan (as the reference)")
        sname = self.scannames[0]
        col = self.scandata[sname]
        res = self._geometry(sname, [k[0] for k in ks],
                             np.array([self.grains[k].translation for k in ks], float),
                             angles=True)
        gv, tth, eta, omc, off = res
        for j, k in enumerate(ks):
            g = self.grains[k]
            sl = slice(off[j], off[j + 1])
            np.put(col.tth_per_grain, g.ind, tth[sl])
            np.put(col.eta_per_grain, g.ind, eta[sl])
            if self.OMEGA_FLOAT:
                np.put(col.omegacalc_per_grain, g.ind, omc[sl])
            gvg = gv[sl]
            for c, name in enumerate(("gx", "gy", "gz")):
                np.put(col[name], g.ind, gvg[:, c])
            hkl_real = np.dot(g.ubi, gvg.T)
            for c, name in enumerate(("hr", "kr", "lr")):
                np.put(col[name], g.ind, hkl_real[c, :])
            hkl = np.floor(hkl_real + 0.5)
            for c, name in enumerate(("h", "k", "l")):
                np.put(col[name], g.ind, hkl[c, :])
            sign_eta = np.sign(col.eta_per_grain[g.ind])
            g.nuniq = len(set((int(h), int(kk), int(ll), int(s))
                              for (h, kk, ll), s in zip(hkl.T, sign_eta)))
        self.set_translation(ks[-1][0], sname)
        grain.write_grain_file(filename, [self.grains[k] for k in ks])

    # -- parameters --------------------------------------------------------
    def applyargs(self, args):
        self.parameterobj.set_variable_values(args)

    def printresult(self, arg):
        for item, value in zip(self.parameterobj.varylist, arg):
            self.parameterobj.parameters[item] = value
            print(item, value)

    def getgrains(self):
        return list(self.grains.keys())

    def makeuniq(self, symmetry):
        """Replace every UBI read and every grain's UBI by its equivalent
        under the rotation group `symmetry` (a name of GROUP_GENERATORS)
        with the largest trace (ImageD11/refinegrains.py:288-300).  The
        device copies of the UBIs read are refreshed; the grains' own UBIs
        reach the device at the next evaluation."""
        ops = point_group(symmetry)
        for k in list(self.ubisread.keys()):
            self.ubisread[k] = find_uniq_u(self.ubisread[k], ops)
        for k in list(self.grains.keys()):
            self.grains[k].set_ubi(find_uniq_u(self.grains[k].ubi, ops))
        for st in self._state.values():
            if st.ubi_read is not None and st.names:
                st.ubi_read = torch.as_tensor(np.array([self.ubisread[n] for n in st.names],
                                                       float), device=st.device)

    def set_translation(self, gr, sc):
        t = self.grains[(gr, sc)].translation
        p = self.parameterobj.parameters
        p["t_x"], p["t_y"], p["t_z"] = t[0], t[1], t[2]

    # -- the batched core --------------------------------------------------
    def _requests(self, sname, positions):
        """CSR entries of the requests: (idx, owner, off (R+1) host and
        device) for grain positions ``positions`` (one request each).  The
        lists of the last set of positions are kept (the lockstep simplex
        asks for the same set round after round) until assignlabels
        rebuilds the grains' lists."""
        st = self._state[sname]
        pos = np.asarray(positions, np.int64)
        if st.requests is not None and np.array_equal(st.requests[0], pos):
            return st.requests[1]
        lengths = st.off_host[pos + 1] - st.off_host[pos]
        off = np.zeros(len(pos) + 1, np.int64)
        off[1:] = np.cumsum(lengths)
        M = int(off[-1])
        dev = st.device
        lengths_d = torch.as_tensor(lengths, device=dev)
        owner = torch.repeat_interleave(torch.arange(len(pos), device=dev), lengths_d,
                                        output_size=M)
        starts = torch.as_tensor(st.off_host[pos], device=dev)
        off_d = torch.as_tensor(off, device=dev)
        within = torch.arange(M, device=dev) - off_d[owner]
        idx = st.idx[(starts[owner] + within)].contiguous()
        out = (idx, owner.to(torch.int32), off, off_d)
        st.requests = (pos.copy(), out, batched.list_order(off_d))
        return out

    def _order_of(self, sname, off_d):
        """The refinement's grain order for lists off_d (kept with the
        lists of the last request set)."""
        st = self._state[sname]
        if st.requests is not None and st.requests[1][3] is off_d:
            return st.requests[2]
        return batched.list_order(off_d)

    def _geometry(self, sname, names, trans, angles=False, ubis=None):
        """g-vectors (host) of the listed grains at the given translations,
        with the omega-float step from ``ubis`` (default: the grains'
        current UBIs).  Returns (gv, tth, eta, omega_calc, off)."""
        st = self._state[sname]
        pos = [self._pos[sname][n] for n in names]
        idx, owner, off, off_d = self._requests(sname, pos)
        gv_d, tth, eta, omc = self._gv_device(sname, idx, owner, off_d, trans, names, angles,
                                              ubis)
        host = lambda t: None if t is None else t.cpu().numpy()
        return host(gv_d), host(tth), host(eta), host(omc), off

    def _gv_device(self, sname, idx, owner, off_d, trans, names, angles, ubis=None):
        st = self._state[sname]
        p = self.parameterobj.parameters
        st.trig(p["omegasign"])
        if isinstance(trans, torch.Tensor):
            trans_d = trans.reshape(-1, 3)
        else:
            trans_d = torch.as_tensor(np.ascontiguousarray(trans, dtype=np.float64),
                                      device=st.device).reshape(-1, 3)
        gv, tth, eta, omc = batched.grain_gv(st.xl, st.so, st.co, st.omega, p, trans_d, idx,
                                             owner, angles=angles)
        if self.OMEGA_FLOAT:
            if ubis is None:
                ubis = np.array([self.grains[(n, sname)].ubi for n in names], float)
            if isinstance(ubis, torch.Tensor):
                u = ubis.contiguous()
            else:
                u = torch.as_tensor(np.ascontiguousarray(ubis, dtype=np.float64),
                                    device=st.device).reshape(-1, 3, 3)
            mats, _, _ = batched.refine_lists(u, gv, None, off_d, self.tolerance, 1,
                                              order=self._order_of(sname, off_d))
            gv, tth, eta, omc = batched.grain_gv(st.xl, st.so, st.co, st.omega, p, trans_d, idx,
                                                 owner, mats=mats, slop=self.slop, angles=angles)
        return gv, tth, eta, omc

    def _evaluate(self, sname, names, trans, start="read", update=True):
        """compute_gv + refine for every (grain, translation) request:
        per-request (npks, avg_drlv2) of the second score_and_refine call,
        the refined UBIs written to the grains when ``update``."""
        st = self._state[sname]
        pos = np.array([self._pos[sname][n] for n in names], np.int64)
        self._push(sname, names, pos)
        npks, avg = self._evaluate_pos(sname, pos, trans, start, update)
        if update:
            self._pull(sname, names, pos)
        return npks, avg

    def _push(self, sname, names, pos):
        """Grain objects' UBIs -> the device stack (they may have been set
        by the caller between calls)."""
        st = self._state[sname]
        st.ubi_cur[torch.as_tensor(pos, device=st.device)] = torch.as_tensor(
            np.array([self.grains[(n, sname)].ubi for n in names], float), device=st.device)

    def _pull(self, sname, names, pos):
        st = self._state[sname]
        cur = st.ubi_cur[torch.as_tensor(pos, device=st.device)].cpu().numpy()
        for j, n in enumerate(names):
            self.grains[(n, sname)].set_ubi(cur[j])

    def _evaluate_pos(self, sname, pos, trans, start="read", update=True):
        """_evaluate on grain positions and the device UBI stacks: no
        per-grain host work (the lockstep simplex calls it every round)."""
        st = self._state[sname]
        idx, owner, off, off_d = self._requests(sname, pos)
        pos_d = torch.as_tensor(pos, device=st.device)
        gv, _, _, _ = self._gv_device(sname, idx, owner, off_d, trans, None, False,
                                      ubis=st.ubi_cur[pos_d])
        u = (st.ubi_read if start == "read" else st.ubi_cur)[pos_d].contiguous()
        order = self._order_of(sname, off_d)
        if self.latticesymmetry is triclinic:
            u, n, m = batched.refine_lists(u, gv, None, off_d, self.tolerance, 2, order=order)
            k = 1
        else:
            u, _, _ = batched.refine_lists(u, gv, None, off_d, self.tolerance, 1, order=order)
            u = torch.as_tensor(symmetrize(u.cpu().numpy(), self.latticesymmetry),
                                device=st.device)
            u, n, m = batched.refine_lists(u, gv, None, off_d, self.tolerance, 1, order=order)
            u = torch.as_tensor(symmetrize(u.cpu().numpy(), self.latticesymmetry),
                                device=st.device)
            k = 0
        if update:
            st.ubi_cur[pos_d] = u
        nm = torch.stack([n[k].double(), m[k]]).cpu().numpy()
        return nm[0].astype(np.int64), nm[1]

    # -- the reference's per-grain entry points ----------------------------
    def compute_gv(self, thisgrain, update_columns=False):
        """g-vectors of one grain's peaks with the parameters' translation
        (refinegrains.py:366-456); sets self.gv, self.tth, self.eta."""
        key = [k for k, g in self.grains.items() if g is thisgrain][0]
        p = self.parameterobj.parameters
        t = np.array([[p["t_x"], p["t_y"], p["t_z"]]], float)
        gv, tth, eta, omc, _ = self._geometry(key[1], [key[0]], t, angles=True)
        self.tth, self.eta = tth, eta
        if self.OMEGA_FLOAT:
            thisgrain.omega_calc = omc
        else:
            thisgrain.omega_calc = np.zeros(len(thisgrain.ind))
        if update_columns:
            col = self.scandata[key[1]]
            np.put(col.tth_per_grain, thisgrain.ind, tth)
            np.put(col.eta_per_grain, thisgrain.ind, eta)
            if self.OMEGA_FLOAT:
                np.put(col.omegacalc_per_grain, thisgrain.ind, omc)
        self.gv = np.ascontiguousarray(gv)

    def refine(self, ubi, quiet=True):
        """Two score_and_refine calls on self.gv (refinegrains.py:459-498)."""
        dev = self.device
        gv = torch.as_tensor(np.ascontiguousarray(self.gv, dtype=np.float64), device=dev)
        off = torch.tensor([0, gv.shape[0]], dtype=torch.int64, device=dev)
        u = torch.as_tensor(np.array(ubi, float).reshape(1, 3, 3), device=dev)
        u, _, _ = batched.refine_lists(u, gv, None, off, self.tolerance, 1)
        mat = symmetrize(u.cpu().numpy(), self.latticesymmetry)
        u, n, m = batched.refine_lists(torch.as_tensor(mat, device=dev), gv, None, off,
                                       self.tolerance, 1)
        self.npks, self.avg_drlv2 = int(n[0, 0]), float(m[0, 0])
        mat = symmetrize(u.cpu().numpy(), self.latticesymmetry)[0]
        if not quiet:
            print("%-8d %.6f" % (self.npks, math.sqrt(self.avg_drlv2)))
        return mat

    def gof(self, args):
        """1e6 * sum(npks * <drlv2>) / sum(npks) over grains_to_refine
        (refinegrains.py:517-560), every grain in one batched evaluation."""
        self.applyargs(args)
        p = self.parameterobj.parameters
        diffs = 0.
        contribs = 0.
        byscan = {}
        for key in self.grains_to_refine:
            byscan.setdefault(key[1], []).append(key[0])
        results = {}
        t = [p["t_x"], p["t_y"], p["t_z"]]
        for sname, names in byscan.items():
            if self.recompute_xlylzl:
                self._state[sname].lab(p)
            npks, avg = self._evaluate(sname, names, np.array([t] * len(names), float))
            for j, n in enumerate(names):
                results[(n, sname)] = (int(npks[j]), float(avg[j]))
        for key in self.grains_to_refine:
            self.npks, self.avg_drlv2 = results[key]
            diffs += self.npks * self.avg_drlv2
            contribs += self.npks
        if contribs > 0:
            return 1e6 * diffs / contribs
        print("No contribs???", self.grains_to_refine)
        return 1e6

    def estimate_steps(self, gof, guess, steps):
        """Scale the steps to equal objective changes (refinegrains.py:
        563-585)."""
        cen = self.gof(guess)
        deriv = []
        for i in range(len(steps)):
            newguess = [g for g in guess]
            newguess[i] = newguess[i] + steps[i]
            here = self.gof(newguess)
            deriv.append(here - cen)
        j = int(np.argmax(np.absolute(deriv)))
        return [s * abs(deriv[j]) / abs(d) for d, s in zip(deriv, steps)]

    def fit(self, maxiters=100):
        """Fit the varied global parameters with a simplex on gof
        (refinegrains.py:588-627)."""
        self.assignlabels()
        guess = self.parameterobj.get_variable_values()
        inc = self.parameterobj.get_variable_stepsizes()
        names = self.parameterobj.varylist
        self.printresult(guess)
        self.grains_to_refine = list(self.grains.keys())
        self.recompute_xlylzl = any(n not in ("t_x", "t_y", "t_z") for n in names)
        inc = self.estimate_steps(self.gof, guess, inc)
        s = simplex.Simplex(self.gof, guess, inc)
        newguess, error, it = s.minimize(maxiters=maxiters, monitor=1)
        for p, v in zip(names, newguess):
            self.parameterobj.set(p, v)
        for i, t in enumerate(("t_x", "t_y", "t_z")):
            if t in names:
                for g in self.getgrains():
                    self.grains[g].translation[i] = newguess[names.index(t)]
        self.printresult(newguess)
        return newguess, error, it

    def refinepositions(self, quiet=True, maxiters=100):
        """Fit every grain's translation with its own simplex
        (refinegrains.py:636-660), all grains stepped together."""
        self.assignlabels(quiet=True)
        ks = sorted(self.grains.keys())
        tolcache = self.tolerance
        self.tolerance = 1.0
        self.parameterobj.varylist = ["t_x", "t_y", "t_z"]
        try:
            inc = self.parameterobj.get_variable_stepsizes()
            results = {}
            for sname in self.scannames:
                keys = [k for k in ks if k[1] == sname]
                if not keys:
                    continue
                names = [k[0] for k in keys]
                guesses = np.array([self.grains[k].translation for k in keys], float)
                pos = np.array([self._pos[sname][n] for n in names], np.int64)
                self._push(sname, names, pos)

                if self.round_graphs and self.latticesymmetry is triclinic:
                    # every round replays one captured evaluation of all
                    # the scan's grains, the ones not asked for at their
                    # latest points (computed and dropped)
                    rounds = _RoundGraph(self, sname, pos)
                    latest = guesses.copy()

                    def run(points, which, latest=latest, rounds=rounds):
                        latest[which] = points
                        return rounds(latest, which)
                else:
                    def run(points, which, sname=sname, pos=pos):
                        return self._evaluate_pos(sname, pos[which], points)

                def evaluate(points, which, run=run):
                    npks, avg = run(points, which)
                    diffs = npks * avg
                    val = np.where(npks > 0, 1e6 * (0. + diffs) / np.maximum(npks, 1), 1e6)
                    return val
                sb = simplex.SimplexBatch(evaluate, guesses, np.asarray(inc, float))
                best, err, iters = sb.minimize(maxiters=maxiters)
                self._pull(sname, names, pos)
                for j, k in enumerate(keys):
                    self.grains[k].translation = sb.last[j].copy()
                    results[k] = (best[j], err[j], int(iters[j]))
                    if not quiet:
                        print(k, self.grains[k].translation, err[j], int(iters[j]))
                self.evaluation_rounds = sb.rounds
            last = ks[-1]
            self.set_translation(last[0], last[1])
        finally:
            self.tolerance = tolcache
        return results

    def refineubis(self, quiet=True, scoreonly=False):
        """compute_gv + refine of every grain at its own translation
        (refinegrains.py:663-681); returns {key: (npks, avg_drlv2)}."""
        ks = sorted(self.grains.keys())
        out = {}
        for sname in self.scannames:
            keys = [k for k in ks if k[1] == sname]
            if not keys:
                continue
            names = [k[0] for k in keys]
            trans = np.array([self.grains[k].translation for k in keys], float)
            npks, avg = self._evaluate(sname, names, trans, start="current",
                                       update=not scoreonly)
            for j, k in enumerate(keys):
                out[k] = (int(npks[j]), float(avg[j]))
                if not quiet:
                    print("%10s %10s %-8d %.6f" % (k[0], k[1], npks[j], math.sqrt(avg[j])))
        if ks:
            self.set_translation(*ks[-1])
            self.npks, self.avg_drlv2 = out[ks[-1]]
        return out

    def assignlabels(self, quiet=False):
        """Assign every peak to the best grain (refinegrains.py:684-820):
        one fused launch for all grains, then the per-grain lists, the
        per-grain tth/eta columns and the intensity summaries."""
        p = self.parameterobj.parameters
        sign, wvln, wedge, chi = geometry.pack_gv(p)
        for sname in self.scannames:
            st = self._state[sname]
            col = self.scandata[sname]
            st.lab(p)
            st.trig(p["omegasign"])
            names = list(self.grainnames)
            K = len(names)
            self._pos[sname] = {n: j for j, n in enumerate(names)}
            dev = st.device
            ubis = torch.as_tensor(np.array([self.grains[(n, sname)].ubi for n in names], float),
                                   device=dev).reshape(K, 3, 3)
            st.names = names
            st.ubi_cur = ubis.clone()
            st.ubi_read = torch.as_tensor(np.array([self.ubisread[n] for n in names], float),
                                          device=dev).reshape(K, 3, 3)
            trans = torch.as_tensor(np.array([self.grains[(n, sname)].translation for n in names],
                                             float), device=dev).reshape(K, 3)
            labels = torch.full((st.n,), -1, dtype=torch.int32, device=dev)
            drlv2 = torch.ones(st.n, dtype=torch.float64, device=dev)
            glabels = torch.as_tensor(np.array([int(n) for n in names], np.int32), device=dev)
            batched.gv_assign_many(st.xl, st.so, st.co, wvln, ubis, trans, self.tolerance,
                                   drlv2, labels, grain_labels=glabels,
                                   mat=batched.wedge_chi_matrix(wedge, chi))
            # per-grain lists in grain order (peaks ascending in each)
            lab_h = labels.cpu().numpy()
            ints = np.array([int(n) for n in names], np.int64)
            lut = np.full(int(ints.max()) + 2 if K else 1, -1, np.int64)
            lut[ints] = np.arange(K)
            ok = (lab_h >= 0) & (lab_h < len(lut) - 1)
            posarr = np.where(ok, lut[np.where(ok, lab_h, len(lut) - 1)], -1)
            st.idx, st.off, _ = batched.lists_from_labels(torch.as_tensor(posarr, device=dev), K)
            st.off_host = st.off.cpu().numpy()
            st.requests = None
            # the g-vector column is compute_gv with the last grain's translation
            last_t = trans[-1].cpu().numpy() if K else np.zeros(3)
            self.gv = batched.gvectors_trig(st.xl, st.so, st.co,
                                            batched.wedge_chi_matrix(wedge, chi), wvln,
                                            t=tuple(last_t)).cpu().numpy()
            col.addcolumn(lab_h, "labels")
            col.addcolumn(drlv2.cpu().numpy(), "drlv2")
            nr = st.n
            col.addcolumn(np.zeros(nr, np.float32) - 1, "tth_per_grain")
            col.addcolumn(np.zeros(nr, np.float32), "eta_per_grain")
            col.addcolumn(col.omega * 0, "omegacalc_per_grain")
            col.addcolumn(self.gv[:, 0], "gx")
            col.addcolumn(self.gv[:, 1], "gy")
            col.addcolumn(self.gv[:, 2], "gz")
            for name in ("hr", "kr", "lr", "h", "k", "l"):
                col.addcolumn(np.zeros(nr, np.float32), name)
            idx_h = st.idx.cpu().numpy().astype(np.int64)
            # observed tth, eta of every grain's peaks with its translation
            owner = torch.repeat_interleave(torch.arange(K, device=dev),
                                            st.off[1:] - st.off[:-1]).to(torch.int32)
            _, tth, eta, _ = batched.grain_gv(st.xl, st.so, st.co, st.omega, p, trans, st.idx,
                                              owner, angles=True)
            # every column in list order once; each grain takes its slice (the
            # lists are disjoint: a peak has one label)
            xl_l = st.xl[st.idx.long()].cpu().numpy()
            sc_l, fc_l, om_l = (np.take(c, idx_h) for c in (col.sc, col.fc, col.omega))
            col.tth_per_grain[idx_h] = tth.cpu().numpy()
            col.eta_per_grain[idx_h] = eta.cpu().numpy()
            bounds = st.off_host.astype(np.int64)
            for j, n in enumerate(names):
                gr = self.grains[(n, sname)]
                a, b = bounds[j], bounds[j + 1]
                gr.ind = idx_h[a:b]
                gr.peaks_xyz = xl_l[a:b]
                gr.sc = sc_l[a:b]
                gr.fc = fc_l[a:b]
                gr.om = om_l[a:b]
                gr.omega_calc = gr.om.copy()
                gr.npks = int(b - a)
                self.set_translation(n, sname)
                if not quiet:
                    print("Grain", n, "Scan", sname, "npks=", gr.npks)
            compute_lp_factor(col)
            w_l = np.take(_weighted_intensity(col), idx_h)
            t_l = np.take(_tth_column(col), idx_h)
            for j, n in enumerate(names):
                a, b = bounds[j], bounds[j + 1]
                self.grains[(n, sname)].intensity_info = _summary_of(
                    w_l[a:b], t_l[a:b], self.intensity_tth_range, 2, quiet)
