"""ImageD11 column files (``.flt`` peak tables) and their geometry update
on the GPU (``ImageD11/columnfile.py``).

Text format: ``# name = value`` header lines carry parameters, one ``#``
line holds the column titles, then one row of numbers per peak
(``readfile`` :293-347, ``writefile`` :266-291; the per-column number
formats are the reference's table, so files written here and there are
alike).  ``updateGeometry`` / ``updateGV`` compute xl, yl, zl, tth, eta, ds
and gx, gy, gz for every peak in one fused kernel pass (``gs_sf2geometry``
/ ``gs_sf2gv``), the work the reference does with
``Ctransform.sf2xyz`` + ``xyz2geometry`` (:469-504).
"""
import numpy as np
import torch

from . import _lib, geometry
from ._lib import darr, dptr, stream_ptr
from .parameters import parameters as Parameters

# number formats per column name (ImageD11/columnfile.py:31-128)
_FLOATS = ["fc", "sc", "omega", "f_raw", "s_raw", "sigf", "sigs", "covsf", "sigo", "covso",
           "covfo", "sum_intensity", "sum_intensity^2", "IMax_int", "IMax_o", "avg_intensity",
           "Min_o", "Max_o", "dety", "detz", "gx", "gy", "gz", "hr", "kr", "zr", "xl", "yl",
           "zl", "drlv2", "tth", "eta", "tth_hist_prob"]
_INTS = ["Number_of_pixels", "IMax_f", "IMax_s", "Min_f", "Max_f", "Min_s", "Max_s",
         "spot3d_id", "spot4d_id", "h", "k", "l", "onfirst", "onlast", "labels", "Grain",
         "grainno", "grain_id", "omega_pair_id", "eta_pair_id", "IKEY", "npk2d"]
_IJ = ["%d%d" % (i, j) for i in range(1, 4) for j in range(1, 4)]
_LONGFLOATS = [s + v for v in _IJ for s in ["U", "UBI"]]
_IJS = [11, 22, 33, 23, 13, 12]
_EXPONENTIALS = [h + str(v) + t for v in _IJS for h, t in
                 [("eps", ""), ("eps", "_s"), ("sig", ""), ("sig", "_s")]]
_EXPONENTIALS += ["%s%d%s%d%s" % (h, _IJS[i], h, _IJS[j], t) for i in range(6)
                  for j in range(i, 6) for h, t in [("e", ""), ("e", "_s"), ("s", ""),
                                                    ("s", "_s")]]
FORMATS = {}
FORMATS.update({k: "%.4f" for k in _FLOATS})
FORMATS.update({k: "%.12f" for k in _LONGFLOATS})
FORMATS.update({k: "%.0f" for k in _INTS})
FORMATS.update({k: "%.4e" for k in _EXPONENTIALS})

GEOMETRY_COLUMNS = ("tth", "eta", "ds", "gx", "gy", "gz")


class columnfile:
    """Named float64 columns of equal length, plus a parameters object."""

    def __init__(self, filename=None, new=False):
        self.__dict__["titles"] = []
        self.__dict__["_cols"] = []
        self.filename = filename
        self.parameters = Parameters()
        self.nrows = 0
        self.ncols = 0
        if not new:
            self.readfile(filename)

    # ---- columns as attributes -------------------------------------------------
    def __getattr__(self, name):
        cols = self.__dict__.get("_cols")
        titles = self.__dict__.get("titles")
        if titles is not None and name in titles:
            return cols[titles.index(name)]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in self.__dict__.get("titles", ()):
            col = self._cols[self.titles.index(name)]
            if np.isscalar(value):
                col[:] = value
            else:
                assert len(value) == self.nrows
                self._cols[self.titles.index(name)] = np.asarray(value)
            return
        self.__dict__[name] = value

    def __getitem__(self, key):
        return self.getcolumn(key)

    def __setitem__(self, key, value):
        if key in self.titles:
            self.getcolumn(key)[:] = value
        else:
            self.addcolumn(value, key)

    def keys(self):
        return self.titles

    @property
    def bigarray(self):
        return np.asarray(self._cols)

    def getcolumn(self, name):
        if name in self.titles:
            return self._cols[self.titles.index(name)]
        raise KeyError("Name " + name + " not in file")

    def addcolumn(self, col, name):
        """Add (or overwrite) column `name`."""
        if len(col) != self.nrows:
            raise ValueError("Wrong length column")
        data = np.asarray(col)
        if name in self.titles:
            self._cols[self.titles.index(name)] = data
        else:
            self.titles.append(name)
            self._cols.append(data)
            self.ncols += 1

    def filter(self, mask):
        """Keep the rows where mask is true."""
        mask = np.asarray(mask, bool)
        assert len(mask) == self.nrows
        self.__dict__["_cols"] = [c[mask] for c in self._cols]
        self.nrows = int(mask.sum())

    def copyrows(self, rows):
        """A new columnfile with the selected rows (a boolean mask or
        indices) and a copy of the parameters (ImageD11/columnfile.py:381-392)."""
        c = columnfile(new=True)
        c.__dict__["titles"] = list(self.titles)
        c.__dict__["_cols"] = [x[rows] for x in self._cols]
        c.nrows = len(c._cols[0]) if c._cols else 0
        c.ncols = self.ncols
        c.parameters = Parameters(**self.parameters.parameters)
        c.filename = self.filename
        return c

    def reorder(self, indices):
        self.__dict__["_cols"] = [c[indices] for c in self._cols]

    def sortby(self, name):
        self.reorder(np.argsort(self.getcolumn(name)))

    def copy(self):
        c = columnfile(new=True)
        c.__dict__["titles"] = list(self.titles)
        c.__dict__["_cols"] = [x.copy() for x in self._cols]
        c.nrows, c.ncols = self.nrows, self.ncols
        c.parameters = Parameters(**self.parameters.parameters)
        c.filename = self.filename
        return c

    # ---- text format -----------------------------------------------------------
    def readfile(self, filename):
        self.__dict__["titles"] = []
        self.parameters = Parameters()
        with open(filename) as f:
            raw = f.readlines()
        i = 0
        while i < len(raw):
            line = raw[i]
            if len(line.strip()) == 0:
                i += 1
                continue
            if line[0] != "#":
                break
            if "=" in line:
                name, value = [s.strip() for s in line[1:].split("=", 1)]
                self.parameters.addpar(name, value)
            else:
                self.__dict__["titles"] = line[1:].split()
            i += 1
        rows = raw[i:]
        if rows:
            first = rows[0].split()
            last = rows[-1].split()
            if len(first) != len(last):  # a truncated last row is dropped
                rows = rows[:-1]
            ncols = len(first)
        else:
            ncols = len(self.titles)
        data = np.empty((len(rows), ncols))
        for r, line in enumerate(rows):
            data[r] = [float(v) for v in line.split()]
        self.__dict__["_cols"] = [np.ascontiguousarray(data[:, j]) for j in range(ncols)]
        self.nrows, self.ncols = len(rows), ncols
        self.parameters.dumbtypecheck()

    def writefile(self, filename):
        with open(filename, "w") as f:
            for p in sorted(self.parameters.get_parameters()):
                f.write("# %s = %s\n" % (p, str(self.parameters.get(p))))
            f.write("#" + "".join("  %s" % t for t in self.titles) + "\n")
            fmt = "".join("  %s" % FORMATS.get(t, "%f") for t in self.titles) + "\n"
            for r in range(self.nrows):
                f.write(fmt % tuple(c[r] for c in self._cols))

    # ---- geometry on the GPU ---------------------------------------------------
    def setparameters(self, pars):
        if isinstance(pars, dict):
            pars = Parameters(**pars)
        self.parameters = pars
        self.parameters.dumbtypecheck()

    def _spot_columns(self):
        if "sc" in self.titles and "fc" in self.titles:
            return self.sc, self.fc
        if "xc" in self.titles and "yc" in self.titles:
            return self.xc, self.yc
        raise ValueError("columnfile file misses xc/yc or sc/fc")

    def _gpu_inputs(self, pars, translation):
        if pars is not None:
            self.setparameters(pars)
        p = self.parameters.get_parameters()
        if "omega" not in self.titles:
            raise ValueError("No omega column")
        sc, fc = self._spot_columns()
        if translation is None:
            translation = (p.get("t_x", 0.0), p.get("t_y", 0.0), p.get("t_z", 0.0))
        dev = torch.device("cuda")
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
        return p, T(sc), T(fc), T(self.omega), [float(v) for v in translation]

    def updateGeometry(self, pars=None, translation=None):
        """Add xl, yl, zl, tth, eta, ds, gx, gy, gz for the current
        parameters (translation: grain position, default t_x, t_y, t_z)."""
        lib = _lib.load()
        _lib.require_gpu()
        p, s, f, w, t = self._gpu_inputs(pars, translation)
        n = self.nrows
        xl = torch.empty((n, 3), dtype=torch.float64, device=s.device)
        out = torch.empty((n, 6), dtype=torch.float64, device=s.device)
        pd, r, dist = geometry.pack_xlylzl(p)
        sign, wvln, wedge, chi = geometry.pack_gv(p)
        _lib.check(lib.gs_sf2geometry(dptr(s), dptr(f), dptr(w), darr(pd), darr(r), darr(dist),
                                      float(sign), float(wvln), float(wedge), float(chi),
                                      darr(t), dptr(xl), dptr(out), n, stream_ptr()),
                   "sf2geometry")
        xl, out = xl.cpu().numpy(), out.cpu().numpy()
        for i, name in enumerate(("xl", "yl", "zl")):
            self.addcolumn(np.ascontiguousarray(xl[:, i]), name)
        for i, name in enumerate(GEOMETRY_COLUMNS):
            self.addcolumn(np.ascontiguousarray(out[:, i]), name)

    def updateGV(self, pars=None, translation=None):
        """Add gx, gy, gz only."""
        from . import batched
        p, s, f, w, t = self._gpu_inputs(pars, translation)
        gv = batched.sf2gv(s, f, w, p, t=t).cpu().numpy()
        for i, name in enumerate(("gx", "gy", "gz")):
            self.addcolumn(np.ascontiguousarray(gv[:, i]), name)


class newcolumnfile(columnfile):
    """An empty column file with the given titles."""

    def __init__(self, titles):
        super().__init__(filename=None, new=True)
        self.__dict__["titles"] = list(titles)
        self.__dict__["_cols"] = [np.zeros(0) for _ in titles]
        self.ncols = len(titles)


def colfile_from_dict(c):
    """A column file holding the equal-length arrays of a dict, titles in
    the dict's order (``ImageD11/columnfile.py:546-555``)."""
    titles = list(c.keys())
    nrows = len(c[titles[0]])
    for t in titles:
        if len(c[t]) != nrows:
            raise ValueError("column %s has %d rows, not %d" % (t, len(c[t]), nrows))
    colf = newcolumnfile(titles)
    colf.nrows = nrows
    colf.__dict__["_cols"] = [np.ascontiguousarray(c[t], dtype=np.float64) for t in titles]
    return colf
