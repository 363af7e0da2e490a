"""Grain (``.ubi``) files (``ImageD11/grain.py``: ``write_grain_file``
:315-342, ``read_grain_file`` :345-373).

Each grain is written as optional ``#translation: x y z``, ``#name``,
``#intensity_info``, ``#npks``, ``#nuniq`` and ``#Rod`` lines, then
``#UBI:`` and the three rows of the matrix (``%.9g``), then a blank line.
"""
import numpy as np


class grain:
    """An orientation matrix UBI with an optional translation."""

    def __init__(self, ubi, translation=None, **kwds):
        self.ubi = np.array(ubi, float).reshape(3, 3)
        self.translation = None if translation is None else np.array(translation, float)
        for k, v in kwds.items():
            setattr(self, k, v)

    def set_ubi(self, ubi):
        """Replace the orientation matrix (``ImageD11/grain.py:60-64``)."""
        self.ubi = np.array(ubi, float).reshape(3, 3)

    @property
    def UB(self):
        return np.linalg.inv(self.ubi)


def write_grain_file(filename, list_of_grains):
    with open(filename, "w") as f:
        for g in list_of_grains:
            t = g.translation
            if t is not None:
                f.write("#translation: %g %g %g\n" % (t[0], t[1], t[2]))
            if hasattr(g, "name"):
                f.write("#name %s\n" % (str(g.name).rstrip()))
            if hasattr(g, "intensity_info"):
                f.write("#intensity_info %s\n" % (str(g.intensity_info).rstrip()))
            if hasattr(g, "npks"):
                f.write("#npks %d\n" % int(g.npks))
            if hasattr(g, "nuniq"):
                f.write("#nuniq %d\n" % int(g.nuniq))
            if hasattr(g, "Rod"):
                f.write("#Rod %f %f %f\n" % tuple(float(r) for r in g.Rod))
            f.write("#UBI:\n")
            u = g.ubi
            for r in range(3):
                f.write("%.9g %.9g %.9g\n" % (u[r, 0], u[r, 1], u[r, 2]))
            f.write("\n")


def read_grain_file(filename):
    """The grains of a .ubi file; name, npks, nuniq and intensity_info are
    kept as the strings the file holds (as the reference does)."""
    grains = []
    rows = []
    t = None
    extra = {}
    with open(filename) as f:
        for line in f:
            if line.startswith("#translation:"):
                t = [float(x) for x in line.split()[1:]]
                continue
            if line.startswith("#") and "UBI" not in line:
                parts = line[1:].split(" ", 1)
                if len(parts) == 2:
                    extra[parts[0]] = parts[1]
                continue
            if line.startswith("#"):
                continue
            vals = [float(x) for x in line.split()]
            if len(vals) == 3:
                rows.append(vals)
            if len(rows) == 3:
                g = grain(rows, t)
                for k in ("name", "npks", "nuniq", "intensity_info"):
                    if k in extra:
                        setattr(g, k, extra[k])
                grains.append(g)
                rows, t, extra = [], None, {}
    return grains
