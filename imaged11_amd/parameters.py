"""ImageD11 parameter files: one ``name value`` pair per line
(``ImageD11/parameters.py``: ``loadparameters`` :502-521, ``saveparameters``
:492-500, ``dumbtypecheck`` :523-555).

Values read as text become ints when they parse as an integer-valued
number, floats when they parse as numbers, and stripped strings otherwise;
the file is written sorted by name with ``str(value)``, so files written
here read back into the same dictionary.
"""


def _coerce(value):
    if not isinstance(value, str):
        return value
    try:
        vf = float(value)
    except ValueError:
        return value.strip()
    try:
        vi = int(value)
    except ValueError:
        return vf
    return vi if abs(vi - vf) < 1e-9 else vf


class parameters:
    """A dictionary of named parameters with the reference's file format,
    and the list of parameters an optimiser varies with their step sizes
    (``ImageD11/parameters.py:398-446``)."""

    def __init__(self, filename=None, **kwds):
        self.parameters = dict(kwds)
        self.varylist = []
        self.stepsizes = {}
        if filename is not None:
            self.loadparameters(filename)

    def loadparameters(self, filename):
        """Lines of exactly two space-separated fields; '-' in names becomes
        '_'; other lines are skipped."""
        with open(filename) as f:
            for line in f:
                fields = line.split(" ")
                if len(fields) != 2:
                    continue
                self.parameters[fields[0].replace("-", "_")] = fields[1]
        self.dumbtypecheck()

    def saveparameters(self, filename):
        with open(filename, "w") as f:
            for key in sorted(self.parameters):
                f.write("%s %s\n" % (key, str(self.parameters[key])))

    def dumbtypecheck(self):
        for k in list(self.parameters):
            self.parameters[k] = _coerce(self.parameters[k])

    def get(self, name, default=None):
        return self.parameters.get(name, default)

    def set(self, name, value):
        self.parameters[name] = value

    def get_parameters(self):
        return self.parameters

    def set_parameters(self, d):
        self.parameters.update(d)
        self.dumbtypecheck()

    def addpar(self, name, value):
        self.parameters[name] = value

    def get_variable_values(self):
        """Values of the parameters in varylist, in its order."""
        return [self.parameters[name] for name in self.varylist]

    def get_variable_stepsizes(self):
        """Step sizes of the parameters in varylist, in its order."""
        return [self.stepsizes[name] for name in self.varylist]

    def set_varylist(self, vl):
        for v in vl:
            if v not in self.parameters:
                raise KeyError("no parameter %r" % (v,))
        self.varylist = list(vl)

    def set_variable_values(self, values):
        if len(values) != len(self.varylist):
            raise ValueError("%d values for %d varied parameters" % (len(values),
                                                                   len(self.varylist)))
        for name, value in zip(self.varylist, values):
            self.parameters[name] = value


def read_par_file(filename):
    """The parameters of a file as a dict."""
    return parameters(filename).parameters
