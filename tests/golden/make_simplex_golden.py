"""Write tests/golden/simplex_golden.json: the evaluation sequence of the
reference's Nelder-Mead (ImageD11/simplex.py) on five small problems.

The simplex's trajectory (which vertex is reflected, expanded, contracted
or shrunk, in which order the objective is called) has no closed form, so
it is pinned by running the reference's own module once, here, on a CPU:
the file is loaded by path (it imports only math, copy and sys), without
the ImageD11 package.  The fixture holds, per problem, every point the
objective was called with, the value returned, and minimize()'s result.

    python tests/golden/make_simplex_golden.py [REFERENCE_ROOT]

Nothing here is imported by the product or the tests; the tests read the
JSON file.
"""
import importlib.util
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REFROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def load_reference_simplex():
    path = os.path.join(REFROOT, "ImageD11", "simplex.py")
    spec = importlib.util.spec_from_file_location("reference_simplex", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cubic_abs(a):
    return abs(a[0] * a[0] * a[0] * 5 - a[1] * a[1] * 7 + math.sqrt(abs(a[0])) - 118)


def rosenbrock(a):
    return (1 - a[0]) ** 2 + 100 * (a[1] - a[0] * a[0]) ** 2


def stepped(a):
    # plateaus: many equal errors, exercising the tie rules
    return round((a[0] - 1) ** 2 + (a[1] + 2) ** 2 + (a[2] - 0.5) ** 2, 1)


def bowl(a):
    return (a[0] - 3) ** 2 + 2 * (a[1] + 1) ** 2 + 0.5 * (a[2] - 2) ** 2 + a[0] * a[1] * 0.1


PROBLEMS = [
    ("cubic_abs", cubic_abs, [1.0, 1.0, 1.0], [2.0, 4.0, 6.0], {}),
    ("rosenbrock", rosenbrock, [-1.2, 1.0], [0.1, 0.1], {"maxiters": 250}),
    ("stepped", stepped, [0.0, 0.0, 0.0], [0.3, 0.3, 0.3], {"maxiters": 60}),
    ("bowl_5_iters", bowl, [0.0, 0.0, 0.0], [0.5, 0.5, 0.5], {"maxiters": 5}),
    ("bowl_0_iters", bowl, [0.0, 0.0, 0.0], [0.5, 0.5, 0.5], {"maxiters": 0}),
]


def main():
    mod = load_reference_simplex()
    out = {}
    for name, f, guess, inc, kw in PROBLEMS:
        calls = []

        def rec(args, f=f, calls=calls):
            v = f(args)
            calls.append([list(map(float, args)), float(v)])
            return v
        s = mod.Simplex(rec, list(guess), list(inc))
        res, err, it = s.minimize(monitor=0, **kw)
        out[name] = {"guess": guess, "increments": inc, "kwargs": kw, "calls": calls,
                     "result": list(map(float, res)), "error": float(err), "iter": int(it)}
    with open(os.path.join(HERE, "simplex_golden.json"), "w") as f:
        json.dump(out, f, indent=0)
    print({k: (len(v["calls"]), v["iter"]) for k, v in out.items()})


if __name__ == "__main__":
    main()
