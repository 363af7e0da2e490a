"""Detector/diffractometer parameter packing for the grain-scoring path.

This mirrors what ``transform.Ctransform`` (``ImageD11/transform.py:655-727``)
does before it calls the C kernels: it turns an ImageD11 parameter dict into
the packed arguments of ``compute_xlylzl`` (``p``, ``r``, ``dist``) and of
``compute_gv`` (``omegasign``, ``wavelength``, ``wedge``, ``chi``).  The
packing is host-side set-up of a handful of scalars; all per-peak work runs
in the HIP kernels.
"""
import math

import numpy as np

# ImageD11/transform.py:656-673 (Ctransform.pnames)
PNAMES = (
    "y_center", "z_center", "y_size", "z_size", "distance", "wavelength",
    "omegasign", "tilt_x", "tilt_y", "tilt_z", "o11", "o12", "o21", "o22",
    "wedge", "chi",
)

DEFAULTS = {
    "y_center": 1024.0, "z_center": 1024.0, "y_size": 50.0, "z_size": 50.0,
    "distance": 150000.0, "wavelength": 0.2, "omegasign": 1.0,
    "tilt_x": 0.0, "tilt_y": 0.0, "tilt_z": 0.0,
    "o11": 1.0, "o12": 0.0, "o21": 0.0, "o22": -1.0,
    "wedge": 0.0, "chi": 0.0,
}


def detector_rotation_matrix(tilt_x, tilt_y, tilt_z):
    """R = Rx(tilt_x) . Ry(tilt_y) . Rz(tilt_z), tilts in radians.

    Same convention as ``transform.detector_rotation_matrix``
    (``ImageD11/transform.py:51-82``): right-handed rotations about z, then
    y, then x applied to detector-plane vectors.
    """
    cz, sz = math.cos(tilt_z), math.sin(tilt_z)
    cy, sy = math.cos(tilt_y), math.sin(tilt_y)
    cx, sx = math.cos(tilt_x), math.sin(tilt_x)
    rz = np.array([[cz, -sz, 0.0], [sz, cz, 0.0], [0.0, 0.0, 1.0]])
    ry = np.array([[cy, 0.0, sy], [0.0, 1.0, 0.0], [-sy, 0.0, cy]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cx, -sx], [0.0, sx, cx]])
    return np.dot(np.dot(rx, ry), rz)


def flip_matrix(o11, o12, o21, o22):
    """Detector orientation matrix acting on (0, y, z) detector vectors
    (``ImageD11/transform.py:687-689``)."""
    return np.array([[1.0, 0.0, 0.0], [0.0, o22, o21], [0.0, o12, o11]])


def full_parameters(pars=None):
    """Return a dict with every name of PNAMES, defaults filled in."""
    out = dict(DEFAULTS)
    if pars:
        for k in PNAMES:
            if k in pars:
                out[k] = float(pars[k])
    return out


def pack_xlylzl(pars):
    """(p, r, dist) for compute_xlylzl, as ``Ctransform.reset`` packs them
    (``ImageD11/transform.py:685-693``):
    p = [z_center, y_center, z_size, y_size],
    r = (detector_rotation_matrix . flip).ravel(), dist = [distance, 0, 0].
    """
    q = full_parameters(pars)
    r = np.dot(detector_rotation_matrix(q["tilt_x"], q["tilt_y"], q["tilt_z"]),
               flip_matrix(q["o11"], q["o12"], q["o21"], q["o22"])).ravel()
    p = np.array([q["z_center"], q["y_center"], q["z_size"], q["y_size"]], float)
    dist = np.array([q["distance"], 0.0, 0.0], float)
    return p, np.ascontiguousarray(r), dist


def pack_gv(pars):
    """(omegasign, wavelength, wedge, chi) for compute_gv
    (``ImageD11/transform.py:717-726``)."""
    q = full_parameters(pars)
    return q["omegasign"], q["wavelength"], q["wedge"], q["chi"]


def wedge_chi_matrix(wedge, chi):
    """Rotation the reference builds from wedge and chi in degrees
    (``src/cdiffraction.c:119-142``): mat = W(wedge) . C(chi) with
    W = [[cw,0,-sw],[0,1,0],[sw,0,cw]], C = [[1,0,0],[0,cc,-sc],[0,sc,cc]].
    Row-major, 9 entries.  Host-side helper for simulation and tests; the
    kernels build the same matrix on the device."""
    w = math.radians(wedge)
    c = math.radians(chi)
    W = np.array([[math.cos(w), 0.0, -math.sin(w)], [0.0, 1.0, 0.0],
                  [math.sin(w), 0.0, math.cos(w)]])
    C = np.array([[1.0, 0.0, 0.0], [0.0, math.cos(c), -math.sin(c)],
                  [0.0, math.sin(c), math.cos(c)]])
    return np.dot(W, C).ravel()


def uncompute_vectors(wavelength, wedge, chi):
    """The two fixed inputs of the omega-float step
    (``transform.uncompute_g_vectors``, ``ImageD11/transform.py:529-564``):
    rb = post^T.(-1/wavelength, 0, 0) with post = wedgechi(wedge, chi) =
    wedgemat(wedge).chimat(chi), and pre = chiwedge(chi, wedge)^T
    (``ImageD11/gv_general.py:270-312``).  When wedge == chi == 0 the
    reference skips both rotations: rb is then the beam itself and pre is
    None."""
    beam = np.array([-1.0 / wavelength, 0.0, 0.0])
    if wedge == chi == 0:
        return beam, None
    cw, sw = np.cos(np.radians(wedge)), np.sin(np.radians(wedge))
    cc, sc = np.cos(np.radians(chi)), np.sin(np.radians(chi))
    wm = np.array([[cw, 0, sw], [0, 1, 0], [-sw, 0, cw]])
    cm = np.array([[1, 0, 0], [0, cc, sc], [0, -sc, cc]])
    post = np.dot(wm, cm)
    pre = np.dot(cm, wm).T
    return np.dot(post.T, beam), np.ascontiguousarray(pre)
