"""Host-side conversions between g-vectors and diffraction angles, for the
reports the indexer writes (``indexer.saveindexing``).

``uncompute_g_vectors`` follows ``transform.uncompute_g_vectors``
(``ImageD11/transform.py:529-564``) with ``gv_general.g_to_k`` /
``k_to_g`` (``ImageD11/gv_general.py:158-312``): for each g-vector the two
rotation angles omega that bring it onto the Ewald sphere, then eta of the
scattered beam for each, and 2theta from |g|.  These are numpy expressions
over a few thousand peaks, run once per report: no kernel is involved.
"""
import math

import numpy as np


def _wedge(w):
    """gv_general.wedgemat (degrees)."""
    c, s = math.cos(math.radians(w)), math.sin(math.radians(w))
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _chi(x):
    """gv_general.chimat (degrees)."""
    c, s = math.cos(math.radians(x)), math.sin(math.radians(x))
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def _about_z(v, deg, sense):
    """Rotate the columns of v (3, n) about the z axis by deg (degrees,
    per column); sense +1 for the axis (0, 0, 1), -1 for (0, 0, -1)."""
    q = np.radians(deg) * sense
    c, s = np.cos(q), np.sin(q)
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]])


def _wrap_pm180(rad):
    """Angles in radians brought into [-pi, pi) (gv_general.angmod), then
    degrees."""
    return np.degrees(np.arctan2(np.sin(rad), np.cos(rad)))


def ewald_angles(g, wavelength, pre=None, post=None):
    """The two rotations (degrees) about -z that satisfy the Laue condition
    for each column of g (3, n), and whether a solution exists.

    With the beam b = post^T (-1/lambda, 0, 0) and rg = pre.g split into its
    part along the axis and the perpendicular pair (axis x rg, (axis x rg) x
    axis), the condition b.k = -|g|^2/2 reads A sin(x) + B cos(x) = C; its
    solutions are x = asin(C/R) + atan2(B, A) and pi - asin(C/R) +
    atan2(B, A) with R = sqrt(A^2 + B^2), valid when R > 0 and |C/R| <= 1
    (gv_general.g_to_k, ImageD11/gv_general.py:192-280)."""
    rg = g if pre is None else np.dot(pre, g)
    b = np.array([-1.0 / wavelength, 0.0, 0.0])
    if post is not None:
        b = np.dot(post.T, b)
    # axis (0, 0, -1): axis x rg = (rg_y, -rg_x, 0), (axis x rg) x axis =
    # (rg_x, rg_y, 0), the component along the axis = (0, 0, rg_z)
    a_perp1 = np.array([rg[1], -rg[0], np.zeros_like(rg[0])])
    a_perp2 = np.array([rg[0], rg[1], np.zeros_like(rg[0])])
    along = rg - a_perp2
    b0 = b @ along
    A = b @ a_perp1
    B = b @ a_perp2
    C = -np.sum(g * g, 0) / 2.0 - b0
    R = np.sqrt(A * A + B * B)
    none = R <= 0
    ratio = C / (R + none)
    valid = (~none) & (ratio >= -1) & (ratio <= 1)
    x = np.arcsin(np.where(valid, ratio, 0.0))
    p = np.arctan2(B, A)
    return _wrap_pm180(x + p), _wrap_pm180(math.pi - x + p), valid


def uncompute_g_vectors(g, wavelength, wedge=0.0, chi=0.0):
    """(tth, [eta1, eta2], [omega1, omega2]) in degrees for g-vectors g
    (3, n); entries without a Laue solution are 0 (ImageD11/transform.py:
    529-564)."""
    g = np.asarray(g, float)
    post = None if wedge == chi == 0 else np.dot(_wedge(wedge), _chi(chi))
    om1, om2, valid = ewald_angles(g, wavelength, pre=None, post=post)
    pre = None if post is None else np.dot(_chi(chi), _wedge(wedge)).T
    etas = []
    for om in (om1, om2):
        k = _about_z(g, om, +1)
        if pre is not None:
            k = np.dot(pre, k)
        etas.append(np.degrees(np.arctan2(-k[1], k[2])) * valid)
    ds = np.sqrt(np.sum(g * g, 0))
    tth = np.degrees(np.arcsin(ds * wavelength / 2.0) * 2.0) * valid
    return tth, etas, [om1 * valid, om2 * valid]
