"""compute_geometry, compute_xlylzl_xpos_variable and the column-file
geometry update on the GPU against the oracle.

d* and g are bit-exact given the per-peak sin/cos and in the drop-in's
parity mode (sin/cos from the host's glibc); with the device sincos
(correctly rounded, glibc's is not everywhere) within 1e-13 relative.  tth and eta go through atan2, whose device and
glibc versions are both within about an ulp of the true value but not
always equal: they are compared within 4 ulps."""
import os

import numpy as np
import pytest
import torch

from imaged11_amd import _lib, cImageD11, columnfile, geometry, parameters
from imaged11_amd._lib import darr, dptr, stream_ptr

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden")


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def _inputs(n=20000, seed=0):
    rng = np.random.default_rng(seed)
    xl = np.stack([rng.uniform(1e5, 3e5, n), rng.uniform(-1e5, 1e5, n),
                   rng.uniform(-1e5, 1e5, n)], 1)
    return xl, rng.uniform(-360, 360, n)


SETS = [(0.0, 0.0, (0.0, 0.0, 0.0), 1.0), (1.3, -2.1, (12.0, -40.0, 3.0), -1.0),
        (-0.5, 0.7, (200.0, 300.0, -50.0), 1.0)]


@pytest.mark.parametrize("wedge,chi,t,sign", SETS)
def test_compute_geometry_given_trig_bitexact(dev, orc, wedge, chi, t, sign):
    xl, om = _inputs()
    so, co = orc.sincos_omega(om, sign)
    mat = orc.wedge_chi_matrix(wedge, chi)
    want = orc.compute_geometry_trig(xl, so, co, mat, 0.2, np.array(t))
    lib = _lib.load()
    # hold every input tensor until the kernel has run: a temporary's memory
    # returns to torch's allocator as soon as its last reference goes
    dxl, dso, dco = (torch.as_tensor(np.ascontiguousarray(a), device=dev) for a in (xl, so, co))
    out = torch.empty((len(om), 6), dtype=torch.float64, device=dev)
    _lib.check(lib.gs_compute_geometry_trig_dev(dptr(dxl), dptr(dso), dptr(dco), darr(mat), 0.2,
                                                darr(t), dptr(out), len(om), stream_ptr()),
               "geometry_trig")
    got = out.cpu().numpy()
    assert np.array_equal(bits(got[:, 2:]), bits(want[:, 2:]))
    assert ulps(got[:, 0], want[:, 0]).max() <= 4
    assert ulps(got[:, 1], want[:, 1]).max() <= 4


@pytest.mark.parametrize("wedge,chi,t,sign", SETS)
def test_compute_geometry_dropin(dev, orc, wedge, chi, t, sign):
    xl, om = _inputs(seed=1)
    want = np.zeros((len(om), 6))
    orc.compute_geometry(xl, om, sign, 0.2, wedge, chi, np.array(t), want)
    got = np.zeros((len(om), 6))
    cImageD11.compute_geometry(xl, om, sign, 0.2, wedge, chi, t, got)
    # parity mode (the default): ds and g bit for bit, tth/eta within 4 ulps
    assert np.array_equal(bits(got[:, 2:]), bits(want[:, 2:]))
    assert ulps(got[:, 0], want[:, 0]).max() <= 4
    assert ulps(got[:, 1], want[:, 1]).max() <= 4
    got_d = torch.zeros((len(om), 6), dtype=torch.float64, device=dev)
    cImageD11.compute_geometry(torch.as_tensor(xl, device=dev), torch.as_tensor(om, device=dev),
                               sign, 0.2, wedge, chi, t, got_d)
    assert np.array_equal(bits(got_d.cpu().numpy()), bits(got))
    # the device sincos: within 1e-13 relative
    fast = np.zeros((len(om), 6))
    cImageD11.compute_geometry(xl, om, sign, 0.2, wedge, chi, t, fast, parity=False)
    np.testing.assert_allclose(fast, want, rtol=1e-13, atol=1e-13)
    with pytest.raises(ValueError):
        cImageD11.compute_geometry(xl, om, sign, 0.2, wedge, chi, t, np.zeros((len(om), 5)))


def test_xpos_variable_bitexact(dev, orc):
    rng = np.random.default_rng(2)
    n = 30000
    s, f, xp = rng.uniform(0, 2048, n), rng.uniform(0, 2048, n), rng.normal(size=n) * 300
    p = np.array([1024.0, 1000.0, 50.0, 49.0])
    r, dist = rng.normal(size=9), np.array([1.5e5, 3.0, -2.0])
    want = np.zeros((n, 3))
    orc.compute_xlylzl_xpos_variable(s, f, p, r, dist, xp, want)
    got = np.zeros((n, 3))
    cImageD11.compute_xlylzl_xpos_variable(s, f, p, r, dist, xp, got)
    assert np.array_equal(bits(got), bits(want))
    got_d = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    cImageD11.compute_xlylzl_xpos_variable(torch.as_tensor(s, device=dev), f, p, r, dist, xp,
                                           got_d)
    assert np.array_equal(bits(got_d.cpu().numpy()), bits(want))


def test_columnfile_update_geometry(dev, orc):
    """updateGeometry of the makemap fixture (test.flt + test.prm) against
    compute_xlylzl + compute_geometry of the oracle; updateGV gives the
    same g-vectors."""
    c = columnfile.columnfile(os.path.join(G, "makemap_test.flt"))
    p = parameters.read_par_file(os.path.join(G, "makemap_test.prm"))
    for t in (None, (120.0, -35.0, 4.0)):
        c.updateGeometry(p, translation=t)
        pd, r, dist = geometry.pack_xlylzl(p)
        sign, wvln, wedge, chi = geometry.pack_gv(p)
        xl = np.zeros((c.nrows, 3))
        orc.compute_xlylzl(c.xc, c.yc, pd, r, dist, xl)
        assert np.array_equal(bits(np.stack([c.xl, c.yl, c.zl], 1)), bits(xl))
        want = np.zeros((c.nrows, 6))
        tt = (p["t_x"], p["t_y"], p["t_z"]) if t is None else t
        orc.compute_geometry(xl, c.omega, sign, wvln, wedge, chi, np.array(tt, float), want)
        got = np.stack([c[k] for k in columnfile.GEOMETRY_COLUMNS], 1)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
        g = np.stack([c.gx, c.gy, c.gz], 1).copy()
        c.updateGV(p, translation=t)
        assert np.array_equal(bits(np.stack([c.gx, c.gy, c.gz], 1)), bits(g))
