"""The reference's own test of the geometry kernels (test/test_cImageD11.py:
test_compute_gv, 203 points x 7 (wedge, chi, sign, t) sets against the
numpy transform chain with np.allclose; test_compute_xlylzl, 10 240 random
pixels x 8 detector flips x 27 tilt triples against
transform.compute_xyz_lab, max error < 1e-6), run against the drop-in
imaged11_amd.cImageD11 on the GPU, with numpy inputs and with device
tensors.  The numpy formula chains are restated in tests/test_oracle_math.py."""
import unittest

import numpy as np
import pytest
import torch

from imaged11_amd import cImageD11, geometry
from test_oracle_math import _gv_numpy, _xyz_lab

pytestmark = pytest.mark.gpu


class test_compute_gv(unittest.TestCase):
    def setUp(self):
        N = 203
        self.XLYLZL = np.array([np.linspace(0, 2000, N), np.linspace(10, 2010, N), np.ones(N)])
        self.XLYLZLT = self.XLYLZL.T.copy()
        self.omega = np.linspace(-180, 180, N)

    def runTest(self):
        t = np.array((self.t_x, self.t_y, self.t_z))
        gve1 = _gv_numpy(self.XLYLZLT, self.omega, self.omegasign, self.wvln, self.wedge,
                         self.chi, t)
        gve2 = np.empty((len(self.omega), 3), float)
        cImageD11.compute_gv(self.XLYLZLT, self.omega, self.omegasign, self.wvln, self.wedge,
                             self.chi, t, gve2)
        self.assertTrue(np.allclose(gve1, gve2))
        dev = torch.device("cuda:0")
        gve3 = torch.empty((len(self.omega), 3), dtype=torch.float64, device=dev)
        cImageD11.compute_gv(torch.as_tensor(self.XLYLZLT, device=dev),
                             torch.as_tensor(self.omega, device=dev), self.omegasign, self.wvln,
                             self.wedge, self.chi, t, gve3)
        self.assertTrue(np.array_equal(gve3.cpu().numpy(), gve2))

    def set(self, wedge, chi, omegasign, wvln, tx, ty, tz):
        self.wedge, self.chi, self.omegasign, self.wvln = wedge, chi, omegasign, wvln
        self.t_x, self.t_y, self.t_z = tx, ty, tz

    def test_5_10(self):
        self.set(5., 10., 1.0, 0.125, 20., 30., 40.)
        self.runTest()

    def test_5_0(self):
        self.set(5., 0., 1.0, 0.125, 20., 30., 40.)
        self.runTest()

    def test_0_5(self):
        self.set(0., 5., 1.0, 0.125, 20., 30., 40.)
        self.runTest()

    def test_5_10m(self):
        self.set(5., 10., -1.0, 0.125, 200., 300., 40.)
        self.runTest()

    def test_5_0_tt(self):
        self.set(5., 0., 1.0, 0.125, 200., 300., -400.)
        self.runTest()

    def test_0_0_t0(self):
        self.set(0., 0., 1.0, 0.125, 0., 0., 0.)
        self.runTest()

    def test_0_0_t10(self):
        self.set(0., 0., 1.0, 0.125, 200., 300., 400.)
        self.runTest()


class test_compute_xlylzl(unittest.TestCase):
    def setUp(self):
        npk = 10240
        rng = np.random.default_rng(0)
        self.sc = rng.random(npk) * 2048
        self.fc = rng.random(npk) * 2048
        self.pars = {"z_center": 980., "y_center": 1010., "z_size": 48., "y_size": 49.,
                     "distance": 100100.}

    def test1(self):
        testtilts = [-0.5, 0., 0.24]
        tilts = [(x, y, z) for x in testtilts for y in testtilts for z in testtilts]
        p = self.pars
        pars = np.array((p["z_center"], p["y_center"], p["z_size"], p["y_size"]))
        dist = np.array((p["distance"], 0., 0.))
        ok = 0
        for o11, o12, o21, o22 in [[1, 0, 0, 1], [-1, 0, 0, 1], [-1, 0, 0, -1], [1, 0, 0, -1],
                                   [0, 1, 1, 0], [0, -1, 1, 0], [0, -1, -1, 0], [0, 1, -1, 0]]:
            for tx, ty, tz in tilts:
                q = dict(p, tilt_x=tx, tilt_y=ty, tilt_z=tz, o11=o11, o12=o12, o21=o21, o22=o22)
                xlylzl = _xyz_lab(self.sc, self.fc, q)
                r = np.dot(geometry.detector_rotation_matrix(tx, ty, tz),
                           geometry.flip_matrix(o11, o12, o21, o22))
                outxyz = np.zeros((len(self.sc), 3), float)
                cImageD11.compute_xlylzl(self.sc, self.fc, pars, r.ravel(), dist, outxyz)
                error = np.abs(outxyz - xlylzl)
                self.assertTrue(error.max() < 1e-6)
                ok += 1
        self.assertEqual(ok, 216)


class test_scoring_roundtrip(unittest.TestCase):
    """verify_rounding (the import-time assert of ImageD11/cImageD11.py:163)
    and a score_and_refine fixed point on an exact lattice."""

    def test_verify_rounding(self):
        self.assertEqual(cImageD11.verify_rounding(20), 0)

    def test_fixed_point(self):
        rng = np.random.default_rng(3)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        ub = q / np.array([4.05, 4.05, 4.05])[None, :]
        r = np.arange(-5, 6)
        hkl = np.array(np.meshgrid(r, r, r, indexing="ij")).reshape(3, -1).T.astype(float)
        gv = hkl @ ub.T
        ubi = np.linalg.inv(ub)
        u = ubi.copy()
        n, mean = cImageD11.score_and_refine(u, gv, 1e-6)
        self.assertEqual(n, len(gv))
        self.assertLess(mean, 1e-25)
        self.assertLess(np.abs(u - ubi).max(), 1e-12)
        self.assertEqual(cImageD11.score(ubi, gv, 1e-6), len(gv))
