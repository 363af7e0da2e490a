// Per-grain g-vectors of the refinegrains driver for many (grain, peak)
// list entries in one launch: refinegrains.compute_gv
// (ImageD11/refinegrains.py:366-456) without its per-grain Python loop.
//
// Each list entry q names a peak idx[q] and a request owner[q]; a request
// is a grain at some translation (and, for the omega-float step, with some
// fitted UBI), so one launch serves every grain of a gof evaluation, or
// every trial point of many simplexes at once.
//
// Observed geometry.  transform.compute_tth_eta_from_xyz (transform.py:
// 203-239) and transform.compute_g_vectors (:475-483) at omega*omegasign
// give, in exact arithmetic, the g-vector of compute_gv
// (src/cdiffraction.c:112-168) with the request's translation: the
// scattering direction d = xl - W.C.R(omega).t, tth = atan2(|d_yz|, d_x),
// eta = atan2(-d_y, d_z), k = (d/|d| - x)/wavelength, g = R(omega)^T.
// (W.C)^T.k.  The kernel evaluates that with compute_gv's arithmetic
// (gs_gv_one), and tth, eta in degrees as compute_geometry does.
//
// Omega float (OmFloat = True, refinegrains.py:388-433).  With the
// request's fitted UBI m: hkl = round(m.g) (half to even, as np.round),
// g_calc = m^-1.hkl, and transform.uncompute_g_vectors(g_calc)
// (transform.py:529-564, gv_general.g_to_k/k_to_g, gv_general.py:158-280)
// gives two diffraction angles omega_1,2 with their eta_1,2; the one whose
// eta is nearer the observed eta is taken (ties and NaN as np.argmin),
// the difference to the observed angle is wrapped into [-180, 180] and
// clipped to +-slop, and g is rotated to the new angle (tth and eta, hence
// the lab-frame k, unchanged).  Invalid solutions (|quot| > 1) count as
// omega = eta = 0 like the reference's `* valid`.
#pragma once
#include "gs_kernels.h"

// fixed vectors of uncompute_g_vectors: rb = post^T.(-1/wavelength, 0, 0)
// with post = wedgechi(wedge, chi); pre = chiwedge(chi, wedge)^T, applied
// only when has_pre (the reference skips both when wedge == chi == 0)
struct gs_omfloat {
    double rb[3];
    double pre[9];
    double slop;
    int has_pre;
};

// one solution of k_to_g followed by the eta of its k (gv_general.py:
// 63-91 with axis (0,0,1): p cos q + a (a.p)(1 - cos q) + sin q (a x p))
GS_HD __forceinline__ double gs_float_eta(const gs_omfloat& of, double gx, double gy, double gz,
                                          double omega_deg) {
    double sq, cq;
    gs_sincos(omega_deg * (GS_PI / 180.0), &sq, &cq);
    const double omc = 1.0 - cq;
    const double r0 = (gx * cq + 0.0 * omc) + sq * (-gy);
    const double r1 = (gy * cq + 0.0 * omc) + sq * gx;
    const double r2 = (gz * cq + gz * omc) + sq * 0.0;
    double k1 = r1, k2 = r2;
    if (of.has_pre) {
        k1 = of.pre[3] * r0 + of.pre[4] * r1 + of.pre[5] * r2;
        k2 = of.pre[6] * r0 + of.pre[7] * r1 + of.pre[8] * r2;
    }
    return (180.0 / GS_PI) * atan2(-k1, k2);
}

// the omega-float angle of one peak (signed frame, degrees), given its
// observed g, observed eta and signed observed angle oms
GS_HD __forceinline__ double gs_omega_float(const gs_omfloat& of, const double* m, double g0,
                                            double g1, double g2, double eta, double oms) {
    const double nan = __builtin_nan("");
    double h[3];
    for (int i = 0; i < 3; i++) h[i] = rint(m[3 * i] * g0 + m[3 * i + 1] * g1 + m[3 * i + 2] * g2);
    double mi[9];
    for (int i = 0; i < 9; i++) mi[i] = m[i];
    if (gs_inverse3x3(mi) != 0) return nan;
    double gc[3];
    for (int i = 0; i < 3; i++) gc[i] = mi[3 * i] * h[0] + mi[3 * i + 1] * h[1] + mi[3 * i + 2] * h[2];
    // g_to_k with axis (0, 0, -1): a1 = axis x g = (gy, -gx, 0),
    // a2 = a1 x axis = (gx, gy, 0), a0 = g - a2 = (0, 0, gz)
    const double a0z = gc[2] - 0.0;
    const double rbda0 = (of.rb[0] * (gc[0] - gc[0]) + of.rb[1] * (gc[1] - gc[1])) + of.rb[2] * a0z;
    const double rbda1 = (of.rb[0] * gc[1] + of.rb[1] * (-gc[0])) + of.rb[2] * 0.0;
    const double rbda2 = (of.rb[0] * gc[0] + of.rb[1] * gc[1]) + of.rb[2] * 0.0;
    const double modg = sqrt((gc[0] * gc[0] + gc[1] * gc[1]) + gc[2] * gc[2]);
    const double kdotbeam = -modg * modg / 2.;
    const double phi = atan2(rbda2, rbda1);
    const double den = sqrt(rbda1 * rbda1 + rbda2 * rbda2);
    const bool msk = den <= 0;
    double quot = (kdotbeam - rbda0) / (den + (msk ? 1.0 : 0.0));
    const bool valid = !msk && quot >= -1 && quot <= 1;
    if (!valid) quot = 0.0;
    const double xpp = asin(quot);
    const double s1 = xpp + phi;
    const double s2 = GS_PI - xpp + phi;
    const double deg = 180.0 / GS_PI;
    double ss, cs;
    gs_sincos(s1, &ss, &cs);
    const double om1 = deg * atan2(ss, cs);
    gs_sincos(s2, &ss, &cs);
    const double om2 = deg * atan2(ss, cs);
    const double v = valid ? 1.0 : 0.0;
    const double eta1 = gs_float_eta(of, gc[0], gc[1], gc[2], om1) * v;
    const double eta2 = gs_float_eta(of, gc[0], gc[1], gc[2], om2) * v;
    const double w1 = om1 * v, w2 = om2 * v;
    const double e1 = fabs(eta1 - eta), e2 = fabs(eta2 - eta);
    // np.argmin over (e1, e2): the first NaN wins, else the first minimum
    const double best = (e1 != e1) ? 0.0 : (e2 != e2) ? 1.0 : (e2 < e1 ? 1.0 : 0.0);
    const double omcalc = best * w2 + (1 - best) * w1;
    double err = oms - omcalc;
    err = err - 360 * rint(err / 360.0);
    const double c = err < -of.slop ? -of.slop : (err > of.slop ? of.slop : err);
    return oms - c;
}

// One list entry: observed tth, eta and g for the translation t; with m
// (the fitted UBI) also the omega-float angle and the g-vector at it.
// oms is omega*omegasign of the peak.  Shared by the kernel and the host
// self-test.
GS_HD __forceinline__ void gs_grain_gv_one(const gs_omfloat& of, const double* mat, double ds,
                                           double x0, double x1, double x2, double so,
                                           double co, double oms, const double* t,
                                           const double* m, double* out) {
    const double u0 = co * t[0] - so * t[1];
    const double u1 = so * t[0] + co * t[1];
    const double o0 = mat[0] * u0 + mat[1] * u1 + mat[2] * t[2];
    const double o1 = mat[3] * u0 + mat[4] * u1 + mat[5] * t[2];
    const double o2 = mat[6] * u0 + mat[7] * u1 + mat[8] * t[2];
    const double d0 = x0 - o0, d1 = x1 - o1, d2 = x2 - o2;
    const double deg = 180.0 / GS_PI;
    out[0] = deg * atan2(sqrt(d1 * d1 + d2 * d2), d0);
    out[1] = deg * atan2(-d1, d2);
    const double modyz = 1. / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double k0 = ds * (d0 * modyz - 1.), k1 = ds * d1 * modyz, k2 = ds * d2 * modyz;
    const double v0 = mat[0] * k0 + mat[3] * k1 + mat[6] * k2;
    const double v1 = mat[1] * k0 + mat[4] * k1 + mat[7] * k2;
    const double v2 = mat[2] * k0 + mat[5] * k1 + mat[8] * k2;
    double g0 = co * v0 + so * v1, g1 = -so * v0 + co * v1;
    out[2] = __builtin_nan("");
    if (m) {
        const double omc = gs_omega_float(of, m, g0, g1, v2, out[1], oms);
        double sc, cc;
        gs_sincos(omc * (GS_PI / 180.0), &sc, &cc);
        g0 = cc * v0 + sc * v1;
        g1 = -sc * v0 + cc * v1;
        out[2] = omc;
    }
    out[3] = g0;
    out[4] = g1;
    out[5] = v2;
}

// FLOAT = 0: observed geometry; FLOAT = 1: omega-float g-vectors (mats =
// fitted UBIs per request).  so/co: sin/cos of ((PI/180)*omega)*omegasign
// per peak.  Outputs per entry: tth, eta (optional), omega_calc (FLOAT,
// optional), g (M, 3).  Entries with an index out of range give NaN.
template <int FLOAT>
__global__ __launch_bounds__(GS_BLOCK) void k_grain_gv(
    const double* __restrict__ xl, const double* __restrict__ so_in,
    const double* __restrict__ co_in, const double* __restrict__ omega, int ng, gs_goniometer gp,
    const double* __restrict__ trans, const double* __restrict__ mats, int R, gs_omfloat of,
    const int* __restrict__ idx, const int* __restrict__ owner, long long M,
    double* __restrict__ tth_out, double* __restrict__ eta_out, double* __restrict__ omc_out,
    double* __restrict__ gv_out) {
    __shared__ double smat[9];
    __shared__ double sds;
    gs_block_setup(gp, smat, &sds);
    const long long q = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (q >= M) return;
    const double nan = __builtin_nan("");
    double out[6] = {nan, nan, nan, nan, nan, nan};
    const int k = idx[q], r = owner[q];
    if (k >= 0 && k < ng && r >= 0 && r < R) {
        double mat[9];
        for (int j = 0; j < 9; j++) mat[j] = smat[j];
        const double t[3] = {trans[3 * (size_t)r], trans[3 * (size_t)r + 1],
                             trans[3 * (size_t)r + 2]};
        gs_grain_gv_one(of, mat, sds, xl[3 * (size_t)k], xl[3 * (size_t)k + 1],
                        xl[3 * (size_t)k + 2], so_in[k], co_in[k],
                        FLOAT ? omega[k] * gp.omegasign : 0.0, t,
                        FLOAT ? mats + 9 * (size_t)r : nullptr, out);
    }
    if (tth_out) tth_out[q] = out[0];
    if (eta_out) eta_out[q] = out[1];
    if (FLOAT && omc_out) omc_out[q] = out[2];
    gv_out[3 * q] = out[3];
    gv_out[3 * q + 1] = out[4];
    gv_out[3 * q + 2] = out[5];
}
