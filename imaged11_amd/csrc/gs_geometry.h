// Column-file geometry: compute_geometry (tth, eta, d*, g per peak) and
// compute_xlylzl_xpos_variable, plus the fused raw-peak pass that
// columnfile.updateGeometry makes (compute_xlylzl -> compute_geometry).
#pragma once
#include "gs_kernels.h"

// One peak of compute_geometry (src/cdiffraction.c:31-94): the grain
// origin and scattering direction as in compute_gv, then
//   tth = DEG * atan2(sqrt(d1^2 + d2^2), d0),  eta = DEG * atan2(-d1, d2),
//   d*  = |k|,  g = Omega^T mat^T k.
// DEG expands to 180.0 / PI before the product, as the reference's macro.
// The g-vector and d* use the reference's operation order (bit parity for
// given sin/cos); atan2 is the device library's (within an ulp or two of
// the host libm's, whose own atan2 is not correctly rounded either).
GS_HD __forceinline__ void gs_geometry_one(double xl0, double xl1, double xl2, double so,
                                           double co, const double* mat, double ds, double t0,
                                           double t1, double t2, double* out) {
    const double deg = 180.0 / 3.14159265358979323846;
    double u0 = co * t0 - so * t1;
    double u1 = so * t0 + co * t1;
    double u2 = t2;
    double o0 = mat[0] * u0 + mat[1] * u1 + mat[2] * u2;
    double o1 = mat[3] * u0 + mat[4] * u1 + mat[5] * u2;
    double o2 = mat[6] * u0 + mat[7] * u1 + mat[8] * u2;
    double d0 = xl0 - o0, d1 = xl1 - o1, d2 = xl2 - o2;
    double modyz = 1. / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    out[0] = deg * atan2(sqrt(d1 * d1 + d2 * d2), d0);
    double k0 = ds * (d0 * modyz - 1.);
    double k1 = ds * d1 * modyz;
    double k2 = ds * d2 * modyz;
    out[1] = deg * atan2(-d1, d2);
    out[2] = sqrt(k0 * k0 + k1 * k1 + k2 * k2);
    double v0 = mat[0] * k0 + mat[3] * k1 + mat[6] * k2;
    double v1 = mat[1] * k0 + mat[4] * k1 + mat[7] * k2;
    double v2 = mat[2] * k0 + mat[5] * k1 + mat[8] * k2;
    out[3] = co * v0 + so * v1;
    out[4] = -so * v0 + co * v1;
    out[5] = v2;
}

// compute_geometry over n peaks; TRIG: sin/cos of omega given per peak.
// out is (n, 6) row-major: tth, eta, ds, gx, gy, gz.
template <int TRIG>
__global__ __launch_bounds__(GS_BLOCK) void k_geometry(
    const double* __restrict__ xl, const double* __restrict__ omega,
    const double* __restrict__ so_in, const double* __restrict__ co_in, gs_goniometer gp,
    double* __restrict__ out, int n) {
    __shared__ double smat[9];
    __shared__ double sds;
    gs_block_setup(gp, smat, &sds);
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= n) return;
    double so, co;
    if (TRIG) {
        so = so_in[i];
        co = co_in[i];
    } else {
        gs_sincos(gs_omega_rad(omega[i], gp.omegasign), &so, &co);
    }
    double mat[9];
    for (int j = 0; j < 9; j++) mat[j] = smat[j];
    double r[6];
    gs_geometry_one(xl[3 * i], xl[3 * i + 1], xl[3 * i + 2], so, co, mat, sds, gp.t[0], gp.t[1],
                    gp.t[2], r);
    for (int j = 0; j < 6; j++) out[6 * i + j] = r[j];
}

// compute_xlylzl_xpos_variable (src/cdiffraction.c:245-277): compute_xlylzl
// with a per-peak offset xpos subtracted from x after the distance.
__global__ __launch_bounds__(GS_BLOCK) void k_xlylzl_xpos(const double* __restrict__ s,
                                                          const double* __restrict__ f,
                                                          gs_detector det,
                                                          const double* __restrict__ xpos,
                                                          double* __restrict__ out, int n) {
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= n) return;
    double x0, x1, x2;
    gs_xlylzl_one(s[i], f[i], det.p, det.r, det.dist, x0, x1, x2);
    out[3 * i + 0] = x0 - xpos[i];
    out[3 * i + 1] = x1;
    out[3 * i + 2] = x2;
}

// columnfile.updateGeometry in one pass (ImageD11/columnfile.py:469-504):
// raw (s, f, omega) -> xl, yl, zl (optional) and tth, eta, ds, gx, gy, gz.
__global__ __launch_bounds__(GS_BLOCK) void k_sf2geometry(
    const double* __restrict__ s, const double* __restrict__ f, const double* __restrict__ omega,
    gs_detector det, gs_goniometer gp, double* __restrict__ xl_out, double* __restrict__ out,
    int n) {
    __shared__ double smat[9];
    __shared__ double sds;
    gs_block_setup(gp, smat, &sds);
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= n) return;
    double x0, x1, x2;
    gs_xlylzl_one(s[i], f[i], det.p, det.r, det.dist, x0, x1, x2);
    double so, co;
    gs_sincos(gs_omega_rad(omega[i], gp.omegasign), &so, &co);
    double mat[9];
    for (int j = 0; j < 9; j++) mat[j] = smat[j];
    double r[6];
    gs_geometry_one(x0, x1, x2, so, co, mat, sds, gp.t[0], gp.t[1], gp.t[2], r);
    if (xl_out) {
        xl_out[3 * i + 0] = x0;
        xl_out[3 * i + 1] = x1;
        xl_out[3 * i + 2] = x2;
    }
    for (int j = 0; j < 6; j++) out[6 * i + j] = r[j];
}
