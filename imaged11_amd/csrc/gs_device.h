// Device arithmetic shared by the grain-scoring kernels (gfx950, FP64 VALU).
//
// Everything that decides a count, a label or a drlv2 value is written as
// the reference writes it: left-to-right sums of separately rounded
// products (this file is compiled with -ffp-contract=off), rounding to the
// nearest integer by two literal adds of 1.5*2^52, and the wedge/chi and
// geometry expressions in the reference's macro-expansion order.  FMA is
// used only where the result is not compared with the reference bit for
// bit: inside the double-double sincos, which is accurate to ~2^-80 and
// then rounded once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_trig_constants.h"

#pragma clang fp contract(off)

#define GS_HD __host__ __device__
#define GS_MAGIC 6755399441055744.0
#define GS_PI 3.14159265358979323846
#define GS_WAVE 64

// Round to nearest, ties to even, for |x| < 2^51; bit-identical to
// (x + MAGIC) - MAGIC for every input (src/closest.c:84-85).
GS_HD __forceinline__ double gs_round(double x) {
    return (x + GS_MAGIC) - GS_MAGIC;
}

// h_i = (u_i0*g0 + u_i1*g1) + u_i2*g2 and its squared distance to the
// nearest integer triple (src/closest.c:236-242, :288-294).
struct gs_ubi {
    double u00, u01, u02, u10, u11, u12, u20, u21, u22;
};

GS_HD __forceinline__ gs_ubi gs_load_ubi(const double* __restrict__ u) {
    gs_ubi r;
    r.u00 = u[0]; r.u01 = u[1]; r.u02 = u[2];
    r.u10 = u[3]; r.u11 = u[4]; r.u12 = u[5];
    r.u20 = u[6]; r.u21 = u[7]; r.u22 = u[8];
    return r;
}

GS_HD __forceinline__ double gs_drlv2(const gs_ubi& u, double x, double y, double z,
                                           double& ih0, double& ih1, double& ih2) {
    double h0 = u.u00 * x + u.u01 * y + u.u02 * z;
    double h1 = u.u10 * x + u.u11 * y + u.u12 * z;
    double h2 = u.u20 * x + u.u21 * y + u.u22 * z;
    ih0 = gs_round(h0);
    ih1 = gs_round(h1);
    ih2 = gs_round(h2);
    double t0 = h0 - ih0, t1 = h1 - ih1, t2 = h2 - ih2;
    return t0 * t0 + t1 * t1 + t2 * t2;
}

GS_HD __forceinline__ double gs_drlv2(const gs_ubi& u, double x, double y, double z) {
    double a, b, c;
    return gs_drlv2(u, x, y, z, a, b, c);
}

// The certified variant of gs_drlv2: h with FMA, rounding with rint
// (v_rndne, equal to the MAGIC rounding for |h| < 2^51), squares summed
// with FMA.  Used only together with gs_cert_window.
GS_HD __forceinline__ double gs_drlv2_fma(const gs_ubi& u, double x, double y, double z) {
    double h0 = __builtin_fma(u.u02, z, __builtin_fma(u.u01, y, u.u00 * x));
    double h1 = __builtin_fma(u.u12, z, __builtin_fma(u.u11, y, u.u10 * x));
    double h2 = __builtin_fma(u.u22, z, __builtin_fma(u.u21, y, u.u20 * x));
    double t0 = h0 - rint(h0), t1 = h1 - rint(h1), t2 = h2 - rint(h2);
    return __builtin_fma(t2, t2, __builtin_fma(t1, t1, t0 * t0));
}

// Half-width E of the window around tol^2 inside which gs_drlv2_fma cannot
// decide `gs_drlv2 < tol^2`, for peaks whose components are bounded by
// gmax in magnitude (derivation at k_score_many_fast).  +inf when the
// bound does not apply (U1*gmax >= 2^49, or a non-finite gmax).
GS_HD __forceinline__ double gs_cert_window(const gs_ubi& u, double gmax, double tolsq) {
    const double u1 = fabs(u.u00) + fabs(u.u01) + fabs(u.u02) + fabs(u.u10) + fabs(u.u11) +
                      fabs(u.u12) + fabs(u.u20) + fabs(u.u21) + fabs(u.u22);
    const double span = u1 * gmax;
    if (!(span < 562949953421312.0)) return __builtin_inf();  // 2^49
    return 1e-15 * span + 1e-15 + 2e-15 * tolsq;
}

// x - floor(x) in [0, 1): v_fract_f64 on the device (exact for |x| < 2^52),
// the same value on the host
GS_HD __forceinline__ double gs_fract(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_fract(x);
#else
    return x - floor(x);
#endif
}

// The score kernel's evaluation, one instruction per pair cheaper than
// gs_drlv2_fma: the rows are formed as h + 1/2 (the 1/2 enters the first
// FMA of each row), f = fract(h + 1/2) so that f - 1/2 is the distance of
// h to its nearest integer (half-up on ties, where the square is the same),
// and sum f (f - 1) = sum (f - 1/2)^2 - 3/4 = drlv2 - 3/4 in exact
// arithmetic, 5 instructions for the three rows.  Compared against
// tol^2 - 3/4 +- gs_cert_window_fract.
GS_HD __forceinline__ double gs_drlv2_fract(const gs_ubi& u, double x, double y, double z) {
    const double h0 = __builtin_fma(u.u02, z, __builtin_fma(u.u01, y, __builtin_fma(u.u00, x, 0.5)));
    const double h1 = __builtin_fma(u.u12, z, __builtin_fma(u.u11, y, __builtin_fma(u.u10, x, 0.5)));
    const double h2 = __builtin_fma(u.u22, z, __builtin_fma(u.u21, y, __builtin_fma(u.u20, x, 0.5)));
    const double f0 = gs_fract(h0), f1 = gs_fract(h1), f2 = gs_fract(h2);
    double t = __builtin_fma(f0, f0, -f0);
    t = __builtin_fma(f1, f1, t) - f1;
    return __builtin_fma(f2, f2, t) - f2;
}

// Half-width of the undecided window of gs_drlv2_fract around tol^2 - 3/4.
// Each row h + 1/2 is within gamma_3 (A_i + 1/2) of exact (A_i = sum_j
// |u_ij||g_j|, gamma_3 <= 6.7e-16); x -> (x - round x)^2 is 1-Lipschitz,
// so the rows move the sum by <= 6.7e-16 (U1 Gmax + 3/2); the five rounded
// operations on terms of magnitude <= 3/4 add <= 5e-16; the reference's
// own sum of squares <= 5e-16.  E = 1e-15 U1 Gmax + 3e-15 + 2e-15 tol^2
// covers it with margin, the last terms also absorbing the rounding of the
// thresholds (tol^2 - 3/4) +- E.  +inf when the bound does not apply
// (U1 Gmax >= 2^49, where v_rndne/fract and the MAGIC rounding may differ,
// or a non-finite Gmax).
GS_HD __forceinline__ double gs_cert_window_fract(const gs_ubi& u, double gmax, double tolsq) {
    const double u1 = fabs(u.u00) + fabs(u.u01) + fabs(u.u02) + fabs(u.u10) + fabs(u.u11) +
                      fabs(u.u12) + fabs(u.u20) + fabs(u.u21) + fabs(u.u22);
    const double span = u1 * gmax;
    if (!(span < 562949953421312.0)) return __builtin_inf();  // 2^49
    return 1e-15 * span + 3e-15 + 2e-15 * tolsq;
}

// Radius of the one-row sieve of k_score_many_sieve: a pair whose row 0,
// evaluated as y = h0 + r by three FMAs (r entering the first), has
// fract(y) >= 2r lies at least r - eps from every integer (eps = the error
// of y and of fract), and then its reference drlv2 >= tol^2, so it is not
// counted.  Equivalently, with d = y - r - round(y - r): |d| >= r.  Let t be
// the reference's row-0 difference h0 - round(h0) (exact for |h0| < 2^51,
// |t| = the distance of h0 to the nearest integer).  The reference's drlv2
// is a sum of three non-negative rounded squares that starts from fl(t^2),
// and rounding is monotone, so drlv2 >= t^2 (1 - u).  For r < 1/2 (else
// 2r > fract(y) always and every pair passes) y is within gamma_3 (A_0 +
// 1/2) of h0 + r and the reference's h0 within gamma_3 A_0 of the exact
// dot product (A_0 = sum_j |u_0j||g_j| <= (|u00| + |u01| + |u02|) Gmax),
// fract adds <= u, and the distance is 1-Lipschitz: ||d| - |t|| <= 6.7e-16
// A_0 + 5e-16.  With r = tol (1 + 1e-15) + 1e-15 A_0 + 2e-15, |d| >= r
// gives |t| >= tol (1 + 1e-15) and t^2 (1 - u) > tol^2; the margins also
// absorb the rounding of
// sqrt(tol^2) and of r itself.  +inf where gs_cert_window_fract is +inf
// (every pair then reaches the exact recomputation).  A NaN d or r lets the
// pair through, so non-finite inputs are decided by the full test.
GS_HD __forceinline__ double gs_sieve_radius(const gs_ubi& u, double gmax, double tolsq) {
    const double a0 = fabs(u.u00) + fabs(u.u01) + fabs(u.u02);
    const double u1 = a0 + fabs(u.u10) + fabs(u.u11) + fabs(u.u12) + fabs(u.u20) + fabs(u.u21) +
                      fabs(u.u22);
    if (!(u1 * gmax < 562949953421312.0)) return __builtin_inf();  // 2^49
    return sqrt(tolsq) * (1.0 + 1e-15) + 1e-15 * (a0 * gmax) + 2e-15;
}

GS_HD __forceinline__ double gs_abs_finite(double a) {
    return a == a ? fabs(a) : 0.0;  // NaN rows never count on either path
}

// ---------------------------------------------------------------------------
// 3x3 inverse in place, the cofactor formula and operation order of
// src/closest.c:37-76; returns 0, or -1 when det == 0 exactly.
GS_HD __forceinline__ int gs_inverse3x3(double* A) {
    double det = A[0] * (A[8] * A[4] - A[7] * A[5]) - A[3] * (A[8] * A[1] - A[7] * A[2]) +
                 A[6] * (A[5] * A[1] - A[4] * A[2]);
    if (!(det != 0.)) return -1;
    double B[9];
    B[0] = (A[8] * A[4] - A[7] * A[5]) / det;
    B[1] = -(A[8] * A[1] - A[7] * A[2]) / det;
    B[2] = (A[5] * A[1] - A[4] * A[2]) / det;
    B[3] = -(A[8] * A[3] - A[6] * A[5]) / det;
    B[4] = (A[8] * A[0] - A[6] * A[2]) / det;
    B[5] = -(A[5] * A[0] - A[3] * A[2]) / det;
    B[6] = (A[7] * A[3] - A[6] * A[4]) / det;
    B[7] = -(A[7] * A[0] - A[6] * A[1]) / det;
    B[8] = (A[4] * A[0] - A[3] * A[1]) / det;
    for (int i = 0; i < 9; i++) A[i] = B[i];
    return 0;
}

// UB = R . H^-1 then UBI = UB^-1, written to ubi only if both inverses
// exist (src/closest.c:320-336).  H is the full symmetric matrix.
GS_HD __forceinline__ void gs_solve_ubi(double* ubi, const double* R, double* H) {
    double UB[9];
    for (int i = 0; i < 9; i++) UB[i] = 0.;
    int k = gs_inverse3x3(H);
    if (k == 0) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++) UB[3 * i + j] = UB[3 * i + j] + R[3 * i + l] * H[3 * l + j];
    }
    if (k == 0 && gs_inverse3x3(UB) == 0)
        for (int i = 0; i < 9; i++) ubi[i] = UB[i];
}

// ---------------------------------------------------------------------------
// Double-double helpers (error-free transformations with FMA).
struct gs_dd {
    double hi, lo;
};

GS_HD __forceinline__ gs_dd gs_two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

GS_HD __forceinline__ gs_dd gs_fast_two_sum(double a, double b) {
    double s = a + b;
    return {s, b - (s - a)};
}

GS_HD __forceinline__ gs_dd gs_two_prod(double a, double b) {
    double p = a * b;
    return {p, __builtin_fma(a, b, -p)};
}

// a*b for double-double a, b (relative error ~2^-104)
GS_HD __forceinline__ gs_dd gs_dd_mul(gs_dd a, gs_dd b) {
    gs_dd p = gs_two_prod(a.hi, b.hi);
    p.lo = __builtin_fma(a.hi, b.lo, p.lo);
    p.lo = __builtin_fma(a.lo, b.hi, p.lo);
    return gs_fast_two_sum(p.hi, p.lo);
}

// a+b when |b| is well below |a| (Horner steps of a convergent series)
GS_HD __forceinline__ gs_dd gs_dd_add_small(gs_dd a, gs_dd b) {
    gs_dd s = gs_two_sum(a.hi, b.hi);
    s.lo += a.lo + b.lo;
    return gs_fast_two_sum(s.hi, s.lo);
}

// a - b for general double-double values
GS_HD __forceinline__ gs_dd gs_dd_sub(gs_dd a, gs_dd b) {
    gs_dd s = gs_two_sum(a.hi, -b.hi);
    gs_dd t = gs_two_sum(a.lo, -b.lo);
    s.lo += t.hi;
    s = gs_fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return gs_fast_two_sum(s.hi, s.lo);
}

// sin and cos of a double argument x (radians), correctly rounded except
// in extremely rare near-midpoint cases.  Cody-Waite reduction by pi/2
// held in four pieces (exact for |k| < 2^20) to r = rh + rl, |r| <= pi/4;
// then a table path that settles almost every argument, and for the rest
// Taylor series on r with the leading terms in double-double, accurate to
// ~2^-80 relative and rounded once.  Arguments beyond 2^19*pi use the
// device library.  Inlined and returned by value, so the pair stays in
// registers (a call through pointers put it in per-lane scratch memory).
struct gs_sc {
    double s, c;
};

// x = k pi/2 + r, r as a double-double; *q = k mod 4
GS_HD __forceinline__ gs_dd gs_reduce_pio2(double x, int* q) {
    const double kf = rint(x * GS_2_OVER_PI);
    *q = ((int)kf) & 3;
    const double t = x - kf * GS_PIO2_1;  // exact: k*PIO2_1 has <= 53 bits, Sterbenz
    const gs_dd p2 = gs_two_prod(kf, GS_PIO2_2);
    gs_dd r = gs_dd_sub({t, 0.0}, p2);
    r = gs_dd_sub(r, {kf * GS_PIO2_3, 0.0});  // exact product (33+20 bits)
    return gs_dd_sub(r, {kf * GS_PIO2_4, 0.0});
}

// sin r and cos r for |r| <= pi/4 by the series, each rounded once from a
// double-double value accurate to ~2^-80 relative
GS_HD __forceinline__ gs_sc gs_sincos_series(gs_dd r) {
    gs_dd r2 = gs_dd_mul(r, r);
    double z = r2.hi;
    // sin(r) = r * (1 + r2 * S(r2)),   cos(r) = 1 + r2 * C(r2)
    double ts = GS_SIN_TAIL[7];
#pragma unroll
    for (int i = 6; i >= 0; i--) ts = __builtin_fma(ts, z, GS_SIN_TAIL[i]);
    double tc = GS_COS_TAIL[7];
#pragma unroll
    for (int i = 6; i >= 0; i--) tc = __builtin_fma(tc, z, GS_COS_TAIL[i]);
    gs_dd S = {ts, 0.0};
#pragma unroll
    for (int i = 3; i >= 0; i--) S = gs_dd_add_small({GS_SIN_DD[i][0], GS_SIN_DD[i][1]}, gs_dd_mul(S, r2));
    gs_dd C = {tc, 0.0};
#pragma unroll
    for (int i = 4; i >= 0; i--) C = gs_dd_add_small({GS_COS_DD[i][0], GS_COS_DD[i][1]}, gs_dd_mul(C, r2));
    gs_dd sr = gs_dd_mul(r, gs_dd_add_small({1.0, 0.0}, gs_dd_mul(r2, S)));
    gs_dd cr = gs_dd_add_small({1.0, 0.0}, gs_dd_mul(r2, C));
    return {sr.hi + sr.lo, cr.hi + cr.lo};
}

// Half the spacing of the doubles next to y on the side toward zero:
// 2^(e-53) for 2^e <= |y| < 2^(e+1), 2^(e-54) when |y| = 2^e; 0 where that
// is not a normal double (the caller then takes the series).
GS_HD __forceinline__ double gs_half_spacing(double y) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, y) & 0x7fffffffffffffffull;
    const unsigned long long ex = b >> 52;
    const unsigned long long sh = (b & 0xfffffffffffffull) ? 53ull : 54ull;
    return ex > sh ? __builtin_bit_cast(double, (ex - sh) << 52) : 0.0;
}

// y = yh + yl (|yl| <= half an ulp of yh) within e of the exact value v:
// RN(v) = yh when |yl| + e stays below the half spacing next to yh (the
// computed sum rounds up to it whenever the exact sum exceeds it).
GS_HD __forceinline__ bool gs_rounds_to(double yh, double yl, double e) {
    return fabs(yl) + e < gs_half_spacing(yh);
}

// The table path.  With a = j/128 the nearest multiple to rh and b = r - a
// (|b| <= 1/256), sin r = Sa + Ca b + Sa (cos b - 1) + Ca (sin b - b) and
// cos r = Ca - Sa b + Ca (cos b - 1) - Sa (sin b - b), Sa and Ca from the
// double-double table, b^2, Ca bh and Sa (-b^2/2) exact by FMA, the rest
// in double, smallest terms first.  Bounds on the error of y = yh + yl
// (b's part rl, the table's 2^-106, the short series' truncation b^9/9!
// and rounding, the summation): below 2^-76 absolute and, for j = 0,
// 2^-69 relative for sin, with |sin r| >= sin(1/256) for j != 0, so
// 2^-68 relative; below 2^-78 absolute for cos, with cos r >= 0.7.  The
// tests use 2^-66 and 2^-72 relative.  Returns false (the caller takes the
// series) when either value is not settled by gs_rounds_to.
GS_HD __forceinline__ bool gs_sincos_table(gs_dd r, gs_sc* out) {
    const double rh = r.hi, rl = r.lo;
    const double jf = rint(rh * GS_TBL_STEP);
    const double aj = fabs(jf);
    if (!(aj < GS_TBL_N)) return false;
    const double* tb = GS_SINCOS_TBL[(int)aj];
    const double bh = rh - jf * (1.0 / GS_TBL_STEP);  // exact (Sterbenz, or j = 0)
    const double sah = jf < 0.0 ? -tb[0] : tb[0], sal = jf < 0.0 ? -tb[1] : tb[1];
    const double cah = tb[2], cal = tb[3];
    // b^2 = q + qe exactly; cos b - 1 = cbh + cbl; sin b - b = sb
    const double q = bh * bh, qe = __builtin_fma(bh, bh, -q);
    const double cbh = -0.5 * q;
    const double cbl = __builtin_fma(-bh, rl, -0.5 * qe) +
                       q * q * (GS_TBL_COS_POLY[0] + q * (GS_TBL_COS_POLY[1] + q * GS_TBL_COS_POLY[2]));
    const double sb = bh * q * (GS_TBL_SIN_POLY[0] + q * (GS_TBL_SIN_POLY[1] + q * GS_TBL_SIN_POLY[2])) +
                      cbh * rl;
    // sin r
    const double p = cah * bh, pe = __builtin_fma(cah, bh, -p);
    const double m = sah * cbh, me = __builtin_fma(sah, cbh, -m);
    const gs_dd s1 = gs_two_sum(sah, p);
    const gs_dd s2 = gs_two_sum(s1.hi, m);
    double lo = s1.lo + s2.lo + sal + pe + me + cah * rl + cal * bh + sal * cbh;
    lo = lo + sah * cbl;
    lo = lo + cah * sb;
    const gs_dd ys = gs_fast_two_sum(s2.hi, lo);
    // cos r
    const double p2 = -sah * bh, pe2 = __builtin_fma(-sah, bh, -p2);
    const double m2 = cah * cbh, me2 = __builtin_fma(cah, cbh, -m2);
    const gs_dd c1 = gs_two_sum(cah, p2);
    const gs_dd c2 = gs_two_sum(c1.hi, m2);
    double lc = c1.lo + c2.lo + cal + pe2 + me2 - sah * rl - sal * bh + cal * cbh;
    lc = lc + cah * cbl;
    lc = lc - sah * sb;
    const gs_dd yc = gs_fast_two_sum(c2.hi, lc);
    out->s = ys.hi;
    out->c = yc.hi;
    return gs_rounds_to(ys.hi, ys.lo, fabs(ys.hi) * 0x1p-66) &&
           gs_rounds_to(yc.hi, yc.lo, fabs(yc.hi) * 0x1p-72);
}

GS_HD __forceinline__ gs_sc gs_sincos_v(double x) {
    if (!(fabs(x) < 823549.6))  // 2^19 * pi, also catches NaN/Inf
        return {sin(x), cos(x)};
    if (x == 0.0)  // keeps the sign of a zero argument
        return {x, 1.0};
    int q;
    const gs_dd r = gs_reduce_pio2(x, &q);
    gs_sc v;
    if (!gs_sincos_table(r, &v)) v = gs_sincos_series(r);
    switch (q) {
        case 0: return {v.s, v.c};
        case 1: return {v.c, -v.s};
        case 2: return {-v.s, -v.c};
        default: return {-v.c, v.s};
    }
}

// The series alone (the path the table path falls back to), for tests.
GS_HD __forceinline__ gs_sc gs_sincos_series_v(double x) {
    if (!(fabs(x) < 823549.6)) return {sin(x), cos(x)};
    if (x == 0.0) return {x, 1.0};
    int q;
    const gs_sc v = gs_sincos_series(gs_reduce_pio2(x, &q));
    switch (q) {
        case 0: return {v.s, v.c};
        case 1: return {v.c, -v.s};
        case 2: return {-v.s, -v.c};
        default: return {-v.c, v.s};
    }
}

GS_HD __forceinline__ void gs_sincos(double x, double* s_out, double* c_out) {
    const gs_sc r = gs_sincos_v(x);
    *s_out = r.s;
    *c_out = r.c;
}

// ---------------------------------------------------------------------------
// Diffraction geometry (src/cdiffraction.c).

// the wedge and chi arguments as the reference forms them: sin(wedge * RAD)
// with RAD = PI / 180.0 unparenthesised is sin((wedge*PI)/180.0)
// (src/cdiffraction.c:7,119-120,130-131)
GS_HD __forceinline__ double gs_wedge_rad(double a) { return a * GS_PI / 180.0; }

// mat = W(wedge).C(chi), formed with the operand order of the matmat macro
// (src/cdiffraction.h:20-29) from sin((wedge*PI)/180.0) etc.
// (src/cdiffraction.c:119-142).
// The matrix from the four sines and cosines.
GS_HD __forceinline__ void gs_wedge_chi_from_trig(double sw, double cw, double sc, double cc,
                                                  double* mat) {
    double w[9] = {cw, 0.0, -sw, 0., 1.0, 0., sw, 0.0, cw};
    double c[9] = {1., 0.0, 0., 0., cc, -sc, 0., sc, cc};
    for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++)
            mat[3 * col + row] = c[row] * w[3 * col + 0] + c[row + 3] * w[3 * col + 1] +
                                 c[row + 6] * w[3 * col + 2];
}

GS_HD __forceinline__ void gs_wedge_chi_matrix(double wedge, double chi, double* mat) {
    double sw, cw, sc, cc;
    gs_sincos(gs_wedge_rad(wedge), &sw, &cw);
    gs_sincos(gs_wedge_rad(chi), &sc, &cc);
    gs_wedge_chi_from_trig(sw, cw, sc, cc, mat);
}

// the omega argument exactly as the reference forms it:
// ((PI/180.0)*omega)*omegasign (src/cdiffraction.c:146-147)
GS_HD __forceinline__ double gs_omega_rad(double omega, double omegasign) {
    return GS_PI / 180.0 * omega * omegasign;
}

// one peak of compute_gv given sin/cos of its omega (src/cdiffraction.c:149-166)
GS_HD __forceinline__ void gs_gv_one(double xl0, double xl1, double xl2, double so, double co,
                                          const double* mat, double ds, double t0, double t1,
                                          double t2, double& g0, double& g1, double& g2) {
    double u0 = co * t0 - so * t1;
    double u1 = so * t0 + co * t1;
    double u2 = t2;
    double o0 = mat[0] * u0 + mat[1] * u1 + mat[2] * u2;
    double o1 = mat[3] * u0 + mat[4] * u1 + mat[5] * u2;
    double o2 = mat[6] * u0 + mat[7] * u1 + mat[8] * u2;
    double d0 = xl0 - o0, d1 = xl1 - o1, d2 = xl2 - o2;
    double modyz = 1. / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double k0 = ds * (d0 * modyz - 1.);
    double k1 = ds * d1 * modyz;
    double k2 = ds * d2 * modyz;
    double v0 = mat[0] * k0 + mat[3] * k1 + mat[6] * k2;
    double v1 = mat[1] * k0 + mat[4] * k1 + mat[7] * k2;
    double v2 = mat[2] * k0 + mat[5] * k1 + mat[8] * k2;
    g0 = co * v0 + so * v1;
    g1 = -so * v0 + co * v1;
    g2 = v2;
}

// one peak of compute_xlylzl (src/cdiffraction.c:213-222)
GS_HD __forceinline__ void gs_xlylzl_one(double s, double f, const double* p, const double* r,
                                              const double* dist, double& x0, double& x1,
                                              double& x2) {
    double v1 = (f - p[1]) * p[3];
    double v2 = (s - p[0]) * p[2];
    x0 = r[1] * v1 + r[2] * v2 + dist[0];
    x1 = r[4] * v1 + r[5] * v2 + dist[1];
    x2 = r[7] * v1 + r[8] * v2 + dist[2];
}

// ---------------------------------------------------------------------------
// Reductions.
__device__ __forceinline__ double gs_wave_sum(double v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, GS_WAVE);
    return v;
}

// A double moved across lanes by one DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double gs_dpp_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// v + v of the partner lane 16 (SPAN 16) or 32 (SPAN 32) lanes away, via
// v_permlane16_swap / v_permlane32_swap: after the swap one register holds
// the lower partner's value and the other the upper partner's, in every
// lane of the pair, so both lanes add the same two operands in the same
// order.
template <int SPAN>
__device__ __forceinline__ double gs_swap_sum(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
    auto l = SPAN == 16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                        : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    auto h = SPAN == 16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                        : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const double lower = __builtin_bit_cast(double, ((unsigned long long)h[0] << 32) | l[0]);
    const double upper = __builtin_bit_cast(double, ((unsigned long long)h[1] << 32) | l[1]);
    return lower + upper;
}

// Sum over aligned groups of WIDTH lanes (8, 16, 32 or 64), all lanes
// active, every lane of a group ending with the bit-identical total: each
// step adds two values already equal within their half-groups (IEEE
// addition commutes).  Pairs: lane ^ 1 and ^ 2 (DPP quad_perm), the other
// quad of an 8 (row_half_mirror), the other 8 of a row (row_mirror), the
// other row of 32 lanes and the other 32 (permlane swaps).  All VALU: no
// LDS round trip per step.
template <int WIDTH>
__device__ __forceinline__ double gs_group_sum(double v) {
    static_assert(WIDTH == 8 || WIDTH == 16 || WIDTH == 32 || WIDTH == 64, "group width");
    v += gs_dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
    v += gs_dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
    v += gs_dpp_f64<0x141>(v);  // row_half_mirror
    if (WIDTH >= 16) v += gs_dpp_f64<0x140>(v);  // row_mirror
    if (WIDTH >= 32) v = gs_swap_sum<16>(v);
    if (WIDTH >= 64) v = gs_swap_sum<32>(v);
    return v;
}

// ---------------------------------------------------------------------------
// gs_solve_ubi for groups of 16 or more lanes that all hold the same R and
// H (the totals of gs_group_sum), bit for bit the same results.  Written
// per lane, every lane of the wave issues all 15 divisions of the two
// inverses (each some ten instructions, five of them slow); here lane i
// of each row of 16 (i < 9) divides element i only, and row_newbcast DPP
// moves hand the nine quotients to the whole row.  lr: the lane's index
// within its row (lane & 15).

// v of lane J of the lane's row of 16
template <int J>
__device__ __forceinline__ double gs_row_bcast(double v) {
    return gs_dpp_f64<0x150 + J>(v);  // row_newbcast:J
}

// q[i] = num[i] / den for i < 9 in every lane of the row (den the same in
// all lanes of the row), each quotient IEEE division.
__device__ __forceinline__ void gs_div9_row(const double* num, double den, double* q, int lr) {
    // num[lr] (num[8] for lr >= 8) by a tree of selects on the bits of lr
    // (a chain of compares is turned into an indexed load from memory)
    const bool b0 = lr & 1, b1 = lr & 2, b2 = lr & 4, b3 = lr & 8;
    const double s01 = b0 ? num[1] : num[0], s23 = b0 ? num[3] : num[2];
    const double s45 = b0 ? num[5] : num[4], s67 = b0 ? num[7] : num[6];
    const double s03 = b1 ? s23 : s01, s47 = b1 ? s67 : s45;
    const double s07 = b2 ? s47 : s03;
    const double x = b3 ? num[8] : s07;
    const double r = x / den;
    q[0] = gs_row_bcast<0>(r); q[1] = gs_row_bcast<1>(r); q[2] = gs_row_bcast<2>(r);
    q[3] = gs_row_bcast<3>(r); q[4] = gs_row_bcast<4>(r); q[5] = gs_row_bcast<5>(r);
    q[6] = gs_row_bcast<6>(r); q[7] = gs_row_bcast<7>(r); q[8] = gs_row_bcast<8>(r);
}

// gs_inverse3x3 with the divisions shared out over the row: the same
// cofactors, determinant and quotients.
__device__ __forceinline__ int gs_inverse3x3_row(double* A, int lr) {
    const double c0 = A[8] * A[4] - A[7] * A[5];
    const double c1 = A[8] * A[1] - A[7] * A[2];
    const double c2 = A[5] * A[1] - A[4] * A[2];
    const double det = A[0] * c0 - A[3] * c1 + A[6] * c2;
    if (!(det != 0.)) return -1;
    const double num[9] = {c0,
                           -c1,
                           c2,
                           -(A[8] * A[3] - A[6] * A[5]),
                           (A[8] * A[0] - A[6] * A[2]),
                           -(A[5] * A[0] - A[3] * A[2]),
                           (A[7] * A[3] - A[6] * A[4]),
                           -(A[7] * A[0] - A[6] * A[1]),
                           (A[4] * A[0] - A[3] * A[1])};
    gs_div9_row(num, det, A, lr);
    return 0;
}

__device__ __forceinline__ void gs_solve_ubi_row(double* ubi, const double* R, double* H, int lr) {
    double UB[9];
    for (int i = 0; i < 9; i++) UB[i] = 0.;
    const int k = gs_inverse3x3_row(H, lr);
    if (k == 0) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++) UB[3 * i + j] = UB[3 * i + j] + R[3 * i + l] * H[3 * l + j];
    }
    if (k == 0 && gs_inverse3x3_row(UB, lr) == 0)
        for (int i = 0; i < 9; i++) ubi[i] = UB[i];
}

__device__ __forceinline__ int gs_wave_sum_int(int v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, GS_WAVE);
    return v;
}

// Map a dispatch-order block id to a logical tile id so that consecutive
// logical tiles land on the same XCD (blocks are dispatched round-robin
// over the 8 XCDs of the MI355X).  A bijection on [0, nb).
GS_HD __forceinline__ int gs_xcd_remap(int id, int nb) {
    const int nx = 8;
    int q = nb / nx, r = nb % nx;
    int x = id % nx, slot = id / nx;
    return x * q + (x < r ? x : r) + slot;
}
