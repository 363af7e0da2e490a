// compute_gv for each grain's translation fused with score_and_assign, for
// every grain in order: the loop of refinegrains.assignlabels
// (ImageD11/refinegrains.py:718-734), one peak per lane, the (drlv2,
// label) state of a peak carried through all grains in registers.
//
// Fast path.  With M_k = Omega(omega_k) . mat^T, the reference's g-vector
// (src/cdiffraction.c:149-166) is, in exact arithmetic,
//     G = ds (a_k - t) / |a_k - t| - c_k,
//     a_k = M_k . xl_k,  c_k = ds M_k . (1, 0, 0),
// because d = xl - mat.R(omega).t and M_k.mat.R(omega) = Omega.R = I while
// M_k is orthogonal.  a_k, c_k are per-peak constants, so a pair costs one
// subtraction, a norm and a scaling before h = UBI.G.
//
// Certification.  The reference's G is within 90 u ds (u = 2^-53, ds =
// 1/wavelength) of the exact G in every component when |t| <= |xl|/2 (so
// a - t and xl - o cannot cancel): k0 = ds (d0/|d| - 1) is the only
// subtraction of like quantities and |d0/|d|| <= 1, and everything else is
// a few rounded products of magnitude <= 2 ds.  The fast G is within
// ~290 u ds: the same terms plus its 1/|a - t| (gs_rsqrt_cert<1>, within
// 1.5 2^-46 + 3u ~ 200u).  h = UBI.G then moves by
// at most ||u_i||_1 (290 u ds) + 6.7e-16 (A_i + 1/2) per row (the fast path
// uses the fract form of the score kernel, drlv2 - 3/4, whose rows carry
// an extra 1/2), and drlv2 by the sum of that over rows (1-Lipschitz, see
// k_score_many_fast) plus 5e-16 from each sum of squares.  E = U1 (1e-13
// ds + 1e-15 Gmax) + 3e-15 + 2e-15 tol^2 bounds it with margin (Gmax = 2
// ds bounds |G|; 290 u = 3.2e-14).  A pair whose fast drlv2 - 3/4 is >= tol^2 - 3/4 + E
// certainly fails `drlv2 < tol^2`, so its only effect is the
// reference's else branch (label == this label -> -1), which needs no
// arithmetic.  The same holds, with the same E, against the peak's current
// drlv2 d < tol^2: fast drlv2 - 3/4 >= d - 3/4 + E certainly fails
// `drlv2 < d`, so the exact route is taken only below min(tol^2, d) - 3/4
// + E (the values there are no larger than at tol^2, so E still bounds
// their rounding).  Every other pair is recomputed with the reference's own
// compute_gv and score arithmetic (gs_gv_one, gs_drlv2) and the reference's
// rule is applied to that exact value, so labels and stored drlv2 are
// those of the reference loop.  A NaN or infinite fast drlv2 is never
// ">= tol^2 + E", so non-finite inputs take the exact route; a grain with
// |t| > |xl|/2 for some peak of the wave takes it for the whole wave.
#pragma once
#include "gs_kernels.h"

// per-peak constants of the fast path: a = Omega(omega) mat^T xl and
// c = ds Omega(omega) mat^T (1,0,0)
GS_HD __forceinline__ void gs_scan_consts(double x0, double x1, double x2, double so, double co,
                                          const double* mat, double ds, double* a, double* c) {
    const double w0 = mat[0] * x0 + mat[3] * x1 + mat[6] * x2;
    const double w1 = mat[1] * x0 + mat[4] * x1 + mat[7] * x2;
    const double w2 = mat[2] * x0 + mat[5] * x1 + mat[8] * x2;
    a[0] = co * w0 + so * w1;
    a[1] = -so * w0 + co * w1;
    a[2] = w2;
    c[0] = ds * (co * mat[0] + so * mat[1]);
    c[1] = ds * (-so * mat[0] + co * mat[1]);
    c[2] = ds * mat[2];
}

// 1/sqrt(x) to within a few ulp: the hardware estimate (v_rsq_f64, 2^29
// ulp = 2^-23 relative by the ISA's statement; measured at most 2^-24.2
// over 4.2e6 arguments spread over 200 binades) and STEPS Newton steps
// y <- y + y/2 (1 - x y^2).  From a start with relative error e <= 2^-14
// each step gives 1.5 e^2 + ~3u, so two steps leave ~6u however coarse
// the estimate, and one step from the hardware's leaves 1.5 2^-46 + 3u
// (~200u).  x = 0 or non-finite gives NaN, which the callers send down
// their exact routes.  On the host the estimate is the float rsqrt of the
// value (a 24-bit start, e <= 2^-23 as well), so the host self-test
// covers the same bounds.
template <int STEPS = 2>
GS_HD __forceinline__ double gs_rsqrt_cert(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rsq(x);
#else
    double y = (double)(1.0f / sqrtf((float)x));
    if (!(x > 0.0) || x > 1e300 || x < 1e-300) y = __builtin_nan("");
#endif
    for (int it = 0; it < STEPS; it++) {
        const double r = __builtin_fma(-(x * y), y, 1.0);
        y = __builtin_fma(0.5 * y, r, y);
    }
    return y;
}

// fast drlv2 - 3/4 of one pair: G = ds (a - t)/|a - t| - c, then the fract
// form of the score kernel (gs_drlv2_fract)
GS_HD __forceinline__ double gs_scan_fast_drlv2(const gs_ubi& u, const double* a,
                                                const double* c, double t0, double t1, double t2,
                                                double ds) {
    const double e0 = a[0] - t0, e1 = a[1] - t1, e2 = a[2] - t2;
    const double inv =
        ds * gs_rsqrt_cert<1>(__builtin_fma(e2, e2, __builtin_fma(e1, e1, e0 * e0)));
    const double G0 = __builtin_fma(e0, inv, -c[0]);
    const double G1 = __builtin_fma(e1, inv, -c[1]);
    const double G2 = __builtin_fma(e2, inv, -c[2]);
    return gs_drlv2_fract(u, G0, G1, G2);
}

// half-width of the undecided window of gs_scan_fast_drlv2 around
// tol^2 - 3/4 for a grain (UBI u, |t|^2 = tt) and peaks with |xl|^2 >=
// n2min; +inf where the bound does not apply.  The G term covers the fast
// G (290 u ds per component, inverse square root included); the
// fract form adds the per-row 1/2 and its five roundings (3e-15, see
// gs_cert_window_fract).
GS_HD __forceinline__ double gs_scan_window(const gs_ubi& u, double ds, double tolsq, double tt,
                                            double n2min) {
    const double gmax = 2.0 * fabs(ds);
    const double u1 = fabs(u.u00) + fabs(u.u01) + fabs(u.u02) + fabs(u.u10) + fabs(u.u11) +
                      fabs(u.u12) + fabs(u.u20) + fabs(u.u21) + fabs(u.u22);
    if (!(u1 * gmax < 562949953421312.0) || !(4.0 * tt <= n2min)) return __builtin_inf();
    return u1 * (1e-13 * fabs(ds) + 1e-15 * gmax) + 3e-15 + 2e-15 * tolsq;
}

#define GS_SCAN_CHUNK 512  // grains per LDS tile of k_gv_assign_many

template <int P>
__global__ __launch_bounds__(GS_BLOCK) void k_gv_assign_many(
    const double* __restrict__ xl, const double* __restrict__ so_in,
    const double* __restrict__ co_in, int ng, gs_goniometer gp,
    const double* __restrict__ ubis, const double* __restrict__ trans,
    const int* __restrict__ glabels, int K, double tolsq, double* __restrict__ drlv2,
    int* __restrict__ labels, int* __restrict__ counts) {
    __shared__ int lcnt[GS_SCAN_CHUNK];
    __shared__ double lgr[4 * GS_SCAN_CHUNK];  // t0, t1, t2, hi per grain of the chunk
    __shared__ double ln2[GS_BLOCK / GS_WAVE];
    __shared__ double smat[9];
    __shared__ double sds;
    gs_block_setup(gp, smat, &sds);
    double mat[9];
    for (int j = 0; j < 9; j++) mat[j] = smat[j];
    const double ds = sds;

    // per-peak constants of the fast path and the (drlv2, label) state;
    // thm = min(tol^2, drlv2) - 3/4: a pair can change the state through
    // the reference arithmetic only if its drlv2 is below both (else it
    // takes the update's else branch, which needs no arithmetic)
    double ax[P], ay[P], az[P], cx[P], cy[P], cz[P], d[P], thm[P];
    int lab[P];
    double n2min = __builtin_inf();
    const long long base = (long long)blockIdx.x * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const long long k = base + (long long)p * GS_BLOCK;
        ax[p] = ay[p] = az[p] = cx[p] = cy[p] = cz[p] = 0.0;
        d[p] = 0.0;
        thm[p] = tolsq - 0.75;
        lab[p] = -1;
        if (k < ng) {
            const double x0 = xl[3 * k], x1 = xl[3 * k + 1], x2 = xl[3 * k + 2];
            double a[3], c[3];
            gs_scan_consts(x0, x1, x2, so_in[k], co_in[k], mat, ds, a, c);
            ax[p] = a[0];
            ay[p] = a[1];
            az[p] = a[2];
            cx[p] = c[0];
            cy[p] = c[1];
            cz[p] = c[2];
            d[p] = drlv2[k];
            thm[p] = fmin(tolsq, d[p]) - 0.75;  // NaN drlv2: tol^2 (never below it)
            lab[p] = labels[k];
            const double n2 = x0 * x0 + x1 * x1 + x2 * x2;
            n2min = n2 == n2 ? fmin(n2min, n2) : n2min;
        }
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    for (int o = GS_WAVE / 2; o > 0; o >>= 1) n2min = fmin(n2min, __shfl_xor(n2min, o, GS_WAVE));
    if (lane == 0) ln2[threadIdx.x / GS_WAVE] = n2min;
    __syncthreads();
    for (int w = 0; w < GS_BLOCK / GS_WAVE; w++) n2min = fmin(n2min, ln2[w]);
    for (int c0 = 0; c0 < K; c0 += GS_SCAN_CHUNK) {
        const int c1 = min(K, c0 + GS_SCAN_CHUNK);
        // per grain of the chunk, once per block: translation and window
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) {
            const size_t g = (size_t)(c0 + i);
            const double t0 = trans[3 * g], t1 = trans[3 * g + 1], t2 = trans[3 * g + 2];
            const double tt = t0 * t0 + t1 * t1 + t2 * t2;
            lgr[4 * i] = t0;
            lgr[4 * i + 1] = t1;
            lgr[4 * i + 2] = t2;
            lgr[4 * i + 3] = gs_scan_window(gs_load_ubi(ubis + 9 * g), ds, tolsq, tt, n2min);
            lcnt[i] = 0;
        }
        __syncthreads();
        for (int g = c0; g < c1; g++) {
            const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)g);
            const int L = glabels[g];
            const double t0 = lgr[4 * (g - c0)], t1 = lgr[4 * (g - c0) + 1],
                         t2 = lgr[4 * (g - c0) + 2], win = lgr[4 * (g - c0) + 3];
#pragma unroll
            for (int p = 0; p < P; p++) {
                const double a[3] = {ax[p], ay[p], az[p]}, c[3] = {cx[p], cy[p], cz[p]};
                const double sf = gs_scan_fast_drlv2(u, a, c, t0, t1, t2, ds);
                // candidates (and non-finite values) take the exact route:
                // the pairs whose drlv2 may be below both tol^2 and the
                // peak's current drlv2 (sf is drlv2 - 3/4 within win)
                const bool cand = !(sf >= thm[p] + win);
                // a certain failure only matters to a peak holding this
                // grain's label (else branch of the update: label -> -1);
                // such peaks are rare, so one compare per pair decides
                // wave-wide whether any lane needs the update
                // (the empty asm statements keep the compiler from turning
                // this branch into selects and from hoisting the exact
                // route's products with t into the common path)
                if (__ballot(lab[p] == L)) {
                    asm volatile("");
                    if (!cand && lab[p] == L) lab[p] = -1;
                }
                if (cand) {  // rare: the reference arithmetic decides
                    const long long k = base + (long long)p * GS_BLOCK;
                    if (k < ng) {
                        double e0 = t0, e1 = t1, e2 = t2, g0, g1, g2;
                        asm volatile("" : "+v"(e0), "+v"(e1), "+v"(e2));
                        gs_gv_one(xl[3 * k], xl[3 * k + 1], xl[3 * k + 2], so_in[k], co_in[k],
                                  smat, ds, e0, e1, e2, g0, g1, g2);
                        const double se = gs_drlv2(u, g0, g1, g2);
                        if ((se < tolsq) && (se < d[p])) {
                            lab[p] = L;
                            d[p] = se;
                            thm[p] = se - 0.75;
                            atomicAdd(&lcnt[g - c0], 1);
                        } else if (lab[p] == L) {
                            lab[p] = -1;
                        }
                    }
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) {
            const int c = lcnt[i];
            if (c) atomicAdd(&counts[c0 + i], c);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        const long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            drlv2[k] = d[p];
            labels[k] = lab[p];
        }
    }
}

// g-vector of every CSR list entry with its grain's translation
// (compute_gv per grain, gathered): out[q] = G(peak idx[q], t_owner[q]).
// Used to refine grains whose g-vectors depend on their position.
__global__ __launch_bounds__(GS_BLOCK) void k_gv_lists(
    const double* __restrict__ xl, const double* __restrict__ so_in,
    const double* __restrict__ co_in, int ng, gs_goniometer gp, const double* __restrict__ trans,
    int K, const int* __restrict__ idx, const int* __restrict__ owner, long long n_idx,
    double* __restrict__ out) {
    __shared__ double smat[9];
    __shared__ double sds;
    gs_block_setup(gp, smat, &sds);
    const long long q = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (q >= n_idx) return;
    const int k = idx[q], g = owner[q];
    double g0 = __builtin_nan(""), g1 = g0, g2 = g0;
    if (k >= 0 && k < ng && g >= 0 && g < K) {
        double mat[9];
        for (int j = 0; j < 9; j++) mat[j] = smat[j];
        gs_gv_one(xl[3 * (size_t)k], xl[3 * (size_t)k + 1], xl[3 * (size_t)k + 2], so_in[k],
                  co_in[k], mat, sds, trans[3 * (size_t)g], trans[3 * (size_t)g + 1],
                  trans[3 * (size_t)g + 2], g0, g1, g2);
    }
    out[3 * q] = g0;
    out[3 * q + 1] = g1;
    out[3 * q + 2] = g2;
}
