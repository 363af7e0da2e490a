// Adjoint of the joint refinement objective (no counterpart in the
// reference, which fits geometry by Nelder-Mead on refinegrains.gof,
// ImageD11/refinegrains.py:517-624).
//
//   L = sum_g sum_{q in list g} || UBI_g . G_q - hkl_q ||^2
//   G_q = compute_gv(compute_xlylzl(s_k, f_k; p, r, dist), omega_k; wvln,
//                    mat, t_g),  k = idx[q]
//
// with hkl_q fixed integers (assignments held fixed) and the per-peak
// sin/cos of omega precomputed (omega is data, not a parameter).  The
// packed geometry the kernels differentiate is
//   geo[0..3]  p = z_center, y_center, z_size, y_size   (compute_xlylzl)
//   geo[4..12] r = detector rotation . flip, row-major
//   geo[13..15] dist
//   geo[16]    wavelength
//   geo[17..25] mat = W(wedge) . C(chi), row-major
// and the physical parameters (distance, centres, tilts, wavelength,
// wedge, chi) are mapped onto it by differentiable host-side torch code.
//
// gspace = true switches to the g-space objective E = 1/2 sum ||UB.hkl -
// G||^2 with UB = UBI^-1, whose gradient in UB is UB.H - R with the R and H
// sums of score_and_refine (src/closest.c:310-316).
#pragma once
#include "gs_kernels.h"

#define GS_GEO 26
#define GS_GEO_OUT 27  // the 26 geometry partials + the loss
#define GS_NSUMS 15    // sum h G^T (9) and sum G G^T (6) for Gauss-Newton

struct gs_pair_fwd {
    double v1, v2;        // detector-plane coordinates
    double xl[3];         // lab position
    double u[3];          // omega-rotated translation
    double d[3], dn;      // d = xl - mat.u and |d|
    double nh[3];         // d / |d|
    double k[3];          // scattering vector ds*(nh - x)
    double v[3];          // mat^T k
    double G[3];          // g-vector
};

// the forward chain of compute_xlylzl + compute_gv for one pair
__device__ __forceinline__ void gs_joint_forward(double s, double f, double so, double co,
                                                 const double* geo, const double* t,
                                                 gs_pair_fwd& w) {
    const double* p = geo;
    const double* r = geo + 4;
    const double* dist = geo + 13;
    const double* mat = geo + 17;
    const double ds = 1.0 / geo[16];
    w.v1 = (f - p[1]) * p[3];
    w.v2 = (s - p[0]) * p[2];
    for (int j = 0; j < 3; j++) w.xl[j] = r[3 * j + 1] * w.v1 + r[3 * j + 2] * w.v2 + dist[j];
    w.u[0] = co * t[0] - so * t[1];
    w.u[1] = so * t[0] + co * t[1];
    w.u[2] = t[2];
    for (int i = 0; i < 3; i++)
        w.d[i] = w.xl[i] - (mat[3 * i] * w.u[0] + mat[3 * i + 1] * w.u[1] + mat[3 * i + 2] * w.u[2]);
    w.dn = sqrt(w.d[0] * w.d[0] + w.d[1] * w.d[1] + w.d[2] * w.d[2]);
    for (int i = 0; i < 3; i++) w.nh[i] = w.d[i] / w.dn;
    w.k[0] = ds * (w.nh[0] - 1.0);
    w.k[1] = ds * w.nh[1];
    w.k[2] = ds * w.nh[2];
    for (int j = 0; j < 3; j++) w.v[j] = mat[j] * w.k[0] + mat[3 + j] * w.k[1] + mat[6 + j] * w.k[2];
    w.G[0] = co * w.v[0] + so * w.v[1];
    w.G[1] = -so * w.v[0] + co * w.v[1];
    w.G[2] = w.v[2];
}

// Reverse sweep: given dL/dG, accumulate dL/dgeo (26), dL/dt (3).
__device__ __forceinline__ void gs_joint_backward(double s, double f, double so, double co,
                                                  const double* geo, const gs_pair_fwd& w,
                                                  const double* gG, double* ggeo, double* gt) {
    const double* p = geo;
    const double* r = geo + 4;
    const double* mat = geo + 17;
    const double ds = 1.0 / geo[16];
    // G = Omega(omega) v
    double gv[3] = {co * gG[0] - so * gG[1], so * gG[0] + co * gG[1], gG[2]};
    // v = mat^T k
    double gk[3];
    for (int i = 0; i < 3; i++) {
        gk[i] = mat[3 * i] * gv[0] + mat[3 * i + 1] * gv[1] + mat[3 * i + 2] * gv[2];
        for (int j = 0; j < 3; j++) ggeo[17 + 3 * i + j] += w.k[i] * gv[j];
    }
    // k = ds (nh - x): d/dds, then d/dwavelength = -dds / wvln^2 at the end
    const double gds = gk[0] * (w.nh[0] - 1.0) + gk[1] * w.nh[1] + gk[2] * w.nh[2];
    ggeo[16] += gds;
    double gn[3] = {ds * gk[0], ds * gk[1], ds * gk[2]};
    // nh = d / |d|
    const double dot = w.nh[0] * gn[0] + w.nh[1] * gn[1] + w.nh[2] * gn[2];
    double gd[3];
    for (int i = 0; i < 3; i++) gd[i] = (gn[i] - w.nh[i] * dot) / w.dn;
    // d = xl - mat.u
    double gu[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            ggeo[17 + 3 * i + j] -= gd[i] * w.u[j];
            gu[j] -= mat[3 * i + j] * gd[i];
        }
    gt[0] += co * gu[0] + so * gu[1];
    gt[1] += -so * gu[0] + co * gu[1];
    gt[2] += gu[2];
    // xl_j = r[3j+1] v1 + r[3j+2] v2 + dist_j
    double gv1 = 0.0, gv2 = 0.0;
    for (int j = 0; j < 3; j++) {
        ggeo[4 + 3 * j + 1] += gd[j] * w.v1;
        ggeo[4 + 3 * j + 2] += gd[j] * w.v2;
        ggeo[13 + j] += gd[j];
        gv1 += r[3 * j + 1] * gd[j];
        gv2 += r[3 * j + 2] * gd[j];
    }
    // v1 = (f - p1) p3, v2 = (s - p0) p2
    ggeo[1] -= p[3] * gv1;
    ggeo[3] += (f - p[1]) * gv1;
    ggeo[0] -= p[2] * gv2;
    ggeo[2] += (s - p[0]) * gv2;
}

// Fixed integer hkl of every list entry at the current state:
// hkl[q] = round(UBI_g . G_q).  One lane per list entry.
template <bool ROWS>
__global__ __launch_bounds__(GS_BLOCK) void k_joint_hkl(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const double* __restrict__ geo_in,
    const double* __restrict__ ubis, const double* __restrict__ trans, int K,
    const int* __restrict__ idx, const int* __restrict__ owner, long long n_idx,
    int* __restrict__ hkl) {
    __shared__ double geo[GS_GEO];
    if (threadIdx.x < GS_GEO) geo[threadIdx.x] = geo_in[threadIdx.x];
    __syncthreads();
    const long long q = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (q >= n_idx) return;
    const long long k = ROWS ? q : (long long)idx[q];
    const int g = owner[q];
    if (k < 0 || k >= npk || g < 0 || g >= K) {
        hkl[3 * q] = hkl[3 * q + 1] = hkl[3 * q + 2] = 0;
        return;
    }
    double gg[GS_GEO];
    for (int i = 0; i < GS_GEO; i++) gg[i] = geo[i];
    gs_pair_fwd w;
    gs_joint_forward(sc[k], fc[k], so[k], co[k], gg, trans + 3 * (size_t)g, w);
    const double* u = ubis + 9 * (size_t)g;
    for (int i = 0; i < 3; i++) {
        double h = u[3 * i] * w.G[0] + u[3 * i + 1] * w.G[1] + u[3 * i + 2] * w.G[2];
        hkl[3 * q + i] = (int)rint(h);
    }
}

// Per-grain forward + reverse sweep.  LPG lanes per grain, grains from a
// work queue; every per-grain output is the gs_group_sum over that grain's
// lanes, so it does not depend on which wave took the grain.
// Outputs per grain g: loss_g[g], grad_ubi[9g..] (dL/dUBI, or dE/dUB in
// g-space), grad_t[3g..], geo_part[26g..] (dL/dgeo restricted to grain g;
// slot 16 holds dL/d(1/wavelength)), and with NSUMS the Gauss-Newton sums
// sum h G^T (9) and sum G G^T (6: 00 01 02 11 12 22).
// ROWS: no index array; sc/fc/so/co are already in list order (entry q is
// element q), so the pass streams them instead of gathering four cache
// lines per entry.
template <int LPG, bool GSPACE, bool NSUMS, bool ROWS>
__global__ __launch_bounds__(GS_BLOCK) void k_joint_adjoint(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const double* __restrict__ geo_in,
    const double* __restrict__ ubis, const double* __restrict__ trans, int K,
    const int* __restrict__ idx, long long n_idx, const long long* __restrict__ off,
    const int* __restrict__ hkl, const int* __restrict__ order, double* __restrict__ loss_g,
    double* __restrict__ grad_ubi, double* __restrict__ grad_t, double* __restrict__ geo_part,
    double* __restrict__ nsums, int* __restrict__ queue) {
    __shared__ double sgeo[GS_GEO];
    if (threadIdx.x < GS_GEO) sgeo[threadIdx.x] = geo_in[threadIdx.x];
    __syncthreads();
    double geo[GS_GEO];
    for (int i = 0; i < GS_GEO; i++) geo[i] = sgeo[i];
    constexpr int GPW = GS_WAVE / LPG;
    const int lane = threadIdx.x & (GS_WAVE - 1);
    const int sub = lane / LPG;
    const int sl = lane - sub * LPG;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(queue, GPW);
        base = __shfl(base, 0, GS_WAVE);
        if (base >= K) break;
        const int slot = base + sub;
        const int gq = slot < K ? order[slot] : -1;
        const bool active = gq >= 0 && gq < K;
        const int g = active ? gq : 0;
        long long o1 = active ? off[g + 1] : 0;
        o1 = o1 < 0 ? 0 : (o1 > n_idx ? n_idx : o1);
        long long o0 = active ? off[g] : 0;
        o0 = o0 < 0 ? 0 : (o0 > o1 ? o1 : o0);
        double U[9], t[3];
        for (int i = 0; i < 9; i++) U[i] = active ? ubis[9 * (size_t)g + i] : (i % 4 == 0 ? 1.0 : 0.0);
        for (int i = 0; i < 3; i++) t[i] = active ? trans[3 * (size_t)g + i] : 0.0;
        if (GSPACE) {  // UB = UBI^-1 (identity stays identity for inactive lanes)
            if (gs_inverse3x3(U) != 0)
                for (int i = 0; i < 9; i++) U[i] = __builtin_nan("");
        }
        double acc_l = 0.0, acc_u[9], acc_t[3], acc_geo[GS_GEO], acc_n[GS_NSUMS];
        for (int i = 0; i < 9; i++) acc_u[i] = 0.0;
        for (int i = 0; i < 3; i++) acc_t[i] = 0.0;
        for (int i = 0; i < GS_GEO; i++) acc_geo[i] = 0.0;
        for (int i = 0; i < GS_NSUMS; i++) acc_n[i] = 0.0;
        for (long long q = o0 + sl; q < o1; q += LPG) {
            const long long k = ROWS ? q : (long long)idx[q];
            if (k < 0 || k >= npk) continue;
            const double h0 = hkl[3 * q], h1 = hkl[3 * q + 1], h2 = hkl[3 * q + 2];
            gs_pair_fwd w;
            gs_joint_forward(sc[k], fc[k], so[k], co[k], geo, t, w);
            double gG[3];
            if (!GSPACE) {
                double e[3];
                e[0] = U[0] * w.G[0] + U[1] * w.G[1] + U[2] * w.G[2] - h0;
                e[1] = U[3] * w.G[0] + U[4] * w.G[1] + U[5] * w.G[2] - h1;
                e[2] = U[6] * w.G[0] + U[7] * w.G[1] + U[8] * w.G[2] - h2;
                acc_l += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) acc_u[3 * i + j] += 2.0 * e[i] * w.G[j];
                for (int j = 0; j < 3; j++)
                    gG[j] = 2.0 * (U[j] * e[0] + U[3 + j] * e[1] + U[6 + j] * e[2]);
            } else {
                // E = 1/2 ||UB.h - G||^2: dE/dUB = (UB.h - G) h^T, dE/dG = -(UB.h - G)
                const double hh[3] = {h0, h1, h2};
                double e[3];
                for (int i = 0; i < 3; i++)
                    e[i] = U[3 * i] * h0 + U[3 * i + 1] * h1 + U[3 * i + 2] * h2 - w.G[i];
                acc_l += 0.5 * (e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) acc_u[3 * i + j] += e[i] * hh[j];
                for (int j = 0; j < 3; j++) gG[j] = -e[j];
            }
            if (NSUMS) {
                const double hh[3] = {h0, h1, h2};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) acc_n[3 * i + j] += hh[i] * w.G[j];
                acc_n[9] += w.G[0] * w.G[0];
                acc_n[10] += w.G[0] * w.G[1];
                acc_n[11] += w.G[0] * w.G[2];
                acc_n[12] += w.G[1] * w.G[1];
                acc_n[13] += w.G[1] * w.G[2];
                acc_n[14] += w.G[2] * w.G[2];
            }
            gs_joint_backward(sc[k], fc[k], so[k], co[k], geo, w, gG, acc_geo, acc_t);
        }
        acc_l = gs_group_sum<LPG>(acc_l);
        for (int i = 0; i < 9; i++) acc_u[i] = gs_group_sum<LPG>(acc_u[i]);
        for (int i = 0; i < 3; i++) acc_t[i] = gs_group_sum<LPG>(acc_t[i]);
        for (int i = 0; i < GS_GEO; i++) acc_geo[i] = gs_group_sum<LPG>(acc_geo[i]);
        if (NSUMS)
            for (int i = 0; i < GS_NSUMS; i++) acc_n[i] = gs_group_sum<LPG>(acc_n[i]);
        if (active && sl == 0) {
            loss_g[g] = acc_l;
            for (int i = 0; i < 9; i++) grad_ubi[9 * (size_t)g + i] = acc_u[i];
            for (int i = 0; i < 3; i++) grad_t[3 * (size_t)g + i] = acc_t[i];
            for (int i = 0; i < GS_GEO; i++) geo_part[GS_GEO * (size_t)g + i] = acc_geo[i];
            if (NSUMS)
                for (int i = 0; i < GS_NSUMS; i++) nsums[GS_NSUMS * (size_t)g + i] = acc_n[i];
        }
    }
}

// Deterministic reduction of the per-grain partials: out[0..25] = dL/dgeo
// (slot 16 converted from d/d(1/wavelength) to d/dwavelength), out[26] =
// loss.  One block of GS_FIN_GROUPS x 32 threads: thread (r, j) sums
// column j over the grains g = r, r + GS_FIN_GROUPS, ... in that order,
// then thread j adds the GS_FIN_GROUPS partial sums in group order.  The
// order depends only on K (not on which waves took which grains), so the
// result is the same on every run.
#define GS_FIN_GROUPS 32
__global__ __launch_bounds__(GS_FIN_GROUPS * 32) void k_joint_finalize(
    const double* __restrict__ geo_part, const double* __restrict__ loss_g, int K,
    const double* __restrict__ geo_in, double* __restrict__ out) {
    __shared__ double part[GS_FIN_GROUPS][GS_GEO_OUT];
    const int j = threadIdx.x & 31;
    const int r = threadIdx.x >> 5;
    if (j < GS_GEO_OUT) {
        double s = 0.0;
        int g = r;
        // loads run ahead in groups of four; additions stay in order
        for (; g + 3 * GS_FIN_GROUPS < K; g += 4 * GS_FIN_GROUPS) {
            double v[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const size_t gg = (size_t)g + (size_t)q * GS_FIN_GROUPS;
                v[q] = j < GS_GEO ? geo_part[GS_GEO * gg + j] : loss_g[gg];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) s += v[q];
        }
        for (; g < K; g += GS_FIN_GROUPS)
            s += j < GS_GEO ? geo_part[GS_GEO * (size_t)g + j] : loss_g[g];
        part[r][j] = s;
    }
    __syncthreads();
    if (r == 0 && j < GS_GEO_OUT) {
        double s = 0.0;
        for (int q = 0; q < GS_FIN_GROUPS; q++) s += part[q][j];
        if (j == 16) s = -s / (geo_in[16] * geo_in[16]);
        out[j] = s;
    }
}
