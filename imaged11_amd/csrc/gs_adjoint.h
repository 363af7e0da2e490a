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
#include "gs_assign_gv.h"  // gs_rsqrt_cert

#define GS_GEO 26
#define GS_GEO_OUT 27  // the 26 geometry partials + the loss
#define GS_NSUMS 15    // sum h G^T (9) and sum G G^T (6) for Gauss-Newton

// The per-pair forward values the reverse sweep needs.
struct gs_pair_fwd {
    double a1, a2;   // f - p1, s - p0 (detector-plane offsets before scaling)
    double u[3];     // omega-rotated translation
    double nh[3];    // d / |d|, d = xl - mat.u
    double rn;       // 1 / |d|
    double G[3];     // g-vector
};

// The forward chain of compute_xlylzl + compute_gv for one pair, in the
// adjoint's own arithmetic: FMA-contracted, 1/|d| from the hardware
// reciprocal square root and two Newton steps (within a few ulps), so G
// is within a few ulps of the reference's compute_gv (the value and
// gradient, not the bits, are what the adjoint needs; the fixed hkl come
// from k_joint_hkl, which keeps the reference's arithmetic).  geo is read
// at each use through a uniform pointer, so the 26 numbers live in scalar
// registers instead of being copied into every lane.
__device__ __forceinline__ void gs_joint_forward(double s, double f, double so, double co,
                                                 const double* __restrict__ geo, double ds,
                                                 double t0, double t1, double t2,
                                                 gs_pair_fwd& w) {
#pragma clang fp contract(fast)
    const double* p = geo;
    const double* r = geo + 4;
    const double* dist = geo + 13;
    const double* mat = geo + 17;
    w.a1 = f - p[1];
    w.a2 = s - p[0];
    const double v1 = w.a1 * p[3], v2 = w.a2 * p[2];
    w.u[0] = co * t0 - so * t1;
    w.u[1] = so * t0 + co * t1;
    w.u[2] = t2;
    double d[3];
    for (int i = 0; i < 3; i++)
        d[i] = (r[3 * i + 1] * v1 + r[3 * i + 2] * v2 + dist[i]) -
               (mat[3 * i] * w.u[0] + mat[3 * i + 1] * w.u[1] + mat[3 * i + 2] * w.u[2]);
    w.rn = gs_rsqrt_cert(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; i++) w.nh[i] = d[i] * w.rn;
    const double k0 = ds * (w.nh[0] - 1.0), k1 = ds * w.nh[1], k2 = ds * w.nh[2];
    double v[3];
    for (int j = 0; j < 3; j++) v[j] = mat[j] * k0 + mat[3 + j] * k1 + mat[6 + j] * k2;
    w.G[0] = co * v[0] + so * v[1];
    w.G[1] = -so * v[0] + co * v[1];
    w.G[2] = v[2];
}

// Per-lane accumulators of one grain's geometry gradient.  The partials
// of p (4) and of the r entries (6; r[3j] does not enter compute_xlylzl)
// are linear in three sums over the grain's pairs, A_j = sum (f - p1) gd_j,
// B_j = sum (s - p0) gd_j and D_j = sum gd_j (gd = dL/dd), so 19 numbers
// are accumulated per lane instead of 26 (gs_geo_expand forms the 26).
struct gs_geo_acc {
    double m[9];  // dL/dmat
    double ids;   // dL/d(1/wavelength)
    double A[3], B[3], D[3];
};

// Reverse sweep of one pair: given dL/dG, accumulate the geometry sums
// and dL/dt (3).
__device__ __forceinline__ void gs_joint_backward(double so, double co,
                                                  const double* __restrict__ geo, double ds,
                                                  const gs_pair_fwd& w, const double* gG,
                                                  gs_geo_acc& acc, double* gt) {
#pragma clang fp contract(fast)
    const double* mat = geo + 17;
    // G = Omega(omega) v
    const double gv[3] = {co * gG[0] - so * gG[1], so * gG[0] + co * gG[1], gG[2]};
    // v = mat^T k, k = ds (nh - x)
    double gk[3];
    const double k[3] = {ds * (w.nh[0] - 1.0), ds * w.nh[1], ds * w.nh[2]};
    for (int i = 0; i < 3; i++) {
        gk[i] = mat[3 * i] * gv[0] + mat[3 * i + 1] * gv[1] + mat[3 * i + 2] * gv[2];
        for (int j = 0; j < 3; j++) acc.m[3 * i + j] += k[i] * gv[j];
    }
    acc.ids += gk[0] * (w.nh[0] - 1.0) + gk[1] * w.nh[1] + gk[2] * w.nh[2];
    // nh = d / |d|: gd = (gn - nh (nh . gn)) / |d| with gn = ds gk
    const double dot = ds * (w.nh[0] * gk[0] + w.nh[1] * gk[1] + w.nh[2] * gk[2]);
    double gd[3];
    for (int i = 0; i < 3; i++) gd[i] = (ds * gk[i] - w.nh[i] * dot) * w.rn;
    // d = xl - mat.u
    double gu[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            acc.m[3 * i + j] -= gd[i] * w.u[j];
            gu[j] -= mat[3 * i + j] * gd[i];
        }
    gt[0] += co * gu[0] + so * gu[1];
    gt[1] += -so * gu[0] + co * gu[1];
    gt[2] += gu[2];
    // xl_j = r[3j+1] (f - p1) p3 + r[3j+2] (s - p0) p2 + dist_j
    for (int j = 0; j < 3; j++) {
        acc.A[j] += w.a1 * gd[j];
        acc.B[j] += w.a2 * gd[j];
        acc.D[j] += gd[j];
    }
}

// The 26 partials dL/dgeo of one grain from its summed accumulators
// (slot 16 holds dL/d(1/wavelength)).
__device__ __forceinline__ void gs_geo_expand(const gs_geo_acc& a, const double* __restrict__ geo,
                                              double* out) {
    const double* p = geo;
    const double* r = geo + 4;
    double sA = 0.0, sB = 0.0, sD1 = 0.0, sD2 = 0.0;
    for (int j = 0; j < 3; j++) {
        out[4 + 3 * j] = 0.0;
        out[4 + 3 * j + 1] = p[3] * a.A[j];
        out[4 + 3 * j + 2] = p[2] * a.B[j];
        out[13 + j] = a.D[j];
        sA += r[3 * j + 1] * a.A[j];
        sB += r[3 * j + 2] * a.B[j];
        sD1 += r[3 * j + 1] * a.D[j];
        sD2 += r[3 * j + 2] * a.D[j];
    }
    out[0] = -p[2] * sD2;
    out[1] = -p[3] * sD1;
    out[2] = sB;
    out[3] = sA;
    out[16] = a.ids;
    for (int i = 0; i < 9; i++) out[17 + i] = a.m[i];
}

// Fixed integer hkl of every list entry at the current state:
// hkl[q] = round(UBI_g . G_q) with G_q from the reference's own
// compute_xlylzl + compute_gv arithmetic (gs_xlylzl_one, gs_gv_one).  One
// lane per list entry.
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
    double x0, x1, x2, G0, G1, G2;
    gs_xlylzl_one(sc[k], fc[k], geo, geo + 4, geo + 13, x0, x1, x2);
    const double* t = trans + 3 * (size_t)g;
    gs_gv_one(x0, x1, x2, so[k], co[k], geo + 17, 1.0 / geo[16], t[0], t[1], t[2], G0, G1, G2);
    const double* u = ubis + 9 * (size_t)g;
    for (int i = 0; i < 3; i++) {
        double h = u[3 * i] * G0 + u[3 * i + 1] * G1 + u[3 * i + 2] * G2;
        hkl[3 * q + i] = (int)rint(h);
    }
}

// One list entry's inputs: the peak's (sc, fc, sin, cos) and its fixed
// hkl; ok = false for an index outside [0, npk) (the entry is skipped).
struct gs_adj_entry {
    double s, f, so, co;
    int h0, h1, h2;
    bool ok;
};

template <bool ROWS>
__device__ __forceinline__ gs_adj_entry gs_adj_fetch(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const int* __restrict__ idx,
    const int* __restrict__ hkl, long long q) {
    gs_adj_entry e;
    const long long k = ROWS ? q : (long long)idx[q];
    e.ok = k >= 0 && k < npk;
    const long long kk = e.ok ? k : 0;  // npk >= 1 whenever an entry is ok
    if (e.ok) {
        e.s = sc[kk];
        e.f = fc[kk];
        e.so = so[kk];
        e.co = co[kk];
        e.h0 = hkl[3 * q];
        e.h1 = hkl[3 * q + 1];
        e.h2 = hkl[3 * q + 2];
    } else {
        e.s = e.f = e.so = e.co = 0.0;
        e.h0 = e.h1 = e.h2 = 0;
    }
    return e;
}

// Per-grain forward + reverse sweep.  LPG lanes per grain, grains from a
// work queue; every per-grain output is the gs_group_sum over that grain's
// lanes, so it does not depend on which wave took the grain.
// Outputs per grain g: loss_g[g], grad_ubi[9g..] (dL/dUBI, or dE/dUB in
// g-space), grad_t[3g..], geo_part[26g..] (dL/dgeo restricted to grain g;
// slot 16 holds dL/d(1/wavelength)), and with NSUMS the Gauss-Newton sums
// sum h G^T (9) and sum G G^T (6: 00 01 02 11 12 22).  With NSUMS (hkl
// space) dL/dUBI = 2 (UBI . sum G G^T - sum h G^T) is formed from those
// sums instead of being accumulated separately.
// ROWS: no index array; sc/fc/so/co are already in list order (entry q is
// element q), so the pass streams them instead of gathering four cache
// lines per entry.
#ifndef GS_ADJ_MIN_BLOCKS
#define GS_ADJ_MIN_BLOCKS 1  // resident blocks per CU the register budget must allow
#endif
template <int LPG, bool GSPACE, bool NSUMS, bool ROWS>
__global__ __launch_bounds__(GS_BLOCK, GS_ADJ_MIN_BLOCKS) void k_joint_adjoint(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const double* __restrict__ geo,
    const double* __restrict__ ubis, const double* __restrict__ trans, int K,
    const int* __restrict__ idx, long long n_idx, const long long* __restrict__ off,
    const int* __restrict__ hkl, const int* __restrict__ order, double* __restrict__ loss_g,
    double* __restrict__ grad_ubi, double* __restrict__ grad_t, double* __restrict__ geo_part,
    double* __restrict__ nsums, int* __restrict__ queue) {
#pragma clang fp contract(fast)
    constexpr bool DERIVE_U = NSUMS && !GSPACE;  // dL/dUBI from the Gauss-Newton sums
    constexpr int GPW = GS_WAVE / LPG;
    const int lane = threadIdx.x & (GS_WAVE - 1);
    const int sub = lane / LPG;
    const int sl = lane - sub * LPG;
    const double ds = 1.0 / geo[16];
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
        double U[9];
        for (int i = 0; i < 9; i++) U[i] = active ? ubis[9 * (size_t)g + i] : (i % 4 == 0 ? 1.0 : 0.0);
        const double t0 = active ? trans[3 * (size_t)g] : 0.0;
        const double t1 = active ? trans[3 * (size_t)g + 1] : 0.0;
        const double t2 = active ? trans[3 * (size_t)g + 2] : 0.0;
        if (GSPACE) {  // UB = UBI^-1 (identity stays identity for inactive lanes)
            if (gs_inverse3x3(U) != 0)
                for (int i = 0; i < 9; i++) U[i] = __builtin_nan("");
        }
        double acc_l = 0.0, acc_u[9], acc_t[3] = {0.0, 0.0, 0.0}, acc_n[GS_NSUMS];
        gs_geo_acc acc;
        for (int i = 0; i < 9; i++) acc_u[i] = 0.0;
        for (int i = 0; i < 9; i++) acc.m[i] = 0.0;
        acc.ids = 0.0;
        for (int i = 0; i < 3; i++) acc.A[i] = acc.B[i] = acc.D[i] = 0.0;
        for (int i = 0; i < GS_NSUMS; i++) acc_n[i] = 0.0;
        // software-pipelined: the next entry's loads are issued before the
        // current entry is computed, so their latency overlaps the ~200
        // instructions of the forward and reverse sweeps
        gs_adj_entry nx;
        if (o0 + sl < o1) nx = gs_adj_fetch<ROWS>(sc, fc, so, co, npk, idx, hkl, o0 + sl);
        for (long long q = o0 + sl; q < o1; q += LPG) {
            const gs_adj_entry cu = nx;
            if (q + LPG < o1) nx = gs_adj_fetch<ROWS>(sc, fc, so, co, npk, idx, hkl, q + LPG);
            if (!cu.ok) continue;
            const double h0 = cu.h0, h1 = cu.h1, h2 = cu.h2;
            const double s_ = cu.s, f_ = cu.f, so_ = cu.so, co_ = cu.co;
            gs_pair_fwd w;
            gs_joint_forward(s_, f_, so_, co_, geo, ds, t0, t1, t2, w);
            double gG[3];
            if (!GSPACE) {
                double e[3];
                e[0] = U[0] * w.G[0] + U[1] * w.G[1] + U[2] * w.G[2] - h0;
                e[1] = U[3] * w.G[0] + U[4] * w.G[1] + U[5] * w.G[2] - h1;
                e[2] = U[6] * w.G[0] + U[7] * w.G[1] + U[8] * w.G[2] - h2;
                acc_l += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                if (!DERIVE_U)
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
            gs_joint_backward(so_, co_, geo, ds, w, gG, acc, acc_t);
        }
        acc_l = gs_group_sum<LPG>(acc_l);
        if (!DERIVE_U)
            for (int i = 0; i < 9; i++) acc_u[i] = gs_group_sum<LPG>(acc_u[i]);
        for (int i = 0; i < 3; i++) acc_t[i] = gs_group_sum<LPG>(acc_t[i]);
        for (int i = 0; i < 9; i++) acc.m[i] = gs_group_sum<LPG>(acc.m[i]);
        acc.ids = gs_group_sum<LPG>(acc.ids);
        for (int i = 0; i < 3; i++) {
            acc.A[i] = gs_group_sum<LPG>(acc.A[i]);
            acc.B[i] = gs_group_sum<LPG>(acc.B[i]);
            acc.D[i] = gs_group_sum<LPG>(acc.D[i]);
        }
        if (NSUMS)
            for (int i = 0; i < GS_NSUMS; i++) acc_n[i] = gs_group_sum<LPG>(acc_n[i]);
        if (active && sl == 0) {
            loss_g[g] = acc_l;
            if (DERIVE_U) {
                const double GG[9] = {acc_n[9],  acc_n[10], acc_n[11], acc_n[10], acc_n[12],
                                      acc_n[13], acc_n[11], acc_n[13], acc_n[14]};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        grad_ubi[9 * (size_t)g + 3 * i + j] =
                            2.0 * ((U[3 * i] * GG[j] + U[3 * i + 1] * GG[3 + j] +
                                    U[3 * i + 2] * GG[6 + j]) - acc_n[3 * i + j]);
            } else {
                for (int i = 0; i < 9; i++) grad_ubi[9 * (size_t)g + i] = acc_u[i];
            }
            for (int i = 0; i < 3; i++) grad_t[3 * (size_t)g + i] = acc_t[i];
            double gp[GS_GEO];
            gs_geo_expand(acc, geo, gp);
            for (int i = 0; i < GS_GEO; i++) geo_part[GS_GEO * (size_t)g + i] = gp[i];
            if (NSUMS)
                for (int i = 0; i < GS_NSUMS; i++) nsums[GS_NSUMS * (size_t)g + i] = acc_n[i];
        }
    }
}

// Deterministic reduction of the per-grain partials: out[0..25] = dL/dgeo
// (slot 16 converted from d/d(1/wavelength) to d/dwavelength), out[26] =
// loss, in two launches whose summation order depends only on K (not on
// which waves took which grains), so the result is the same on every run.
// Stage 1: nb = gs_fin_blocks(K) blocks; block b takes the grains
// [b*per, min(K, (b+1)*per)), per = ceil(K/nb); thread (r, j) of its
// GS_FIN_GROUPS x 32 threads sums column j over the grains start + r,
// start + r + GS_FIN_GROUPS, ... in that order, then thread j adds the
// GS_FIN_GROUPS group sums in group order into part[27 b + j].
// Stage 2: thread (r, j) of one block sums part[27 b + j] over b = r, r +
// GS_FIN_GROUPS, ..., then thread j adds the group sums in order.
#define GS_FIN_GROUPS 8
#define GS_FIN_BLOCKS 128  // most stage-1 blocks (the workspace holds 27 per block)
#define GS_FIN_MIN_GRAINS 32

GS_HD __forceinline__ int gs_fin_blocks(int K) {
    int nb = (K + GS_FIN_MIN_GRAINS - 1) / GS_FIN_MIN_GRAINS;
    return nb < 1 ? 1 : (nb > GS_FIN_BLOCKS ? GS_FIN_BLOCKS : nb);
}

__global__ __launch_bounds__(GS_FIN_GROUPS * 32) void k_joint_reduce1(
    const double* __restrict__ geo_part, const double* __restrict__ loss_g, int K,
    double* __restrict__ part) {
    __shared__ double grp[GS_FIN_GROUPS][GS_GEO_OUT];
    const int j = threadIdx.x & 31;
    const int r = threadIdx.x >> 5;
    const int nb = gridDim.x;
    const int per = (K + nb - 1) / nb;
    const long long g0 = (long long)blockIdx.x * per;
    const long long g1 = g0 + per < K ? g0 + per : K;
    if (j < GS_GEO_OUT) {
        double s = 0.0;
        long long g = g0 + r;
        // loads run ahead in groups of four; additions stay in order
        for (; g + 3 * GS_FIN_GROUPS < g1; g += 4 * GS_FIN_GROUPS) {
            double v[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long long gg = g + (long long)q * GS_FIN_GROUPS;
                v[q] = j < GS_GEO ? geo_part[GS_GEO * gg + j] : loss_g[gg];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) s += v[q];
        }
        for (; g < g1; g += GS_FIN_GROUPS) s += j < GS_GEO ? geo_part[GS_GEO * g + j] : loss_g[g];
        grp[r][j] = s;
    }
    __syncthreads();
    if (r == 0 && j < GS_GEO_OUT) {
        double s = 0.0;
        for (int q = 0; q < GS_FIN_GROUPS; q++) s += grp[q][j];
        part[GS_GEO_OUT * (size_t)blockIdx.x + j] = s;
    }
}

__global__ __launch_bounds__(GS_FIN_GROUPS * 32) void k_joint_reduce2(
    const double* __restrict__ part, int nb, const double* __restrict__ geo_in,
    double* __restrict__ out) {
    __shared__ double grp[GS_FIN_GROUPS][GS_GEO_OUT];
    const int j = threadIdx.x & 31;
    const int r = threadIdx.x >> 5;
    if (j < GS_GEO_OUT) {
        double s = 0.0;
        for (int b = r; b < nb; b += GS_FIN_GROUPS) s += part[GS_GEO_OUT * (size_t)b + j];
        grp[r][j] = s;
    }
    __syncthreads();
    if (r == 0 && j < GS_GEO_OUT) {
        double s = 0.0;
        for (int q = 0; q < GS_FIN_GROUPS; q++) s += grp[q][j];
        if (j == 16) s = -s / (geo_in[16] * geo_in[16]);
        out[j] = s;
    }
}
