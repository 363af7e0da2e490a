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
#pragma clang fp contract(on)
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
    w.rn = gs_rsqrt_cert<2>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
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
// are accumulated per lane instead of 26 (gs_geo_expand_one forms the 26).
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
#pragma clang fp contract(on)
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

// The geometry sums as 19 consecutive numbers (the order of a piece's
// partials in the workspace): m[9], ids, A[3], B[3], D[3].
#define GS_ACC_N 19
#define GS_ACC_IDS 9
#define GS_ACC_A 10
#define GS_ACC_B 13
#define GS_ACC_D 16

__device__ __forceinline__ void gs_geo_acc_flat(const gs_geo_acc& a, double* v) {
    for (int i = 0; i < 9; i++) v[i] = a.m[i];
    v[GS_ACC_IDS] = a.ids;
    for (int j = 0; j < 3; j++) {
        v[GS_ACC_A + j] = a.A[j];
        v[GS_ACC_B + j] = a.B[j];
        v[GS_ACC_D + j] = a.D[j];
    }
}

// Partial k (0..25) of dL/dgeo from summed geometry sums a[19] (slot 16
// is dL/d(1/wavelength)).  r[3j] does not enter compute_xlylzl: its
// partials are 0.
__device__ __forceinline__ double gs_geo_expand_one(int k, const double* a,
                                                    const double* __restrict__ geo) {
    const double* p = geo;
    const double* r = geo + 4;
    if (k < 4) {
        // p0, p1 enter through (s - p0), (f - p1); p2, p3 scale them
        const int col = (k == 0 || k == 2) ? 2 : 1;
        const int src = k < 2 ? GS_ACC_D : (k == 2 ? GS_ACC_B : GS_ACC_A);
        double s = 0.0;
        for (int j = 0; j < 3; j++) s += r[3 * j + col] * a[src + j];
        return k == 0 ? -p[2] * s : (k == 1 ? -p[3] * s : s);
    }
    if (k < 13) {
        const int j = (k - 4) / 3, c = (k - 4) - 3 * j;
        return c == 0 ? 0.0 : (c == 1 ? p[3] * a[GS_ACC_A + j] : p[2] * a[GS_ACC_B + j]);
    }
    if (k < 16) return a[GS_ACC_D + (k - 13)];
    if (k == 16) return a[GS_ACC_IDS];
    return a[k - 17];
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

// ROWS: entry q is element q of the per-entry arrays; the caller keeps q
// below n_idx <= npk, so no entry is skipped and no bound is tested.
template <bool ROWS>
__device__ __forceinline__ gs_adj_entry gs_adj_fetch(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const int* __restrict__ idx,
    const int* __restrict__ hkl, long long q) {
    gs_adj_entry e;
    const long long k = ROWS ? q : (long long)idx[q];
    e.ok = ROWS || (k >= 0 && k < npk);
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

// ---- the pass over pieces ---------------------------------------------------
//
// Every grain's list [off[g], off[g+1]) is cut into pieces of at most
// piece_len consecutive entries, numbered in grain order: grain g owns
// pieces piece_off[g] .. piece_off[g+1]-1, piece p of grain g starts at
// entry off[g] + (p - piece_off[g]) piece_len, and piece_grain[p] = g.
// slots[] is the launch order (longest piece first).  No piece is longer
// than piece_len, so no grain's list length sets the pass's duration, and
// a wave's pieces have about the same length.
//
// Workspace: per piece GS_ADJ_STRIDE partial sums, then per block of the
// piece pass GS_ADJ_BLK sums (loss and the 19 geometry sums).
#define GS_ADJ_STRIDE 48
#define GS_ADJ_LOSS 0
#define GS_ADJ_U 1    // dL/dUBI (9), when it is not formed from the Gauss-Newton sums
#define GS_ADJ_N 10   // Gauss-Newton sums (15)
#define GS_ADJ_T 25   // dL/dt (3)
#define GS_ADJ_GEO 28 // the 19 geometry sums
#define GS_ADJ_BLK 20

// Blocks of the piece pass for n_pieces pieces at lanes per piece, and
// the workspace doubles (the largest block count, at 64 lanes per piece,
// so one workspace serves every lane count).
GS_HD __forceinline__ long long gs_adj_blocks(long long n_pieces, int lanes) {
    const long long per_block = (long long)(GS_BLOCK / GS_WAVE) * (GS_WAVE / lanes);
    return (n_pieces + per_block - 1) / per_block;
}
GS_HD __forceinline__ long long gs_adj_ws_doubles(long long n_pieces) {
    return n_pieces * GS_ADJ_STRIDE + gs_adj_blocks(n_pieces, GS_WAVE) * GS_ADJ_BLK;
}

// One 64-bit value of another lane (a uniform lane number).
__device__ __forceinline__ double gs_readlane_d(double x, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

#ifndef GS_ADJ_DEPTH
#define GS_ADJ_DEPTH 1  // list entries whose loads are in flight ahead of the current one
#endif

// Forward + reverse sweep over pieces: LPG lanes per piece, 64/LPG pieces
// per wave, slots taken in order (slot s by wave s / (64/LPG)).  Each
// piece's sums (gs_group_sum over its lanes: the same totals in every
// lane) go to part[GS_ADJ_STRIDE * p ..]: loss, dL/dUBI (or the
// Gauss-Newton sums, from which dL/dUBI = 2 (UBI . sum G G^T - sum h G^T)
// is formed per grain), dL/dt and the geometry sums.  The block adds its
// pieces' loss and geometry sums in slot order into blk[GS_ADJ_BLK * b ..].
// Every summation order depends only on the schedule, so the results are
// the same on every run.  Out-of-range schedule entries are not trusted: a
// slot whose piece or grain number is out of range does nothing, and list
// bounds are clipped to [0, n_idx].
// ROWS: no index array; sc/fc/so/co are already in list order (entry q is
// element q), so the pass streams them instead of gathering four cache
// lines per entry.
template <int LPG, bool GSPACE, bool NSUMS, bool ROWS>
__global__ __launch_bounds__(GS_BLOCK) void k_joint_pieces(
    const double* __restrict__ sc, const double* __restrict__ fc, const double* __restrict__ so,
    const double* __restrict__ co, int npk, const double* __restrict__ geo,
    const double* __restrict__ ubis, const double* __restrict__ trans, int K,
    const int* __restrict__ idx, long long n_idx, const long long* __restrict__ off,
    const int* __restrict__ hkl, const long long* __restrict__ piece_off,
    const int* __restrict__ piece_grain, const int* __restrict__ slots, long long n_pieces,
    int piece_len, double* __restrict__ part, double* __restrict__ blk) {
#pragma clang fp contract(on)
    constexpr bool DERIVE_U = NSUMS && !GSPACE;  // dL/dUBI from the Gauss-Newton sums
    constexpr int GPW = GS_WAVE / LPG;
    constexpr int WPB = GS_BLOCK / GS_WAVE;
    __shared__ double lds[WPB][GS_ADJ_BLK];
    const int lane = threadIdx.x & (GS_WAVE - 1);
    const int wv = threadIdx.x / GS_WAVE;
    const int sub = lane / LPG;
    const int sl = lane - sub * LPG;
    const double ds = 1.0 / geo[16];
    const long long slot = ((long long)blockIdx.x * WPB + wv) * GPW + sub;
    const int pid = slot < n_pieces ? slots[slot] : -1;
    const int gq = (pid >= 0 && pid < n_pieces) ? piece_grain[pid] : -1;
    const bool active = gq >= 0 && gq < K;
    const int g = active ? gq : 0;
    long long q0 = 0, q1 = 0;
    if (active) {
        long long o1 = off[g + 1];
        o1 = o1 < 0 ? 0 : (o1 > n_idx ? n_idx : o1);
        long long o0 = off[g];
        o0 = o0 < 0 ? 0 : (o0 > o1 ? o1 : o0);
        q0 = o0 + ((long long)pid - piece_off[g]) * (long long)piece_len;
        q0 = q0 < o0 ? o0 : (q0 > o1 ? o1 : q0);
        q1 = o1 - q0 > piece_len ? q0 + piece_len : o1;
    }
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
    // software-pipelined: the loads of the entries GS_ADJ_DEPTH steps ahead
    // are issued before the current entry is computed, so their latency
    // overlaps the ~200 instructions of the forward and reverse sweeps (one
    // step ahead measured fastest on MI355X: 34.4 us per evaluation at the
    // 5000-grain shape, 34.7 at two, 38.0 at three)
    constexpr int D = GS_ADJ_DEPTH;
    gs_adj_entry ahead[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        const long long qq = q0 + sl + (long long)d * LPG;
        ahead[d] = {0.0, 0.0, 0.0, 0.0, 0, 0, 0, false};
        if (qq < q1) ahead[d] = gs_adj_fetch<ROWS>(sc, fc, so, co, npk, idx, hkl, qq);
    }
    for (long long q = q0 + sl; q < q1; q += LPG) {
        const gs_adj_entry cu = ahead[0];
#pragma unroll
        for (int d = 0; d + 1 < D; d++) ahead[d] = ahead[d + 1];
        if (q + (long long)D * LPG < q1)
            ahead[D - 1] = gs_adj_fetch<ROWS>(sc, fc, so, co, npk, idx, hkl, q + (long long)D * LPG);
        if (!cu.ok) continue;
        const double h0 = cu.h0, h1 = cu.h1, h2 = cu.h2;
        gs_pair_fwd w;
        gs_joint_forward(cu.s, cu.f, cu.so, cu.co, geo, ds, t0, t1, t2, w);
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
        gs_joint_backward(cu.so, cu.co, geo, ds, w, gG, acc, acc_t);
    }
    double a[GS_ACC_N];
    gs_geo_acc_flat(acc, a);
    acc_l = gs_group_sum<LPG>(acc_l);
    for (int i = 0; i < GS_ACC_N; i++) a[i] = gs_group_sum<LPG>(a[i]);
    for (int i = 0; i < 3; i++) acc_t[i] = gs_group_sum<LPG>(acc_t[i]);
    if (!DERIVE_U)
        for (int i = 0; i < 9; i++) acc_u[i] = gs_group_sum<LPG>(acc_u[i]);
    if (NSUMS)
        for (int i = 0; i < GS_NSUMS; i++) acc_n[i] = gs_group_sum<LPG>(acc_n[i]);
    if (active && sl == 0) {
        double* w = part + (size_t)pid * GS_ADJ_STRIDE;
        w[GS_ADJ_LOSS] = acc_l;
        if (!DERIVE_U)
            for (int i = 0; i < 9; i++) w[GS_ADJ_U + i] = acc_u[i];
        if (NSUMS)
            for (int i = 0; i < GS_NSUMS; i++) w[GS_ADJ_N + i] = acc_n[i];
        for (int i = 0; i < 3; i++) w[GS_ADJ_T + i] = acc_t[i];
        for (int i = 0; i < GS_ACC_N; i++) w[GS_ADJ_GEO + i] = a[i];
    }
    // the block's loss and geometry sums: the wave's pieces in slot order
    // (an inactive piece's sums are exactly zero), then the waves in order
    double wsum[GS_ADJ_BLK];
    wsum[0] = acc_l;
    for (int i = 0; i < GS_ACC_N; i++) wsum[1 + i] = a[i];
    for (int i = 0; i < GS_ADJ_BLK; i++) {
        double s = gs_readlane_d(wsum[i], 0);
        for (int k = 1; k < GPW; k++) s += gs_readlane_d(wsum[i], k * LPG);
        wsum[i] = s;
    }
    if (lane == 0)
        for (int i = 0; i < GS_ADJ_BLK; i++) lds[wv][i] = wsum[i];
    __syncthreads();
    if (threadIdx.x < GS_ADJ_BLK) {
        double s = lds[0][threadIdx.x];
        for (int k = 1; k < WPB; k++) s += lds[k][threadIdx.x];
        blk[(size_t)blockIdx.x * GS_ADJ_BLK + threadIdx.x] = s;
    }
}

// x[i * stride] summed over i = i0, i0 + step, ... below n, in that order;
// the loads are issued GS_SUM_BATCH at a time (slots past n add +0.0), so
// a sum of up to GS_SUM_BATCH terms costs one memory latency.
#define GS_SUM_BATCH 16
__device__ __forceinline__ double gs_ordered_sum(const double* __restrict__ x, long long i0,
                                                 long long n, long long step, long long stride) {
    double s = 0.0;
    for (long long i = i0; i < n; i += GS_SUM_BATCH * step) {
        double v[GS_SUM_BATCH];
#pragma unroll
        for (int q = 0; q < GS_SUM_BATCH; q++) {
            const long long k = i + q * step;
            v[q] = k < n ? x[k * stride] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < GS_SUM_BATCH; q++) s += v[q];
    }
    return s;
}

// Per-grain results and the totals, one launch of GS_FIN_THREADS threads
// per block.
// Blocks 0 .. ceil(K / GS_FIN_GRAINS) - 1: wave w of block b takes grain
// g = GS_FIN_GRAINS b + w; lane j adds partial j over the grain's pieces in
// piece order, then the lanes write loss_g[g], grad_ubi[9g..] (dL/dUBI, or
// dE/dUB in g-space), nsums[15g..] (sum h G^T, then sum G G^T as 00 01 02
// 11 12 22; when requested), grad_t[3g..] and geo_part[26g..] (dL/dgeo
// restricted to grain g, slot 16 being dL/d(1/wavelength); when
// requested).
// The last block: out[0..25] = dL/dgeo (slot 16 converted to
// d/dwavelength), out[26] = loss, from the nblk block sums of the piece
// pass: thread (r, j) adds column j over the blocks r, r + 32, ... in that
// order, then thread j adds the 32 group sums in order.
#define GS_FIN_THREADS 1024
#define GS_FIN_GRAINS (GS_FIN_THREADS / GS_WAVE)

GS_HD __forceinline__ int gs_fin_grain_blocks(int K) {
    return (K + GS_FIN_GRAINS - 1) / GS_FIN_GRAINS;
}

template <bool GSPACE, bool NSUMS>
__global__ __launch_bounds__(GS_FIN_THREADS) void k_joint_finish(
    const double* __restrict__ part, const long long* __restrict__ piece_off, long long n_pieces,
    int K, const double* __restrict__ ubis, const double* __restrict__ geo,
    double* __restrict__ loss_g, double* __restrict__ grad_ubi, double* __restrict__ grad_t,
    double* __restrict__ geo_part, double* __restrict__ nsums, const double* __restrict__ blk,
    long long nblk, double* __restrict__ out) {
#pragma clang fp contract(on)
    constexpr bool DERIVE_U = NSUMS && !GSPACE;
    __shared__ double s[GS_FIN_GRAINS][GS_ADJ_STRIDE];
    const int t = threadIdx.x;
    if ((int)blockIdx.x < gs_fin_grain_blocks(K)) {
        const int lane = t & (GS_WAVE - 1);
        const int wv = t / GS_WAVE;
        const int g = (int)blockIdx.x * GS_FIN_GRAINS + wv;
        const bool ok = g < K;
        const bool used = lane == GS_ADJ_LOSS || (!DERIVE_U && lane >= GS_ADJ_U && lane < GS_ADJ_N) ||
                          (NSUMS && lane >= GS_ADJ_N && lane < GS_ADJ_T) ||
                          (lane >= GS_ADJ_T && lane < GS_ADJ_GEO + GS_ACC_N);
        if (ok && lane < GS_ADJ_STRIDE) {
            long long p1 = piece_off[g + 1];
            p1 = p1 < 0 ? 0 : (p1 > n_pieces ? n_pieces : p1);
            long long p0 = piece_off[g];
            p0 = p0 < 0 ? 0 : (p0 > p1 ? p1 : p0);
            s[wv][lane] = used ? gs_ordered_sum(part + lane, p0, p1, 1, GS_ADJ_STRIDE) : 0.0;
        }
        __syncthreads();
        if (!ok) return;
        const double* a = s[wv];
        if (lane == GS_ADJ_LOSS) {
            loss_g[g] = a[GS_ADJ_LOSS];
        } else if (lane < GS_ADJ_N) {
            const int i = lane - GS_ADJ_U, r = i / 3, c = i - 3 * r;
            double v;
            if (DERIVE_U) {
                const double* n = a + GS_ADJ_N;
                const double GG[9] = {n[9], n[10], n[11], n[10], n[12], n[13], n[11], n[13], n[14]};
                const double* u = ubis + 9 * (size_t)g + 3 * r;
                v = 2.0 * ((u[0] * GG[c] + u[1] * GG[3 + c] + u[2] * GG[6 + c]) - n[3 * r + c]);
            } else {
                v = a[GS_ADJ_U + i];
            }
            grad_ubi[9 * (size_t)g + i] = v;
        } else if (lane < GS_ADJ_T) {
            if (NSUMS && nsums) nsums[GS_NSUMS * (size_t)g + (lane - GS_ADJ_N)] = a[lane];
        } else if (lane < GS_ADJ_GEO) {
            grad_t[3 * (size_t)g + (lane - GS_ADJ_T)] = a[lane];
        } else if (lane < GS_ADJ_GEO + GS_GEO) {
            if (geo_part)
                geo_part[GS_GEO * (size_t)g + (lane - GS_ADJ_GEO)] =
                    gs_geo_expand_one(lane - GS_ADJ_GEO, a + GS_ADJ_GEO, geo);
        }
        return;
    }
    // the totals
    double* grp = &s[0][0];  // 32 x GS_ADJ_BLK
    double* tot = grp + 32 * GS_ADJ_BLK;
    static_assert(33 * GS_ADJ_BLK <= GS_FIN_GRAINS * GS_ADJ_STRIDE, "totals fit the grain buffer");
    const int r = t >> 5, j = t & 31;
    if (j < GS_ADJ_BLK) grp[r * GS_ADJ_BLK + j] = gs_ordered_sum(blk + j, r, nblk, 32, GS_ADJ_BLK);
    __syncthreads();
    if (t < GS_ADJ_BLK) {
        double sum = 0.0;
        for (int q = 0; q < 32; q++) sum += grp[q * GS_ADJ_BLK + t];
        tot[t] = sum;
    }
    __syncthreads();
    if (t < GS_GEO) {
        double v = gs_geo_expand_one(t, tot + 1, geo);
        if (t == 16) v = -v / (geo[16] * geo[16]);
        out[t] = v;
    } else if (t == GS_GEO) {
        out[GS_GEO] = tot[0];
    }
}

// The per-grain Gauss-Newton UBI for fixed g-vectors and hkl: the UBI that
// minimises sum ||UBI.G - h||^2 is A M^-1 with A = sum h G^T (row-major)
// and M = sum G G^T (00 01 02 11 12 22), the 15 sums k_joint_finish
// writes.  M^-1 = adj(M)/det by cofactors (M is symmetric).  A grain with
// det == 0 keeps prev (NaN when there is none), as score_and_refine leaves
// a UBI whose normal matrix is singular (src/closest.c:320-336, where the
// fit is of UB in g-space).  out may be prev: every read precedes the
// first write.
GS_HD __forceinline__ void gs_gn_ubi(const double* __restrict__ n, const double* prev,
                                     double* out) {
    const double s0 = n[9], s1 = n[10], s2 = n[11], s3 = n[12], s4 = n[13], s5 = n[14];
    const double c00 = s3 * s5 - s4 * s4, c01 = s2 * s4 - s1 * s5, c02 = s1 * s4 - s2 * s3;
    const double c11 = s0 * s5 - s2 * s2, c12 = s1 * s2 - s0 * s4, c22 = s0 * s3 - s1 * s1;
    const double det = s0 * c00 + s1 * c01 + s2 * c02;
    double u[9];
    if (det == 0.0) {
        for (int i = 0; i < 9; i++) u[i] = prev ? prev[i] : __builtin_nan("");
    } else {
        const double i00 = c00 / det, i01 = c01 / det, i02 = c02 / det;
        const double i11 = c11 / det, i12 = c12 / det, i22 = c22 / det;
        const double inv[9] = {i00, i01, i02, i01, i11, i12, i02, i12, i22};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++)
                u[3 * r + c] = n[3 * r] * inv[c] + n[3 * r + 1] * inv[3 + c] +
                               n[3 * r + 2] * inv[6 + c];
    }
    for (int i = 0; i < 9; i++) out[i] = u[i];
}

// one thread per grain
__global__ __launch_bounds__(GS_BLOCK) void k_gn_ubis(const double* __restrict__ nsums,
                                                      const double* prev, int K, double* out) {
    const int g = (int)blockIdx.x * GS_BLOCK + (int)threadIdx.x;
    if (g >= K) return;
    gs_gn_ubi(nsums + GS_NSUMS * (size_t)g, prev ? prev + 9 * (size_t)g : nullptr,
              out + 9 * (size_t)g);
}
