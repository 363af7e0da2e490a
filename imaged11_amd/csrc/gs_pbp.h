// Point-by-point refinement (ImageD11/sinograms/point_by_point.py:
// 1432-1670, refine_map): the per-voxel selection and scoring of peaks,
// for every (voxel, UBI) task of a map in one pass over the peak table.
//
// For a voxel at sample position (xi0, yi0) a peak k is local when the
// beam passes through the voxel at the peak's rotation and sample
// y-translation (get_voxel_idx, :359-368):
//     ydist = |y0 - xi0 sin(omega_k) - yi0 cos(omega_k) - dty_k| <= ystep,
// evaluated left to right as numpy does.  A local peak is assigned to a
// UBI of the voxel when sum((h - round(h))^2) < tol^2 with h = UBI.g_k and
// round half to even (np.round), the first selection of refine_map
// (:1537-1545).  The kernel runs twice: COUNT adds up the assigned peaks
// of each task, FILL writes (peak, h, k, l) of each, at
// task_off[t] + (a slot taken from cursor[t]); the caller orders each
// task's entries by peak index afterwards.  Both passes make the same
// decisions (the same code on the same inputs), and FILL never writes past
// a task's count.
#pragma once
#include "gs_kernels.h"

#define GS_PBP_POINTS 64  // voxels per block (one point chunk)

template <int FILL>
__global__ __launch_bounds__(GS_BLOCK) void k_pbp_assign(
    const double* __restrict__ so, const double* __restrict__ co,
    const double* __restrict__ dty, const double* __restrict__ gv, int ng,
    const double* __restrict__ pts, const int* __restrict__ ptask, int P,
    const double* __restrict__ ubis, double y0, double ystep, double tolsq,
    int* __restrict__ counts, const long long* __restrict__ task_off, int* __restrict__ cursor,
    int* __restrict__ out_peak, int* __restrict__ out_hkl) {
    const long long k = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    const int p0 = blockIdx.y * GS_PBP_POINTS;
    const int p1 = min(P, p0 + GS_PBP_POINTS);
    const bool have = k < ng;
    double s = 0.0, c = 0.0, d = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (have) {
        s = so[k];
        c = co[k];
        d = dty[k];
        g0 = gv[3 * k];
        g1 = gv[3 * k + 1];
        g2 = gv[3 * k + 2];
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    for (int p = p0; p < p1; p++) {
        const double xi0 = pts[2 * p], yi0 = pts[2 * p + 1];
        const double ydist = fabs(((y0 - xi0 * s) - yi0 * c) - d);
        const bool local = have && ydist <= ystep;
        if (!__any(local)) continue;
        for (int t = ptask[p]; t < ptask[p + 1]; t++) {
            const double* u = ubis + 9 * (size_t)t;
            bool take = false;
            double hr[3];
            if (local) {
                double e = 0.0;
                for (int i = 0; i < 3; i++) {
                    const double h = u[3 * i] * g0 + u[3 * i + 1] * g1 + u[3 * i + 2] * g2;
                    hr[i] = rint(h);
                    const double r = h - hr[i];
                    e = i == 0 ? r * r : e + r * r;
                }
                // hkl are stored as int32 (the reference's .astype(np.int32))
                take = e < tolsq && fabs(hr[0]) < 2147483647.0 && fabs(hr[1]) < 2147483647.0 &&
                       fabs(hr[2]) < 2147483647.0;
            }
            const unsigned long long m = __ballot(take);
            if (!m) continue;
            if (!FILL) {
                if (lane == 0) atomicAdd(&counts[t], __popcll(m));
            } else if (take) {
                // one slot per assigned lane: the wave's first lane reserves
                // the block of slots, each lane takes its rank in the mask
                const int leader = __ffsll((long long)m) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&cursor[t], __popcll(m));
                base = __shfl(base, leader, GS_WAVE);
                const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
                if (slot < counts[t]) {
                    const long long o = task_off[t] + slot;
                    out_peak[o] = (int)k;
                    out_hkl[3 * o] = (int)hr[0];
                    out_hkl[3 * o + 1] = (int)hr[1];
                    out_hkl[3 * o + 2] = (int)hr[2];
                }
            }
        }
    }
}

// Per-segment column reductions, row by row from 0.0 in row order:
// out[s][c] = the sum (MAX false) or the maximum (MAX true) of v[r][c] and
// 0.0 for r in [off[s], off[s+1]).  The sums are added in the order in which
// np.bincount(labels, weights) accumulates, so grouped weighted sums
// (refine_map's merge, :338-356) come out bit for bit as the reference's,
// and the same on every run; the maximum propagates NaN.  One thread per
// (segment, column); consecutive threads read consecutive columns of a row.
template <bool MAX>
__global__ __launch_bounds__(GS_BLOCK) void k_segment_sum(const double* __restrict__ v, int C,
                                                          const long long* __restrict__ off,
                                                          int S, double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= (long long)S * C) return;
    const long long s = i / C;
    const int c = (int)(i - s * C);
    double acc = 0.0;
    long long r = off[s];
    const long long r1 = off[s + 1];
    auto add = [&](double x) {
        if (!MAX) {
            acc += x;
        } else if (x != x || x > acc) {
            if (acc == acc) acc = x;  // a NaN stays
        }
    };
    // loads run ahead in groups of eight; the reduction stays in row order
    for (; r + 8 <= r1; r += 8) {
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = v[(r + j) * C + c];
#pragma unroll
        for (int j = 0; j < 8; j++) add(x[j]);
    }
    for (; r < r1; r++) add(v[r * C + c]);
    out[i] = acc;
}

// The per-peak columns the merge of refine_map averages.
struct gs_merge_cols {
    const double *si, *sc, *fc, *om, *dty, *xpos, *eta;
};

// refine_map's merge (:338-356) for every (task, h, k, l, sign eta, dtyi)
// group at once: out[g] = (sum w, sum sc w, sum fc w, sum omega w, sum dty w,
// sum xpos w, sum eta w) over rows r in [goff[g], goff[g+1]) of the grouped
// order, row r being peak p = peak[perm[r]] with weight w = si[p].  Each
// sum starts at 0.0 and adds in row order, and each product is x * w: the
// gathered, multiplied, permuted and segment-summed columns of the
// unfused form, bit for bit.  One thread per group; a peak outside [0, npk)
// is skipped.
__global__ __launch_bounds__(GS_BLOCK) void k_pbp_merge_sums(const long long* __restrict__ perm,
                                                             const long long* __restrict__ peak,
                                                             gs_merge_cols c,
                                                             const long long* __restrict__ goff,
                                                             int G, long long npk,
                                                             double* __restrict__ out) {
    const long long g = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (g >= G) return;
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long r = goff[g]; r < goff[g + 1]; r++) {
        const long long p = peak[perm[r]];
        if (p < 0 || p >= npk) continue;
        const double w = c.si[p];
        acc[0] += w;
        acc[1] += c.sc[p] * w;
        acc[2] += c.fc[p] * w;
        acc[3] += c.om[p] * w;
        acc[4] += c.dty[p] * w;
        acc[5] += c.xpos[p] * w;
        acc[6] += c.eta[p] * w;
    }
#pragma unroll
    for (int j = 0; j < 7; j++) out[7 * g + j] = acc[j];
}

// ---------------------------------------------------------------------------
// The per-task least-squares fits of refine_map, one thread per task,
// solved as np.linalg.lstsq solves them: the minimum-norm solution with
// the singular values of the weighted design matrix below rcond * the
// largest taken as zero, rcond = eps * max(rows, columns).  Each task's
// rows stream into the triangular factor R of a QR decomposition (Givens
// rotations, the right-hand sides carried along as extra columns), and R,
// which has the design matrix's singular values, is decomposed by one-sided
// Jacobi rotations.  Nothing squares the condition number.

// Rotate one row x[0..NC+NR) into the upper-triangular [R | C].
template <int NC, int NR>
__device__ __forceinline__ void gs_qr_add_row(double (&R)[NC][NC + NR], double (&x)[NC + NR]) {
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const double a = R[k][k], b = x[k];
        if (b == 0.0) continue;
        const double r = sqrt(a * a + b * b);
        const double c = a / r, s = b / r;
#pragma unroll
        for (int j = k; j < NC + NR; j++) {
            const double rj = R[k][j], xj = x[j];
            R[k][j] = c * rj + s * xj;
            x[j] = c * xj - s * rj;
        }
    }
}

// Stream rows [r0, r1) of a row-major (rows, NC+NR) array into [R | C],
// CHUNK rows in flight: their loads are issued together (the row index
// clamped to r1 - 1, so always in bounds), then they are rotated in in
// row order and the copies past r1 dropped.
template <int NC, int NR, int CHUNK>
__device__ __forceinline__ void gs_qr_add_rows(double (&R)[NC][NC + NR],
                                               const double* __restrict__ rows, long long r0,
                                               long long r1) {
    for (long long r = r0; r < r1; r += CHUNK) {
        double x[CHUNK][NC + NR];
#pragma unroll
        for (int c = 0; c < CHUNK; c++) {
            const long long rr = r + c < r1 ? r + c : r1 - 1;
#pragma unroll
            for (int j = 0; j < NC + NR; j++) x[c][j] = rows[(NC + NR) * rr + j];
        }
#pragma unroll
        for (int c = 0; c < CHUNK; c++)
            if (r + c < r1) gs_qr_add_row<NC, NR>(R, x[c]);
    }
}

// One-sided Jacobi SVD of the N x N matrix A: on return A's column j is
// U_j * sv[j] and V holds the right singular vectors.  At most 60 sweeps.
template <int N>
__device__ __forceinline__ void gs_jacobi_svd(double (&A)[N][N], double (&V)[N][N],
                                              double (&sv)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < N - 1; p++) {
#pragma unroll
            for (int q = p + 1; q < N; q++) {
                double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) {
                    al += A[i][p] * A[i][p];
                    be += A[i][q] * A[i][q];
                    ga += A[i][p] * A[i][q];
                }
                if (!(fabs(ga) > 2.220446049250313e-16 * sqrt(al * be))) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = fabs(zeta) > 1e150 ? 0.5 / zeta
                                                    : copysign(1.0, zeta) /
                                                          (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
#pragma unroll
                for (int i = 0; i < N; i++) {
                    const double ap = A[i][p], aq = A[i][q];
                    A[i][p] = c * ap - s * aq;
                    A[i][q] = s * ap + c * aq;
                    const double vp = V[i][p], vq = V[i][q];
                    V[i][p] = c * vp - s * vq;
                    V[i][q] = s * vp + c * vq;
                }
            }
        }
        if (!rotated) break;
    }
#pragma unroll
    for (int j = 0; j < N; j++) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) s += A[i][j] * A[i][j];
        sv[j] = sqrt(s);
    }
}

// Minimum-norm least-squares solution X (NC x NR) of the rows streamed into
// [R | C], singular values <= rcond * max taken as zero; returns the rank.
template <int NC, int NR>
__device__ __forceinline__ int gs_lstsq_solve(double (&R)[NC][NC + NR], double rcond,
                                              double (&X)[NC][NR]) {
    bool finite = true;
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
        for (int j = 0; j < NC + NR; j++) finite = finite && isfinite(R[i][j]);
    if (!finite) {  // numpy's lstsq fails on such rows: no solution
#pragma unroll
        for (int i = 0; i < NC; i++)
#pragma unroll
            for (int k = 0; k < NR; k++) X[i][k] = __builtin_nan("");
        return 0;
    }
    double A[NC][NC], V[NC][NC], sv[NC];
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
        for (int j = 0; j < NC; j++) A[i][j] = R[i][j];
    gs_jacobi_svd<NC>(A, V, sv);
    double smax = 0.0;
#pragma unroll
    for (int j = 0; j < NC; j++) smax = fmax(smax, sv[j]);
    const double cut = rcond * smax;
#pragma unroll
    for (int i = 0; i < NC; i++)
#pragma unroll
        for (int k = 0; k < NR; k++) X[i][k] = 0.0;
    int rank = 0;
#pragma unroll
    for (int j = 0; j < NC; j++) {
        if (!(sv[j] > cut)) continue;
        rank++;
        const double inv = 1.0 / (sv[j] * sv[j]);  // A_j = U_j sv_j
#pragma unroll
        for (int k = 0; k < NR; k++) {
            double d = 0.0;
#pragma unroll
            for (int i = 0; i < NC; i++) d += A[i][j] * R[i][NC + k];
            d *= inv;
#pragma unroll
            for (int i = 0; i < NC; i++) X[i][k] += V[i][j] * d;
        }
    }
    return rank;
}

// Singular values of a 3x3 matrix (row-major).
__device__ __forceinline__ void gs_sv3(const double* m, double (&sv)[3]) {
    double A[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) A[i][j] = m[3 * i + j];
    gs_jacobi_svd<3>(A, V, sv);
}

// U.B0 of a fitted UBI: the cell of UBI.UBI^T, its Busing-Levy B and
// U = (B.UBI)^T (ImageD11/sinograms/point_by_point.py:443-490).
__device__ __forceinline__ void gs_ubi_u_b0(const double* u, const double* B0, double* ub0) {
    const double r2d = 180.0 / GS_PI, d2r = GS_PI / 180.0;
    double G[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            G[3 * i + j] = u[3 * i] * u[3 * j] + u[3 * i + 1] * u[3 * j + 1] +
                           u[3 * i + 2] * u[3 * j + 2];
    const double a = sqrt(G[0]), b = sqrt(G[4]), c = sqrt(G[8]);
    const double al = acos(G[5] / b / c) * r2d, be = acos(G[2] / a / c) * r2d,
                 ga = acos(G[1] / a / b) * r2d;
    const double ca = cos(al * d2r), cb = cos(be * d2r), cg = cos(ga * d2r);
    double g[9] = {a * a, a * b * cg, a * c * cb, a * b * cg, b * b, b * c * ca,
                   a * c * cb, b * c * ca, c * c};
    if (gs_inverse3x3(g) != 0) {
        for (int i = 0; i < 9; i++) ub0[i] = __builtin_nan("");
        return;
    }
    const double as = sqrt(g[0]), bs = sqrt(g[4]), cs = sqrt(g[8]);
    const double betas = acos(g[2] / as / cs) * r2d, gammas = acos(g[1] / as / bs) * r2d;
    const double B[9] = {as, bs * cos(gammas * d2r), cs * cos(betas * d2r),
                         0.0, bs * sin(gammas * d2r), -cs * sin(betas * d2r) * ca,
                         0.0, 0.0, 1.0 / c};
    double U[9];  // (B.u)^T
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            U[3 * j + i] = B[3 * i] * u[j] + B[3 * i + 1] * u[3 + j] + B[3 * i + 2] * u[6 + j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            ub0[3 * i + j] = U[3 * i] * B0[j] + U[3 * i + 1] * B0[3 + j] + U[3 * i + 2] * B0[6 + j];
}

struct gs_mat9 {
    double m[9];
};

// The UBI fit of refine_map (weighted_lstsq_ubi_fit, :408-419, and the
// quality test of :1595-1597) for every task with fit[t] != 0: rows
// design[r] = (w g, w hkl) for r in [off[t], off[t+1]).  Outputs per task:
// ubifit (NaN when not fitted), lstsq's rank, worth = rank 3, cond < 1e14,
// det > 0 and matrix_rank 3; ubi_out = ubifit when worth, else ubi_in; and
// ub0 = U.B0 of ubifit when worth (NaN otherwise).
__global__ __launch_bounds__(64) void k_pbp_ubi_fit(
    const double* __restrict__ design, const long long* __restrict__ off,
    const int* __restrict__ fit, const double* __restrict__ ubi_in, gs_mat9 B0, int T,
    double* __restrict__ ubifit, int* __restrict__ rank_out, int* __restrict__ worth_out,
    double* __restrict__ ubi_out, double* __restrict__ ub0) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= T) return;
    double u[9], uo[9], b0[9];
    for (int i = 0; i < 9; i++) {
        u[i] = __builtin_nan("");
        uo[i] = ubi_in[9 * (size_t)t + i];
        b0[i] = __builtin_nan("");
    }
    int rank = 0, worth = 0;
    if (fit[t]) {
        double R[3][6];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 6; j++) R[i][j] = 0.0;
        const long long r0 = off[t], r1 = off[t + 1];
        gs_qr_add_rows<3, 3, 8>(R, design, r0, r1);
        const double n = (double)(r1 - r0);
        double X[3][3];  // ubifit^T
        rank = gs_lstsq_solve<3, 3>(R, 2.220446049250313e-16 * fmax(n, 3.0), X);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) u[3 * i + j] = X[j][i];
        double sv[3];
        gs_sv3(u, sv);
        const double smax = fmax(sv[0], fmax(sv[1], sv[2]));
        const double smin = fmin(sv[0], fmin(sv[1], sv[2]));
        const double cond = smax / smin;  // inf for a singular matrix, NaN with NaN
        int mrank = 0;
        for (int j = 0; j < 3; j++) mrank += sv[j] > smax * 3.0 * 2.220446049250313e-16;
        const double det = u[0] * (u[4] * u[8] - u[5] * u[7]) -
                           u[1] * (u[3] * u[8] - u[5] * u[6]) +
                           u[2] * (u[3] * u[7] - u[4] * u[6]);
        worth = rank == 3 && cond < 1e14 && det > 0.0 && mrank == 3;
        if (worth) {
            for (int i = 0; i < 9; i++) uo[i] = u[i];
            gs_ubi_u_b0(u, B0.m, b0);
        }
    }
    for (int i = 0; i < 9; i++) {
        ubifit[9 * (size_t)t + i] = u[i];
        ubi_out[9 * (size_t)t + i] = uo[i];
        ub0[9 * (size_t)t + i] = b0[i];
    }
    rank_out[t] = rank;
    worth_out[t] = worth;
}

// The strain fit of refine_map (:1652-1656) for every task with
// worth[t] != 0: rows design[r] = (w M, w strain) for r in [off[t],
// off[t+1]), eps[t] = lstsq's minimum-norm solution (NaN for other tasks).
__global__ __launch_bounds__(64) void k_pbp_strain_fit(const double* __restrict__ design,
                                                       const long long* __restrict__ off,
                                                       const int* __restrict__ worth, int T,
                                                       double* __restrict__ eps) {
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= T) return;
    double e[6];
    for (int i = 0; i < 6; i++) e[i] = __builtin_nan("");
    if (worth[t]) {
        double R[6][7];
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = 0; j < 7; j++) R[i][j] = 0.0;
        const long long r0 = off[t], r1 = off[t + 1];
        gs_qr_add_rows<6, 1, 4>(R, design, r0, r1);
        const double n = (double)(r1 - r0);
        double X[6][1];
        gs_lstsq_solve<6, 1>(R, 2.220446049250313e-16 * fmax(n, 6.0), X);
        for (int i = 0; i < 6; i++) e[i] = X[i][0];
    }
    for (int i = 0; i < 6; i++) eps[6 * (size_t)t + i] = e[i];
}

// Row products of refine_map's merged-peak scoring and strain fit: per row
// r, h = U g with U = ubis[task[r]] (or row r's own 3x3 when task is
// NULL), each component ((u_i0 g_0 + u_i1 g_1) + u_i2 g_2) with every
// product and sum rounded on its own (ubi.dot(gve_voxel_merged) and
// U.dot(B0).dot(hkl), ImageD11/sinograms/point_by_point.py:1582,1611,1639,
// in the row order and arithmetic of the device implementation they
// replace).  With hkl given: hkl = h rounded half to even (np.round) and
// ok = ((e_0^2 + e_1^2) + e_2^2) < tolsq for e = h - hkl (:1581-1590).
__global__ __launch_bounds__(GS_BLOCK) void k_pbp_rows_dot(
    const double* __restrict__ ubis, const long long* __restrict__ task, long long R,
    const double* __restrict__ g, double tolsq, double* __restrict__ h,
    long long* __restrict__ hkl, unsigned char* __restrict__ ok) {
    const long long r = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (r >= R) return;
    const double* u = ubis + 9 * (size_t)(task ? task[r] : r);
    const double g0 = g[3 * r], g1 = g[3 * r + 1], g2 = g[3 * r + 2];
    double e2 = 0.0;
    for (int i = 0; i < 3; i++) {
        const double a = u[3 * i] * g0;
        const double b = u[3 * i + 1] * g1;
        const double c = u[3 * i + 2] * g2;
        const double v = (a + b) + c;
        if (h) h[3 * r + i] = v;
        if (hkl) {
            const double k = rint(v);
            const double e = v - k;
            e2 = i == 0 ? e * e : e2 + e * e;
            hkl[3 * r + i] = (long long)k;
        }
    }
    if (ok) ok[r] = e2 < tolsq;
}
