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

// Per-segment column sums, each summed row by row from 0.0 in row order:
// out[s][c] = sum of v[r][c] for r in [off[s], off[s+1]).  That is the
// order in which np.bincount(labels, weights) accumulates, so grouped
// weighted sums (refine_map's merge, :338-356) come out bit for bit as the
// reference's, and the same on every run.  One thread per (segment,
// column); consecutive threads read consecutive columns of a row.
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
    // loads run ahead in groups of eight; the additions stay in row order
    for (; r + 8 <= r1; r += 8) {
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = v[(r + j) * C + c];
#pragma unroll
        for (int j = 0; j < 8; j++) acc += x[j];
    }
    for (; r < r1; r++) acc += v[r * C + c];
    out[i] = acc;
}
