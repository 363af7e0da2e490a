// Indexer kernels: ring assignment, the all-pairs cosine search of
// indexer.find, and batched trial orientations (unitcell.orient +
// quickorient) whose UBIs k_score_many then scores in one launch.
#pragma once
#include "gs_device.h"

#define GS_MAX_COSES 4096  // cosine table held in LDS by k_find_hits / k_find_all
#define GS_FIND_ROWS 8     // ring-1 peaks per block of k_find_hits

// quickorient (src/cdiffraction.c:293-328).  On entry rows 0 and 1 of u
// hold two observed g-vectors; bt = B^-1 [t1 t2 t3] of the hkl pair
// (unitcell.BTmat).  The observed frame is t1 = g1/|g1|,
// t3 = g1 x g2 / |g1 x g2|, t2 = t1 x t3, and u becomes bt . [t1; t2; t3].
// Operation order as in the reference (bit parity with its C).
GS_HD inline void gs_quickorient(double* u, const double* bt) {
    const double ax = u[0], ay = u[1], az = u[2];
    const double bx = u[3], by = u[4], bz = u[5];
    double t3x = ay * bz - az * by;
    double t3y = az * bx - ax * bz;
    double t3z = ax * by - ay * bx;
    const double na = sqrt(ax * ax + ay * ay + az * az);
    const double t1x = ax / na, t1y = ay / na, t1z = az / na;
    const double n3 = sqrt(t3x * t3x + t3y * t3y + t3z * t3z);
    t3x /= n3;
    t3y /= n3;
    t3z /= n3;
    const double m[9] = {t1x, t1y, t1z,
                         t1y * t3z - t1z * t3y, t1z * t3x - t1x * t3z, t1x * t3y - t1y * t3x,
                         t3x, t3y, t3z};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            u[3 * r + c] = bt[3 * r] * m[c] + bt[3 * r + 1] * m[3 + c] + bt[3 * r + 2] * m[6 + c];
}

// |g| in numpy's order for np.sqrt((gv * gv).sum(axis=1))
// (ImageD11/indexing.py:345, :696): ((x*x + y*y) + z*z).
GS_HD inline double gs_norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// The two 3-term dot products of the indexer as numpy evaluates them
// through OpenBLAS 0.3.29 (x86-64 Haswell/SkylakeX kernels, the build of
// numpy 2.2 here): row-of-matrix times vector (dgemv, np.dot(n2, n1[i]),
// ImageD11/indexing.py:711) is fma(x2,y2, fma(x0,y0, x1*y1)); vector
// times vector (ddot, np.dot(g1, g2), ImageD11/unitcell.py:618) is
// fma(x2,y2, fma(x1,y1, x0*y0)).  Both verified bit-for-bit against numpy.
GS_HD inline double gs_dot3_gemv(double x0, double x1, double x2, double y0, double y1,
                                 double y2) {
    return fma(x2, y2, fma(x0, y0, x1 * y1));
}
GS_HD inline double gs_dot3_ddot(const double* x, const double* y) {
    return fma(x[2], y[2], fma(x[1], y[1], x[0] * y[0]));
}

// closest's running minimum (src/closest.c:188-207): strict '<' while
// scanning x-major, so the result is the smallest x index reaching the
// global minimum.  Merging two partial states keeps that rule.
GS_HD inline void gs_best_merge(double& d, int& i, double d2, int i2) {
    if (d2 < d || (d2 == d && i2 < i)) {
        d = d2;
        i = i2;
    }
}

// Block-wide merge of (d, i) held per thread; result in thread 0.
template <int NT>
__device__ inline void gs_block_best(double* sd, int* si, double& d, int& i) {
    const int t = threadIdx.x;
    sd[t] = d;
    si[t] = i;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) {
            double dd = sd[t];
            int ii = si[t];
            gs_best_merge(dd, ii, sd[t + s], si[t + s]);
            sd[t] = dd;
            si[t] = ii;
        }
        __syncthreads();
    }
    d = sd[0];
    i = si[0];
}

// closest(x, v) for one call: one block, threads stride over x.
__global__ __launch_bounds__(GS_BLOCK) void k_closest(const double* __restrict__ x, int nx,
                                                      const double* __restrict__ v, int nv,
                                                      int* __restrict__ ibest,
                                                      double* __restrict__ best) {
    __shared__ double sd[GS_BLOCK];
    __shared__ int si[GS_BLOCK];
    double d = 99.0;
    int ib = 0;
    for (int i = threadIdx.x; i < nx; i += GS_BLOCK) {
        const double xi = x[i];
        for (int j = 0; j < nv; j++) {
            const double e = fabs(xi - v[j]);
            if (e < d) {
                d = e;
                ib = i;
            }
        }
    }
    gs_block_best<GS_BLOCK>(sd, si, d, ib);
    if (threadIdx.x == 0) {
        *ibest = ib;
        *best = d;
    }
}

// One-thread quickorient for the host prototype.
__global__ void k_quickorient(double* __restrict__ u, const double* __restrict__ bt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) gs_quickorient(u, bt);
}

// ds = |g| per peak (indexer.__init__, ImageD11/indexing.py:345).
__global__ void k_gv_ds(const double* __restrict__ gv, int ng, double* __restrict__ ds) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ng) ds[k] = gs_norm3(gv[3 * k], gv[3 * k + 1], gv[3 * k + 2]);
}

// indexer.assigntorings (ImageD11/indexing.py:447-457): the ring j with the
// smallest |ds - ringds[j]| below tol, the first one on ties, else -1.
__global__ void k_ring_assign(const double* __restrict__ ds, int n,
                              const double* __restrict__ ringds, int nr, double tol,
                              int* __restrict__ ra) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double d = ds[k];
    double best = tol;
    int r = -1;
    for (int j = 0; j < nr; j++) {
        const double e = fabs(d - ringds[j]);
        if (e < best) {
            r = j;
            best = e;
        }
    }
    ra[k] = r;
}

// Unit g-vectors of the peaks idx[] (indexer.find, ImageD11/indexing.py:
// 696-706: n = g / |g| componentwise).  Out-of-range indices give NaN rows.
__global__ void k_unit_rows(const double* __restrict__ gv, int ng, const int* __restrict__ idx,
                            int n, double* __restrict__ out) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int k = idx[r];
    double x = __builtin_nan(""), y = x, z = x;
    if (k >= 0 && k < ng) {
        x = gv[3 * (long long)k];
        y = gv[3 * (long long)k + 1];
        z = gv[3 * (long long)k + 2];
        const double m = gs_norm3(x, y, z);
        x = x / m;
        y = y / m;
        z = z / m;
    }
    out[3 * r] = x;
    out[3 * r + 1] = y;
    out[3 * r + 2] = z;
}

// indexer.find with cosine_tol > 0 (ImageD11/indexing.py:710-716): for
// each ring-1 peak i, costheta_j = n2[j] . n1[i] over the ring-2 peaks, and
// closest(costheta, cs) -> (ibest[i], best[i]).  A block owns
// GS_FIND_ROWS rows; its threads stride over j, each n2[j] is read once per
// block and used for every row; the cosine table sits in LDS.
__global__ __launch_bounds__(GS_BLOCK) void k_find_hits(
    const double* __restrict__ n1, int N1, const double* __restrict__ n2, int N2,
    const double* __restrict__ cs, int nv, int* __restrict__ ibest, double* __restrict__ best) {
    __shared__ double lcs[GS_MAX_COSES];
    __shared__ double sd[GS_BLOCK];
    __shared__ int si[GS_BLOCK];
    for (int m = threadIdx.x; m < nv; m += GS_BLOCK) lcs[m] = cs[m];
    const int r0 = blockIdx.x * GS_FIND_ROWS;
    double ax[GS_FIND_ROWS], ay[GS_FIND_ROWS], az[GS_FIND_ROWS], d[GS_FIND_ROWS];
    int ib[GS_FIND_ROWS];
#pragma unroll
    for (int r = 0; r < GS_FIND_ROWS; r++) {
        const int row = min(r0 + r, N1 - 1);
        ax[r] = n1[3 * (long long)row];
        ay[r] = n1[3 * (long long)row + 1];
        az[r] = n1[3 * (long long)row + 2];
        d[r] = 99.0;
        ib[r] = 0;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < N2; j += GS_BLOCK) {
        const double bx = n2[3 * (long long)j], by = n2[3 * (long long)j + 1],
                     bz = n2[3 * (long long)j + 2];
#pragma unroll
        for (int r = 0; r < GS_FIND_ROWS; r++) {
            // np.dot(n2, n1[i]): one row of n2 against n1[i]
            const double c = gs_dot3_gemv(bx, by, bz, ax[r], ay[r], az[r]);
            double e = 99.0;
            for (int m = 0; m < nv; m++) e = fmin(e, fabs(c - lcs[m]));
            if (e < d[r]) {  // e is NaN-free: fmin drops NaN operands
                d[r] = e;
                ib[r] = j;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < GS_FIND_ROWS; r++) {
        double dd = d[r];
        int ii = ib[r];
        gs_block_best<GS_BLOCK>(sd, si, dd, ii);
        if (threadIdx.x == 0 && r0 + r < N1) {
            ibest[r0 + r] = ii;
            best[r0 + r] = dd;
        }
        __syncthreads();
    }
}

// indexer.find with cosine_tol < 0 (ImageD11/indexing.py:717-723): every
// (i, m, j) with |cs[m] - costheta_ij| < mtol is a hit [0, i1[i], i2[j]],
// listed i-major, then m, then j.  Matches are counted into *counter and,
// when keys is given, written as key = (i*nv + m)*N2 + j at atomically
// taken slots (< cap); sorting the keys restores the reference order.
__global__ __launch_bounds__(GS_BLOCK) void k_find_all(
    const double* __restrict__ n1, int N1, const double* __restrict__ n2, int N2,
    const double* __restrict__ cs, int nv, double mtol, unsigned long long* __restrict__ counter,
    long long* __restrict__ keys, long long cap) {
    __shared__ double lcs[GS_MAX_COSES];
    for (int m = threadIdx.x; m < nv; m += GS_BLOCK) lcs[m] = cs[m];
    __syncthreads();
    const int i = blockIdx.x;
    if (i >= N1) return;
    const double ax = n1[3 * (long long)i], ay = n1[3 * (long long)i + 1],
                 az = n1[3 * (long long)i + 2];
    unsigned long long local = 0;
    for (int j = threadIdx.x; j < N2; j += GS_BLOCK) {
        const double c = gs_dot3_gemv(n2[3 * (long long)j], n2[3 * (long long)j + 1],
                                      n2[3 * (long long)j + 2], ax, ay, az);
        for (int m = 0; m < nv; m++) {
            if (fabs(lcs[m] - c) < mtol) {
                if (keys) {
                    const unsigned long long at = atomicAdd(counter, 1ull);
                    if ((long long)at < cap) keys[at] = ((long long)i * nv + m) * N2 + j;
                } else {
                    local++;
                }
            }
        }
    }
    if (!keys && local) atomicAdd(counter, local);
}

// unitcell.orient (ImageD11/unitcell.py:617-653) for many (i, j) peak
// pairs: costheta = g1.g2 / sqrt(|g1|^2 |g2|^2) and UBI = quickorient([g1;
// g2], bt[b]).  Without bsel, b is the nearest entry of the sorted table
// c2ab (searchsorted 'left', then the lower neighbour when strictly
// closer: the crange <= 0 case); with bsel, b = bsel[h] (one of the
// crange > 0 variants).  choice[h] = b and costheta[h] when given.
// Out-of-range peak or table indices give a NaN UBI and choice -1.
__global__ void k_orient_hits(const double* __restrict__ gv, int ng, const int* __restrict__ hi,
                              const int* __restrict__ hj, int nh, const double* __restrict__ c2ab,
                              int nc, const double* __restrict__ bt, const int* __restrict__ bsel,
                              double* __restrict__ ubis, int* __restrict__ choice,
                              double* __restrict__ costheta) {
    const long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const int i = hi[h], j = hj[h];
    double u[9];
    int b = -1;
    double ct = __builtin_nan("");
    const int bfix = bsel ? bsel[h] : 0;
    if (i >= 0 && i < ng && j >= 0 && j < ng && bfix >= 0 && bfix < nc) {
        for (int c = 0; c < 3; c++) {
            u[c] = gv[3 * (long long)i + c];
            u[3 + c] = gv[3 * (long long)j + c];
        }
        const double dot = gs_dot3_ddot(u, u + 3);
        const double s1 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double s2 = u[3] * u[3] + u[4] * u[4] + u[5] * u[5];
        ct = dot / sqrt(s1 * s2);
        int lo = 0;
        if (bsel) {
            lo = nc;  // unused: the table entry is given
        } else if (ct != ct) {
            lo = nc;  // numpy orders NaN after every number
        } else {
            int hi_ = nc;
            while (lo < hi_) {
                const int mid = (lo + hi_) >> 1;
                if (c2ab[mid] < ct)
                    lo = mid + 1;
                else
                    hi_ = mid;
            }
        }
        b = lo;
        if (bsel)
            b = bfix;
        else if (lo > 0 && (lo == nc || fabs(ct - c2ab[lo - 1]) < fabs(ct - c2ab[lo])))
            b = lo - 1;
        u[6] = u[7] = u[8] = 0.0;
        gs_quickorient(u, bt + 9 * (long long)b);
    } else {
        for (int c = 0; c < 9; c++) u[c] = __builtin_nan("");
    }
    for (int c = 0; c < 9; c++) ubis[9 * h + c] = u[c];
    if (choice) choice[h] = b;
    if (costheta) costheta[h] = ct;
}

// Peaks each of K UBIs indexes, decided as the reference's arithmetic
// decides them: the pairs with gs_drlv2 < thr (thr = tol^2 for
// score_and_refine's selection, min(tol^2, 1) for getind's score_and_assign
// from drlv2 = 1).  cursor[g] (zeroed by the caller) counts grain g's
// peaks; when idx is given they are also written at idx[off[g] + slot] for
// slot < off[g+1] - off[g], in no particular order.  Tiles as
// k_score_many_fast: P peaks per lane in registers, a split of gsplit
// grains per block, the certified fract form against thr - 3/4 +- E
// (gs_cert_window_fract, E per grain and block in LDS), the undecided
// pairs settled with gs_drlv2 itself; lane 0 of a wave takes the wave's
// slots with one atomic per grain.
template <int P>
__global__ __launch_bounds__(GS_BLOCK) void k_select_many(
    const double* __restrict__ ubis, int K, int gsplit, int nsplit, const double* __restrict__ gv,
    int ng, double thr, const long long* __restrict__ off, int* __restrict__ cursor,
    int* __restrict__ idx) {
    __shared__ double lwin[2 * GS_SPLIT_MAX];  // (lo, hi) per grain of the split
    __shared__ double lgm[GS_BLOCK / GS_WAVE];
    const int tile = gs_xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile / nsplit;
    const int split = tile - chunk * nsplit;
    const int g0 = split * gsplit;
    const int g1 = min(K, g0 + gsplit);
    double x[P], y[P], z[P];
    double gm = 0.0;
    const long long base = (long long)chunk * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            x[p] = gv[3 * k + 0];
            y[p] = gv[3 * k + 1];
            z[p] = gv[3 * k + 2];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");  // never selected
        }
        gm = fmax(gm, fmax(gs_abs_finite(x[p]), fmax(gs_abs_finite(y[p]), gs_abs_finite(z[p]))));
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    gm = gs_wave_max(gm);  // +inf if any peak of the wave is infinite
    if (lane == 0) lgm[threadIdx.x / GS_WAVE] = gm;
    __syncthreads();
    double gmax = lgm[0];
#pragma unroll
    for (int w = 1; w < GS_BLOCK / GS_WAVE; w++) gmax = fmax(gmax, lgm[w]);
    const double mid = thr - 0.75;  // gs_drlv2_fract is drlv2 - 3/4
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) {
        const double e = gs_cert_window_fract(gs_load_ubi(ubis + 9 * (size_t)(g0 + i)), gmax, thr);
        lwin[2 * i] = mid - e;
        lwin[2 * i + 1] = mid + e;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int g = g0; g < g1; g++) {
        const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)g);
        const double lo = lwin[2 * (g - g0)], hi = lwin[2 * (g - g0) + 1];
#pragma unroll
        for (int p = 0; p < P; p++) {
            const double s = gs_drlv2_fract(u, x[p], y[p], z[p]);
            bool pass = s < lo;
            const bool und = (s < hi) && !pass;
            if (__ballot(und))  // wave-uniform and rare: the reference arithmetic decides
                pass = pass || (und && gs_drlv2(u, x[p], y[p], z[p]) < thr);
            const unsigned long long m = __ballot(pass);
            if (m == 0) continue;
            int first = 0;
            if (lane == 0) first = atomicAdd(&cursor[g], __popcll(m));
            if (idx) {
                first = __shfl(first, 0, GS_WAVE);
                const long long cap = off[g + 1] - off[g];
                const long long slot = first + __popcll(m & below);
                if (pass && slot < cap) idx[off[g] + slot] = (int)(base + (long long)p * GS_BLOCK);
            }
        }
    }
}
