// HIP kernels of the grain-scoring path, written for gfx950 (wave64,
// FP64 on the vector ALUs, 256 CUs in 8 XCDs).
//
// Data layout in HBM: g-vectors and lab positions are (n, 3) row-major
// float64 exactly as the reference passes them (an array of 24-byte rows);
// UBIs are (K, 3, 3) row-major; per-peak drlv2 (float64) and labels (int32)
// are flat arrays; per-grain peak lists are CSR (int32 peak index, int64
// offsets).  All-pairs kernels keep UBIs wave-uniform so they live in
// scalar registers, hold a few peaks per lane in VGPRs, and count with a
// wave ballot + popcount feeding LDS counters, one global atomic per
// (block, grain) with a non-zero count.
#pragma once
#include "gs_device.h"

#define GS_BLOCK 256
#define GS_GRAIN_CHUNK 1024  // grains per LDS counter tile
#define GS_SPLIT_MAX 512     // grains per split of the all-pairs score/select tiles
#define GS_SIEVE_SPLIT_MAX 128  // grains per split of k_score_many_sieve (UBI table in LDS)
#define GS_SIEVE_THR_MAX 0.01  // gs_select_many takes the sieve for thr <= 0.01 (tol <= 0.1)
#ifndef GS_SIEVE_TILES_PER_CU
#define GS_SIEVE_TILES_PER_CU 64  // (chunk, split) tiles per CU of k_score_many_sieve (32: 0.413 ms, 64: 0.403, 128: 0.416, 256: 0.447 at 1000 x 1e6)
#endif
#define GS_NPART 17          // n, sum drlv2, R[9], H[6]
#ifndef GS_REFINE_BATCH
#define GS_REFINE_BATCH 8    // list rows per lane in flight in k_refine_lists (rows mode)
#endif

// ---------------------------------------------------------------------------
// small utility kernels (used instead of hipMemset so every launch is
// graph-capture safe)
__global__ void k_fill_i32(int* __restrict__ p, int v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ p, double v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// verify_rounding on the device (src/closest.c:100-119): 600 integer-valued
// points, MAGIC rounding vs floor(x + 0.5).
__global__ void k_verify_rounding(int n, int* __restrict__ bad) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 600) return;
    int base = t < 300 ? n : -n;
    int i = (t % 300) - 100;
    double v = base + i * 50.0;
    int a = (int)gs_round(v);
    int b = (int)floor(v + 0.5);
    if (a != b) atomicAdd(bad, 1);
}

// ---------------------------------------------------------------------------
// score for K UBIs x ng g-vectors -> counts[K] (counts must be zeroed).
// Logical tile = (peak chunk of GS_BLOCK*P rows, grain split of gsplit
// UBIs); the tiles of one chunk are adjacent after the XCD remap so the
// chunk is fetched into one XCD's L2.
template <int P>
__global__ __launch_bounds__(GS_BLOCK) void k_score_many(const double* __restrict__ ubis, int K,
                                                         int gsplit, int nsplit,
                                                         const double* __restrict__ gv, int ng,
                                                         double tolsq, int* __restrict__ counts) {
    __shared__ int lcnt[GS_SPLIT_MAX];
    const int tile = gs_xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile / nsplit;
    const int split = tile - chunk * nsplit;
    const int g0 = split * gsplit;
    const int g1 = min(K, g0 + gsplit);
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) lcnt[i] = 0;

    double x[P], y[P], z[P];
    const long long base = (long long)chunk * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            x[p] = gv[3 * k + 0];
            y[p] = gv[3 * k + 1];
            z[p] = gv[3 * k + 2];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");  // never counted
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & (GS_WAVE - 1);
    for (int g = g0; g < g1; g++) {
        const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)g);
        int c = 0;
#pragma unroll
        for (int p = 0; p < P; p++) {
            double s = gs_drlv2(u, x[p], y[p], z[p]);
            c += __popcll(__ballot(s < tolsq));
        }
        if (lane == 0 && c) atomicAdd(&lcnt[g - g0], c);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) {
        int c = lcnt[i];
        if (c) atomicAdd(&counts[g0 + i], c);
    }
}

// ---------------------------------------------------------------------------
// Certified fast score: the same counts as k_score_many for every input.
//
// The kernel evaluates gs_drlv2_fract (rows h + 1/2 by FMA, fract, and
// sum f (f - 1) = drlv2 - 3/4: 19 instructions per pair with the two
// compares) against tol^2 - 3/4 +- gs_cert_window_fract; the bound below
// is stated for the plain FMA form gs_drlv2_fma, which the certified assign
// kernel uses, and carries over with the extra 1/2 per row (see
// gs_cert_window_fract).  Let S_e be the sum the
// reference computes (separately rounded products, MAGIC rounding) and S_f
// this one.  With A_i = sum_j |u_ij| |g_j|, each h differs from the exact
// dot product by at most gamma_3 A_i (gamma_3 = 3u/(1-3u), u = 2^-53), so
// |h_f - h_e| <= 6.7e-16 A_i.  The
// distance to the nearest integer is 1-Lipschitz and at most 1/2, so
// x -> (x - round x)^2 changes by at most |dx| (it is continuous across
// the half-integers): the rows move S by at most 6.7e-16 sum_i A_i <=
// 6.7e-16 U1 Gmax, with U1 = sum |u_ij| and Gmax = max |g_j| over the
// block's peaks.  Both sums of squares are <= 3/4, so rounding them adds at
// most 5e-16 in absolute terms.  E = 1e-15 U1 Gmax + 1e-15 + 2e-15 tol^2
// covers all of it with margin (the last term absorbs the rounding of
// tol^2 +- E).  S_f < tol^2 - E proves S_e < tol^2; S_f >= tol^2 + E proves
// the opposite; lanes in between are recomputed with the reference
// arithmetic.
// v_rndne equals the MAGIC rounding for |h| < 2^51, so a (block, grain)
// with U1 Gmax >= 2^49 (or a non-finite g) takes the exact path entirely.
// The window (tol^2 - E, tol^2 + E) of each grain is computed once per
// block into LDS, so the grain loop spends its VALU issue on the pairs.
__device__ __forceinline__ double gs_wave_max(double v) {
    for (int o = GS_WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, GS_WAVE));
    return v;
}

// At 4 peaks per lane the kernel fits 64 VGPRs: 8 waves per SIMD.
template <int P>
__global__ __launch_bounds__(GS_BLOCK, P <= 4 ? 8 : 1) void k_score_many_fast(
    const double* __restrict__ ubis, int K, int gsplit, int nsplit,
    const double* __restrict__ gv, int ng, double tolsq, int* __restrict__ counts) {
    __shared__ int lcnt[GS_SPLIT_MAX];
    __shared__ double lwin[2 * GS_SPLIT_MAX];  // (lo, hi) per grain of the split
    __shared__ double lgm[GS_BLOCK / GS_WAVE];
    const int tile = gs_xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile / nsplit;
    const int split = tile - chunk * nsplit;
    const int g0 = split * gsplit;
    const int g1 = min(K, g0 + gsplit);
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) lcnt[i] = 0;

    double x[P], y[P], z[P];
    double gm = 0.0;
    const long long base = (long long)chunk * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            x[p] = gv[3 * k + 0];
            y[p] = gv[3 * k + 1];
            z[p] = gv[3 * k + 2];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");
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
    const double mid = tolsq - 0.75;  // gs_drlv2_fract is drlv2 - 3/4
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) {
        const double e =
            gs_cert_window_fract(gs_load_ubi(ubis + 9 * (size_t)(g0 + i)), gmax, tolsq);
        lwin[2 * i] = mid - e;
        lwin[2 * i + 1] = mid + e;
    }
    __syncthreads();
    gs_ubi un = gs_load_ubi(ubis + 9 * (size_t)g0);
    for (int g = g0; g < g1; g++) {
        const gs_ubi u = un;
        if (g + 1 < g1) un = gs_load_ubi(ubis + 9 * (size_t)(g + 1));  // prefetch
        const double lo = lwin[2 * (g - g0)], hi = lwin[2 * (g - g0) + 1];
        int c = 0;
        unsigned long long undecided = 0;
#pragma unroll
        for (int p = 0; p < P; p++) {
            const double s = gs_drlv2_fract(u, x[p], y[p], z[p]);
            const unsigned long long in = __ballot(s < lo);
            undecided |= __ballot(s < hi) ^ in;
            c += __popcll(in);
        }
        if (undecided) {  // wave-uniform and rare: settle undecided lanes exactly
#pragma unroll
            for (int p = 0; p < P; p++) {
                const double s = gs_drlv2_fract(u, x[p], y[p], z[p]);
                const bool und = (s < hi) && !(s < lo);
                const double se = gs_drlv2(u, x[p], y[p], z[p]);
                c += __popcll(__ballot(und && se < tolsq));
            }
        }
        if (lane == 0 && c) atomicAdd(&lcnt[g - g0], c);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < g1 - g0; i += GS_BLOCK) {
        int c = lcnt[i];
        if (c) atomicAdd(&counts[g0 + i], c);
    }
}

// ---------------------------------------------------------------------------
// Sieved score: the counts of k_score_many for every input, with most pairs
// decided by one row.
//
// At tol = 0.05 a random pair has |h0 - round h0| < tol with probability
// ~0.1, and any pair outside that band is not counted (gs_sieve_radius).
// Stage one evaluates only row 0 for each pair of the grain step: y = h0 + r
// by three FMAs (r enters the first), and the pair passes unless
// fract(y) >= 2r, i.e. unless h0 is at least r from every integer (five
// FP64 instructions; a NaN passes); the survivors (grain, peak) are
// appended to per-wave queues in LDS, one per peak slot p of the lanes
// (ballot + mbcnt positions).  Whenever a queue holds a wave's worth, 64
// survivors are settled at once, one per lane, by the full certified test
// of k_score_many_fast (its window in LDS, the reference arithmetic for
// undecided pairs) with the UBI from an LDS table and the g-vector re-read
// through the cache.  A queue is drained from its top, so entries never
// move; capacity 128 covers the at most 63 left over plus one grain step's
// 64.  Survivors are counted with LDS atomics per grain (rare: only true
// hits), then one global atomic per (block, grain).
//
// Queue entry (16 bits): (grain offset in the split) << 6 | lane; the slot
// p is the queue's, so a grain step forms one entry value for all its
// slots.
//
// SELECT = true is the sieved k_select_many: a settled pair that passes takes
// a slot of its grain with one global atomic on cursor[g] (passing pairs are
// the rare true hits) and, when idx is given, writes its peak index at
// idx[off[g] + slot] for slot < off[g+1] - off[g].
// kbase: the first peak (in the chunk) of the queue's slot for this wave,
// p GS_BLOCK + 64 wave.
template <bool SELECT>
__device__ __forceinline__ void gs_sieve_settle(int e, const double* __restrict__ gvc, long long nrem,
                                                int kbase, const double* lubi, const double* lwin,
                                                double tolsq, int* lcnt, int g0, long long c0,
                                                const long long* __restrict__ off,
                                                int* __restrict__ cursor, int* __restrict__ idx) {
    const int gl = e >> 6;
    const long long k = kbase + (e & 63);
    if (k >= nrem) return;  // NaN padding rows of the last chunk
    const double x = gvc[3 * k + 0], y = gvc[3 * k + 1], z = gvc[3 * k + 2];
    const double* ur = lubi + 9 * gl;
    const gs_ubi u = {ur[0], ur[1], ur[2], ur[3], ur[4], ur[5], ur[6], ur[7], ur[8]};
    const double s = gs_drlv2_fract(u, x, y, z);
    bool in = s < lwin[2 * gl];
    if (!in && s < lwin[2 * gl + 1]) in = gs_drlv2(u, x, y, z) < tolsq;  // undecided: exact
    if (!in) return;
    if constexpr (SELECT) {
        const int g = g0 + gl;
        const int slot = atomicAdd(&cursor[g], 1);
        if (idx && slot < off[g + 1] - off[g]) idx[off[g] + slot] = (int)(c0 + k);
    } else {
        atomicAdd(&lcnt[gl], 1);
    }
}

template <int P, bool SELECT = false>
__global__ __launch_bounds__(GS_BLOCK, 8) void k_score_many_sieve(
    const double* __restrict__ ubis, int K, int gsplit, int nsplit,
    const double* __restrict__ gv, int ng, double tolsq, int* __restrict__ counts,
    const long long* __restrict__ off = nullptr, int* __restrict__ cursor = nullptr,
    int* __restrict__ idx = nullptr) {
    static_assert(P >= 1 && P <= 8, "at most 8 peak slots per lane");
    static_assert(GS_SIEVE_SPLIT_MAX <= 1024, "a 16-bit queue entry holds the grain offset");
    constexpr int NW = GS_BLOCK / GS_WAVE;
    constexpr int QCAP = 2 * GS_WAVE;  // per slot: 63 left over + one grain step's 64
    __shared__ int lcnt[GS_SIEVE_SPLIT_MAX];
    __shared__ double lubi[9 * GS_SIEVE_SPLIT_MAX];
    __shared__ double lwin[2 * GS_SIEVE_SPLIT_MAX];  // (lo, hi) per grain of the split
    __shared__ double lrad[2 * GS_SIEVE_SPLIT_MAX];  // (r, 2r) per grain of the split
    __shared__ double lgm[NW];
    __shared__ unsigned short lq[NW * P * QCAP];
    const int tile = gs_xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile / nsplit;
    const int split = tile - chunk * nsplit;
    const int g0 = split * gsplit;
    const int ns = min(K, g0 + gsplit) - g0;  // grains of this split, <= GS_SIEVE_SPLIT_MAX

    double x[P], y[P], z[P];
    double gm = 0.0;
    const long long c0 = (long long)chunk * (GS_BLOCK * P);  // first peak of the chunk
    const long long nrem = (long long)ng - c0;
#pragma unroll
    for (int p = 0; p < P; p++) {
        const long long k = (long long)p * GS_BLOCK + threadIdx.x;
        if (k < nrem) {
            x[p] = gv[3 * (c0 + k) + 0];
            y[p] = gv[3 * (c0 + k) + 1];
            z[p] = gv[3 * (c0 + k) + 2];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");  // passes the sieve, skipped when settled
        }
        gm = fmax(gm, fmax(gs_abs_finite(x[p]), fmax(gs_abs_finite(y[p]), gs_abs_finite(z[p]))));
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    const int wave = threadIdx.x / GS_WAVE;
    gm = gs_wave_max(gm);
    if (lane == 0) lgm[wave] = gm;
    __syncthreads();
    double gmax = lgm[0];
#pragma unroll
    for (int w = 1; w < NW; w++) gmax = fmax(gmax, lgm[w]);
    const double mid = tolsq - 0.75;  // gs_drlv2_fract is drlv2 - 3/4
    for (int i = threadIdx.x; i < ns; i += GS_BLOCK) {
        const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)(g0 + i));
        const double e = gs_cert_window_fract(u, gmax, tolsq);
        lcnt[i] = 0;
        lwin[2 * i] = mid - e;
        lwin[2 * i + 1] = mid + e;
        const double r = gs_sieve_radius(u, gmax, tolsq);
        lrad[2 * i] = r;
        lrad[2 * i + 1] = 2.0 * r;
        double* ur = lubi + 9 * i;
        ur[0] = u.u00; ur[1] = u.u01; ur[2] = u.u02;
        ur[3] = u.u10; ur[4] = u.u11; ur[5] = u.u12;
        ur[6] = u.u20; ur[7] = u.u21; ur[8] = u.u22;
    }
    __syncthreads();
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    unsigned short* q = lq + wv * (P * QCAP);  // scalar base; slot p's queue at q + p QCAP
    const double* gvc = gv + 3 * c0;
    int qn[P];  // wave-uniform fill of each slot's queue
#pragma unroll
    for (int p = 0; p < P; p++) qn[p] = 0;
    for (int gl = 0; gl < ns; gl++) {
        const double* ug = ubis + 9 * (size_t)(g0 + gl);
        const double u00 = ug[0], u01 = ug[1], u02 = ug[2];
        const double r = lrad[2 * gl], r2 = lrad[2 * gl + 1];
        const unsigned short e = (unsigned short)((gl << 6) | lane);  // this step's entry, every slot
#pragma unroll
        for (int p = 0; p < P; p++) {
            const double h = __builtin_fma(u02, z[p], __builtin_fma(u01, y[p], __builtin_fma(u00, x[p], r)));
            const bool sv = !(gs_fract(h) >= r2);
            const unsigned long long m = __ballot(sv);
            if (sv) {  // rank among the survivors, then the scalar top of the queue
                const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                    (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                (q + p * QCAP + qn[p])[pos] = e;
            }
            qn[p] += __popcll(m);
        }
#pragma unroll
        for (int p = 0; p < P; p++) {
            if (qn[p] >= GS_WAVE) {  // settle the top 64 entries of slot p
                qn[p] -= GS_WAVE;
                __builtin_amdgcn_wave_barrier();
                gs_sieve_settle<SELECT>(q[p * QCAP + qn[p] + lane], gvc, nrem,
                                        p * GS_BLOCK + wv * GS_WAVE, lubi, lwin, tolsq, lcnt, g0,
                                        c0, off, cursor, idx);
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int p = 0; p < P; p++)
        if (lane < qn[p])
            gs_sieve_settle<SELECT>(q[p * QCAP + lane], gvc, nrem, p * GS_BLOCK + wv * GS_WAVE,
                                    lubi, lwin, tolsq, lcnt, g0, c0, off, cursor, idx);
    if constexpr (SELECT) return;
    __syncthreads();
    for (int i = threadIdx.x; i < ns; i += GS_BLOCK) {
        const int c = lcnt[i];
        if (c) atomicAdd(&counts[g0 + i], c);
    }
}

// ---------------------------------------------------------------------------
// score_and_assign for K grains in order (src/closest.c:366-397 called once
// per grain, as refinegrains.assignlabels and indexer.fight_over_peaks do).
// Each lane owns P peaks and runs every grain in order on their (drlv2,
// label) state, so the per-peak state machine is exactly the sequential
// one.  counts[g] (zeroed by the caller) = the value the g-th call returns.
template <int P>
__global__ __launch_bounds__(GS_BLOCK) void k_assign_many(
    const double* __restrict__ ubis, const int* __restrict__ glabels, int K,
    const double* __restrict__ gv, int ng, double tolsq, double* __restrict__ drlv2,
    int* __restrict__ labels, int* __restrict__ counts) {
    __shared__ int lcnt[GS_GRAIN_CHUNK];
    double x[P], y[P], z[P], d[P];
    int lab[P];
    const long long base = (long long)blockIdx.x * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            x[p] = gv[3 * k + 0];
            y[p] = gv[3 * k + 1];
            z[p] = gv[3 * k + 2];
            d[p] = drlv2[k];
            lab[p] = labels[k];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");
            d[p] = 0.0;
            lab[p] = -1;
        }
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    for (int c0 = 0; c0 < K; c0 += GS_GRAIN_CHUNK) {
        const int c1 = min(K, c0 + GS_GRAIN_CHUNK);
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) lcnt[i] = 0;
        __syncthreads();
        for (int g = c0; g < c1; g++) {
            const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)g);
            const int L = glabels[g];
            int c = 0;
#pragma unroll
            for (int p = 0; p < P; p++) {
                double s = gs_drlv2(u, x[p], y[p], z[p]);
                bool take = (s < tolsq) && (s < d[p]);
                if (take) {
                    lab[p] = L;
                    d[p] = s;
                } else if (lab[p] == L) {
                    lab[p] = -1;
                }
                c += __popcll(__ballot(take));
            }
            if (lane == 0 && c) atomicAdd(&lcnt[g - c0], c);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) {
            int c = lcnt[i];
            if (c) atomicAdd(&counts[c0 + i], c);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            drlv2[k] = d[p];
            labels[k] = lab[p];
        }
    }
}

// The same loop with the certified decision of k_score_many_fast: a pair
// whose fast drlv2 - 3/4 is >= tol^2 - 3/4 + E certainly fails
// `drlv2 < tol^2`, so only
// the reference's else branch (label == L -> -1) can apply; every other
// pair is recomputed with the reference arithmetic and the reference's
// rule is applied to that value, so labels, stored drlv2 and counts are
// those of k_assign_many.
template <int P>
__global__ __launch_bounds__(GS_BLOCK) void k_assign_many_fast(
    const double* __restrict__ ubis, const int* __restrict__ glabels, int K,
    const double* __restrict__ gv, int ng, double tolsq, double* __restrict__ drlv2,
    int* __restrict__ labels, int* __restrict__ counts) {
    __shared__ int lcnt[GS_GRAIN_CHUNK];
    __shared__ double lhi[GS_GRAIN_CHUNK];  // tol^2 - 3/4 + E per grain of the chunk
    __shared__ double lgm[GS_BLOCK / GS_WAVE];
    double x[P], y[P], z[P], d[P];
    int lab[P];
    double gm = 0.0;
    const long long base = (long long)blockIdx.x * (GS_BLOCK * P) + threadIdx.x;
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            x[p] = gv[3 * k + 0];
            y[p] = gv[3 * k + 1];
            z[p] = gv[3 * k + 2];
            d[p] = drlv2[k];
            lab[p] = labels[k];
        } else {
            x[p] = y[p] = z[p] = __builtin_nan("");
            d[p] = 0.0;
            lab[p] = -1;
        }
        gm = fmax(gm, fmax(gs_abs_finite(x[p]), fmax(gs_abs_finite(y[p]), gs_abs_finite(z[p]))));
    }
    const int lane = threadIdx.x & (GS_WAVE - 1);
    gm = gs_wave_max(gm);
    if (lane == 0) lgm[threadIdx.x / GS_WAVE] = gm;
    __syncthreads();
    double gmax = lgm[0];
#pragma unroll
    for (int w = 1; w < GS_BLOCK / GS_WAVE; w++) gmax = fmax(gmax, lgm[w]);
    for (int c0 = 0; c0 < K; c0 += GS_GRAIN_CHUNK) {
        const int c1 = min(K, c0 + GS_GRAIN_CHUNK);
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) {
            lcnt[i] = 0;
            lhi[i] = (tolsq - 0.75) +
                     gs_cert_window_fract(gs_load_ubi(ubis + 9 * (size_t)(c0 + i)), gmax, tolsq);
        }
        __syncthreads();
        for (int g = c0; g < c1; g++) {
            const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)g);
            const int L = glabels[g];
            const double hi = lhi[g - c0];
            int c = 0;
#pragma unroll
            for (int p = 0; p < P; p++) {
                const double sf = gs_drlv2_fract(u, x[p], y[p], z[p]);  // drlv2 - 3/4
                bool take = false;
                if (!(sf >= hi)) {  // candidate (or non-finite): reference arithmetic
                    const double s = gs_drlv2(u, x[p], y[p], z[p]);
                    take = (s < tolsq) && (s < d[p]);
                    if (take) {
                        lab[p] = L;
                        d[p] = s;
                    } else if (lab[p] == L) {
                        lab[p] = -1;
                    }
                } else if (lab[p] == L) {
                    lab[p] = -1;
                }
                c += __popcll(__ballot(take));
            }
            if (lane == 0 && c) atomicAdd(&lcnt[g - c0], c);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < c1 - c0; i += GS_BLOCK) {
            int c = lcnt[i];
            if (c) atomicAdd(&counts[c0 + i], c);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
        long long k = base + (long long)p * GS_BLOCK;
        if (k < ng) {
            drlv2[k] = d[p];
            labels[k] = lab[p];
        }
    }
}

// ---------------------------------------------------------------------------
// One-grain least-squares refinement over all ng peaks, two passes:
// per-block partial sums (n, sum drlv2, R, H) then a one-wave finalize that
// adds the block partials in block order (deterministic for a given grid),
// solves UB = R.H^-1, UBI = UB^-1 in place and writes n and the mean drlv2.
// MODE 0: peaks with drlv2 < tol^2 (score_and_refine, src/closest.c:267-345)
// MODE 1: peaks with labels == label (refine_assigned, src/closest.c:416-474)
// FUSED (a grid of one block): the block finishes the job itself, with the
// finalize's arithmetic, so a fused launch gives the bits of the two-pass
// one-block launch; ubi is then read by every thread first and written by
// thread 0 after the block's barrier.
__device__ __forceinline__ void gs_unpack_RH(const double* t, double* R, double* H);
__device__ __forceinline__ void gs_refine_finish(const double* tot, double* ubi, int* n_out,
                                                 double* mean_out);

template <int MODE, bool FUSED>
__global__ __launch_bounds__(GS_BLOCK) void k_refine_partials(
    double* ubi, const double* __restrict__ gv, int ng, double tolsq,
    const int* __restrict__ labels, int label, double* __restrict__ partials,
    int* __restrict__ n_out, double* __restrict__ mean_out) {
    const gs_ubi u = gs_load_ubi(ubi);
    double a[GS_NPART];
#pragma unroll
    for (int i = 0; i < GS_NPART; i++) a[i] = 0.0;
    for (long long k = (long long)blockIdx.x * GS_BLOCK + threadIdx.x; k < ng;
         k += (long long)gridDim.x * GS_BLOCK) {
        double g0 = gv[3 * k + 0], g1 = gv[3 * k + 1], g2 = gv[3 * k + 2];
        double i0, i1, i2;
        double s = gs_drlv2(u, g0, g1, g2, i0, i1, i2);
        bool keep = (MODE == 0) ? (s < tolsq) : (labels[k] == label);
        if (keep) {
            a[0] += 1.0;
            a[1] += s;
            a[2] += i0 * g0; a[3] += i1 * g0; a[4] += i2 * g0;   // R[0][j] += ih[j]*g[0]
            a[5] += i0 * g1; a[6] += i1 * g1; a[7] += i2 * g1;
            a[8] += i0 * g2; a[9] += i1 * g2; a[10] += i2 * g2;
            a[11] += i0 * i0; a[12] += i1 * i0; a[13] += i2 * i0;  // H[0][0..2]
            a[14] += i1 * i1; a[15] += i2 * i1; a[16] += i2 * i2;  // H[1][1..2], H[2][2]
        }
    }
    __shared__ double red[GS_BLOCK / GS_WAVE][GS_NPART];
    const int lane = threadIdx.x & (GS_WAVE - 1), wave = threadIdx.x / GS_WAVE;
#pragma unroll
    for (int i = 0; i < GS_NPART; i++) {
        double v = gs_wave_sum(a[i], GS_WAVE);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    __shared__ double tot[GS_NPART];
    if (threadIdx.x < GS_NPART) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < GS_BLOCK / GS_WAVE; w++) v += red[w][threadIdx.x];
        if (FUSED)
            tot[threadIdx.x] = 0.0 + v;  // the finalize's sum over one block
        else
            partials[(size_t)blockIdx.x * GS_NPART + threadIdx.x] = v;
    }
    if (FUSED) {
        __syncthreads();
        if (threadIdx.x == 0) gs_refine_finish(tot, ubi, n_out, mean_out);
    }
}

// R and H (full 3x3) from the 17 partial sums
__device__ __forceinline__ void gs_unpack_RH(const double* t, double* R, double* H) {
    for (int i = 0; i < 9; i++) R[i] = t[2 + i];
    H[0] = t[11]; H[1] = t[12]; H[2] = t[13];
    H[3] = t[12]; H[4] = t[14]; H[5] = t[15];
    H[6] = t[13]; H[7] = t[15]; H[8] = t[16];
}

// Solve from the 17 totals, UBI in place, n and the mean drlv2 (one thread).
__device__ __forceinline__ void gs_refine_finish(const double* tot, double* ubi, int* n_out,
                                                 double* mean_out) {
    double R[9], H[9], u[9];
    gs_unpack_RH(tot, R, H);
    for (int i = 0; i < 9; i++) u[i] = ubi[i];
    gs_solve_ubi(u, R, H);
    for (int i = 0; i < 9; i++) ubi[i] = u[i];
    int n = (int)tot[0];
    double s = tot[1];
    *n_out = n;
    *mean_out = n > 0 ? s / n : s;
}

__global__ void k_refine_finalize(const double* __restrict__ partials, int nblocks,
                                  double* __restrict__ ubi, int* __restrict__ n_out,
                                  double* __restrict__ mean_out) {
    __shared__ double tot[GS_NPART];
    if (threadIdx.x < GS_NPART) {
        double v = 0.0;
        for (int b = 0; b < nblocks; b++) v += partials[(size_t)b * GS_NPART + threadIdx.x];
        tot[threadIdx.x] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) gs_refine_finish(tot, ubi, n_out, mean_out);
}

// One score_and_refine term (src/closest.c:296-316): a peak within tol adds
// to n, sum drlv2, R = sum g h^T and H = sum h h^T (the six distinct
// entries).  The keep decision and drlv2 are the reference's arithmetic;
// the products go into the sums by FMA (the batched kernel's sums are
// split over lanes, so their rounding is not the reference's anyway): H's
// products of integers below 2^26 are exact, so its sums are the same,
// and R's terms are each rounded once instead of twice.
__device__ __forceinline__ void gs_refine_add(const gs_ubi& U, double g0, double g1, double g2,
                                              double tolsq, double* a) {
    double i0, i1, i2;
    double s = gs_drlv2(U, g0, g1, g2, i0, i1, i2);
    if (s < tolsq) {
        a[0] += 1.0;
        a[1] += s;
        a[2] = __builtin_fma(i0, g0, a[2]); a[3] = __builtin_fma(i1, g0, a[3]);
        a[4] = __builtin_fma(i2, g0, a[4]); a[5] = __builtin_fma(i0, g1, a[5]);
        a[6] = __builtin_fma(i1, g1, a[6]); a[7] = __builtin_fma(i2, g1, a[7]);
        a[8] = __builtin_fma(i0, g2, a[8]); a[9] = __builtin_fma(i1, g2, a[9]);
        a[10] = __builtin_fma(i2, g2, a[10]);
        a[11] = __builtin_fma(i0, i0, a[11]); a[12] = __builtin_fma(i1, i0, a[12]);
        a[13] = __builtin_fma(i2, i0, a[13]); a[14] = __builtin_fma(i1, i1, a[14]);
        a[15] = __builtin_fma(i2, i1, a[15]); a[16] = __builtin_fma(i2, i2, a[16]);
    }
}

// ---------------------------------------------------------------------------
// rows[q] = gv[idx[q]] for the n entries of CSR peak lists, a NaN row where
// idx[q] is outside [0, ng) (a NaN row scores NaN and is never kept, as an
// out-of-range index is never kept).
__global__ __launch_bounds__(GS_BLOCK) void k_gather_rows(const double* __restrict__ gv, int ng,
                                                          const int* __restrict__ idx, long long n,
                                                          double* __restrict__ rows) {
    const long long q = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (q >= n) return;
    const long long k = idx[q];
    double g0 = __builtin_nan(""), g1 = g0, g2 = g0;
    if (k >= 0 && k < ng) {
        g0 = gv[3 * k + 0];
        g1 = gv[3 * k + 1];
        g2 = gv[3 * k + 2];
    }
    rows[3 * q + 0] = g0;
    rows[3 * q + 1] = g1;
    rows[3 * q + 2] = g2;
}

// Rows q, q + L, ..., q + (B-1) L of a lane into buf, each clamped to the
// list's last row o1 - 1 (in bounds: o1 <= the number of rows); nothing is
// loaded when q >= o1.
template <int B, int L>
__device__ __forceinline__ void gs_rows_load(double (*buf)[3], const double* __restrict__ rows,
                                             long long q, long long o1) {
    if (!(q < o1)) return;
#pragma unroll
    for (int j = 0; j < B; j++) {
        long long r = q + (long long)j * L;
        r = r < o1 ? r : o1 - 1;
        buf[j][0] = rows[3 * r + 0];
        buf[j][1] = rows[3 * r + 1];
        buf[j][2] = rows[3 * r + 2];
    }
}

// The terms of the rows in buf that lie below o1, in row order.
template <int B, int L>
__device__ __forceinline__ void gs_rows_add(const gs_ubi& U, const double (*buf)[3], long long q,
                                            long long o1, double tolsq, double* a) {
#pragma unroll
    for (int j = 0; j < B; j++)
        if (q + (long long)j * L < o1) gs_refine_add(U, buf[j][0], buf[j][1], buf[j][2], tolsq, a);
}

// ---------------------------------------------------------------------------
// One work item of k_refine_lists: the grains in slots base .. base +
// 64/L - 1 of `order` below `limit` (<= K), L lanes each, every grain's
// n_iter iterations.
template <int L, bool ROWS>
__device__ __forceinline__ void gs_refine_item(
    int base, int limit, int lane, double* __restrict__ ubis, int K,
    const double* __restrict__ gv, int ng, const int* __restrict__ idx, long long n_idx,
    const long long* __restrict__ off, const int* __restrict__ order, double tolsq, int n_iter,
    int* __restrict__ n_out, double* __restrict__ mean_out) {
    const int sub = lane / L;
    const int sl = lane - sub * L;
    const int slot = base + sub;
    const int gq = slot < limit ? order[slot] : -1;
    const bool active = gq >= 0 && gq < K;
    const int g = active ? gq : 0;
    long long o1 = active ? off[g + 1] : 0;
    o1 = o1 < 0 ? 0 : (o1 > n_idx ? n_idx : o1);
    long long o0 = active ? off[g] : 0;
    o0 = o0 < 0 ? 0 : (o0 > o1 ? o1 : o0);
    double u[9];
    for (int i = 0; i < 9; i++) u[i] = active ? ubis[9 * (size_t)g + i] : 0.0;
    double X[GS_REFINE_BATCH][3], Y[GS_REFINE_BATCH][3];  // rows mode only
    if (ROWS) gs_rows_load<GS_REFINE_BATCH, L>(X, gv, o0 + sl, o1);
    for (int it = 0; it < n_iter; it++) {
        const gs_ubi U = {u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7], u[8]};
        double a[GS_NPART];
#pragma unroll
        for (int i = 0; i < GS_NPART; i++) a[i] = 0.0;
        if (ROWS) {
            // two buffers of GS_REFINE_BATCH rows: the next batch is in
            // flight while this one is added; the lane adds its rows in
            // list order
            long long q = o0 + sl;
            while (q < o1) {
                const long long qn = q + (long long)GS_REFINE_BATCH * L;
                gs_rows_load<GS_REFINE_BATCH, L>(Y, gv, qn, o1);
                gs_rows_add<GS_REFINE_BATCH, L>(U, X, q, o1, tolsq, a);
                if (!(qn < o1)) break;
                q = qn + (long long)GS_REFINE_BATCH * L;
                gs_rows_load<GS_REFINE_BATCH, L>(X, gv, q, o1);
                gs_rows_add<GS_REFINE_BATCH, L>(U, Y, qn, o1, tolsq, a);
            }
            // the next iteration's first rows, in flight during the sums
            // and the solve
            if (it + 1 < n_iter) gs_rows_load<GS_REFINE_BATCH, L>(X, gv, o0 + sl, o1);
        } else {
            for (long long q = o0 + sl; q < o1; q += L) {
                const long long k = idx[q];
                if (k < 0 || k >= ng) continue;
                gs_refine_add(U, gv[3 * k + 0], gv[3 * k + 1], gv[3 * k + 2], tolsq, a);
            }
        }
#pragma unroll
        for (int i = 0; i < GS_NPART; i++) a[i] = gs_group_sum<L>(a[i]);
        double R[9], H[9];
        gs_unpack_RH(a, R, H);
        // groups of whole DPP rows share each inverse's divisions out
        // over the row (the same results as gs_solve_ubi)
        if constexpr (L >= 16)
            gs_solve_ubi_row(u, R, H, lane & 15);
        else
            gs_solve_ubi(u, R, H);
        if (active && sl == 0) {
            const int n = (int)a[0];
            n_out[(size_t)it * K + g] = n;
            mean_out[(size_t)it * K + g] = n > 0 ? a[1] / n : a[1];
        }
    }
    if (active && sl == 0)
        for (int i = 0; i < 9; i++) ubis[9 * (size_t)g + i] = u[i];
}

// Batched score_and_refine over CSR peak lists: a persistent grid whose
// waves pull work items from a queue.  The slots of `order` fall into
// three tiers: [0, n64) one grain per item on 64 lanes, [n64, n32) two
// grains per item on 32 lanes, [n32, K) 64/LPG grains per item on LPG
// lanes, with n64 = tiers[0] and n32 = tiers[1] (clipped to 0 <= n64 <=
// n32 <= K; tiers NULL: every grain on LPG lanes).  With `order` longest
// list first and the tier counts the numbers of lists of at least given
// lengths (k_count_tiers), each list gets lanes in proportion to its
// length, so no list sets the makespan.  Each grain runs n_iter
// successive score_and_refine calls on its own list (the calls of
// different grains are independent).  Sums are reduced across the grain's
// lanes by gs_group_sum, which leaves bit-identical totals in every lane,
// so each lane solves the 3x3 system itself and keeps the updated UBI in
// registers.  n_out/mean_out are [n_iter][K]; queue[0] must be zero at
// launch.
// Out-of-range entries are ignored rather than trusted: a grain number
// outside [0, K) in `order` is skipped, list bounds are clipped to
// [0, n_idx], and peak indices outside [0, ng) are not kept.
// ROWS: there is no index array; list entry q is row q of gv (the rows
// gathered once by k_gather_rows), so every iteration streams contiguous
// rows instead of gathering one cache line per entry.
template <int LPG, bool ROWS>
__global__ __launch_bounds__(GS_BLOCK) void k_refine_lists(
    double* __restrict__ ubis, int K, const double* __restrict__ gv, int ng,
    const int* __restrict__ idx, long long n_idx, const long long* __restrict__ off,
    const int* __restrict__ order, double tolsq, int n_iter, int* __restrict__ n_out,
    double* __restrict__ mean_out, const int* __restrict__ tiers, int* __restrict__ queue) {
    constexpr int GPW = GS_WAVE / LPG;
    const int lane = threadIdx.x & (GS_WAVE - 1);
    int n64 = tiers ? tiers[0] : 0;
    n64 = n64 < 0 ? 0 : (n64 > K ? K : n64);
    int n32 = tiers ? tiers[1] : n64;
    n32 = n32 < n64 ? n64 : (n32 > K ? K : n32);
    const int items64 = n64;
    const int items32 = (n32 - n64 + 1) / 2;
    const int nitems = items64 + items32 + (K - n32 + GPW - 1) / GPW;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(queue, 1);
        item = __shfl(item, 0, GS_WAVE);
        if (item >= nitems) break;
        if (item < items64)
            gs_refine_item<GS_WAVE, ROWS>(item, n64, lane, ubis, K, gv, ng, idx, n_idx, off,
                                          order, tolsq, n_iter, n_out, mean_out);
        else if (item < items64 + items32)
            gs_refine_item<32, ROWS>(n64 + 2 * (item - items64), n32, lane, ubis, K, gv, ng,
                                     idx, n_idx, off, order, tolsq, n_iter, n_out, mean_out);
        else
            gs_refine_item<LPG, ROWS>(n32 + (item - items64 - items32) * GPW, K, lane, ubis, K,
                                      gv, ng, idx, n_idx, off, order, tolsq, n_iter, n_out,
                                      mean_out);
    }
}

// tiers[0] = the number of grains whose clipped list length is at least
// long_min (0 when long_min <= 0), tiers[1] = the number at least mid_min
// (tiers[0] when mid_min <= 0); queue[0] = 0 (the refine launch's work
// counter, cleared here so that no separate fill is needed; NULL: none).
// One block of GS_TIER_THREADS; each thread loads its grains' offsets
// GS_TIER_BATCH at a time, so K up to GS_TIER_THREADS x GS_TIER_BATCH costs
// one memory latency; each wave counts with ballots and adds its totals
// to LDS once.  The counts do not depend on the grain order.
#define GS_TIER_THREADS 1024
#define GS_TIER_BATCH 8
__global__ __launch_bounds__(GS_TIER_THREADS) void k_count_tiers(
    const long long* __restrict__ off, int K, long long n_idx, long long long_min,
    long long mid_min, int* __restrict__ tiers, int* __restrict__ queue) {
    __shared__ int cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0 && queue) queue[0] = 0;
    __syncthreads();
    int c64 = 0, c32 = 0;
    for (int g0 = threadIdx.x; g0 < K; g0 += GS_TIER_THREADS * GS_TIER_BATCH) {
        long long a[GS_TIER_BATCH], b[GS_TIER_BATCH];
#pragma unroll
        for (int j = 0; j < GS_TIER_BATCH; j++) {
            const int g = g0 + j * GS_TIER_THREADS;
            a[j] = g < K ? off[g] : 0;
            b[j] = g < K ? off[g + 1] : 0;
        }
#pragma unroll
        for (int j = 0; j < GS_TIER_BATCH; j++) {
            long long o1 = b[j] < 0 ? 0 : (b[j] > n_idx ? n_idx : b[j]);
            long long o0 = a[j] < 0 ? 0 : (a[j] > o1 ? o1 : a[j]);
            const long long len = o1 - o0;  // 0 past K: counted in no tier
            const bool is64 = long_min > 0 && len >= long_min;
            const bool is32 = mid_min > 0 ? (len >= mid_min || is64) : is64;
            c64 += __popcll(__ballot(is64));
            c32 += __popcll(__ballot(is32));
        }
    }
    if ((threadIdx.x & (GS_WAVE - 1)) == 0) {  // c64, c32: the wave's totals
        atomicAdd(&cnt[0], c64);
        atomicAdd(&cnt[1], c32);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        tiers[0] = cnt[0];
        tiers[1] = cnt[1];
    }
}

// ---------------------------------------------------------------------------
// Geometry kernels.
struct gs_detector {
    double p[4];     // z_center, y_center, z_size, y_size
    double r[9];     // detector rotation . flip, row-major
    double dist[3];  // distance vector
};

struct gs_goniometer {
    double omegasign, wvln, wedge, chi;
    double t[3];    // grain translation
    double mat[9];  // wedge/chi matrix, used only when mat_given
    int mat_given;
};

// per-block wedge/chi matrix and 1/wavelength in LDS
__device__ __forceinline__ void gs_block_setup(const gs_goniometer& gp, double* smat, double* sds) {
    if (threadIdx.x == 0) {
        if (gp.mat_given) {
            for (int i = 0; i < 9; i++) smat[i] = gp.mat[i];
        } else {
            gs_wedge_chi_matrix(gp.wedge, gp.chi, smat);
        }
        *sds = 1. / gp.wvln;
    }
    __syncthreads();
}

// compute_xlylzl (src/cdiffraction.c:190-224)
__global__ __launch_bounds__(GS_BLOCK) void k_xlylzl(const double* __restrict__ s,
                                                     const double* __restrict__ f, gs_detector det,
                                                     double* __restrict__ out, int n) {
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= n) return;
    double x0, x1, x2;
    gs_xlylzl_one(s[i], f[i], det.p, det.r, det.dist, x0, x1, x2);
    out[3 * i + 0] = x0;
    out[3 * i + 1] = x1;
    out[3 * i + 2] = x2;
}

// sin/cos of ((PI/180)*omega)*omegasign per peak
__global__ __launch_bounds__(GS_BLOCK) void k_sincos_omega(const double* __restrict__ omega,
                                                           double omegasign, double* __restrict__ so,
                                                           double* __restrict__ co, int n) {
    const long long i = (long long)blockIdx.x * GS_BLOCK + threadIdx.x;
    if (i >= n) return;
    double a, b;
    gs_sincos(gs_omega_rad(omega[i], omegasign), &a, &b);
    so[i] = a;
    co[i] = b;
}

// A block's GS_BLOCK rows of three doubles, row t held by thread t (rows
// past n, i.e. t >= n - i0, are not stored), written to out[3*i0 ..] through
// LDS: consecutive threads store consecutive doubles, so each wave writes
// whole contiguous 128-byte lines instead of three strided 8-byte stores per
// lane.  Every thread of the block must call it (it synchronises).
__device__ __forceinline__ void gs_store_rows3(double* __restrict__ out, long long i0, long long n,
                                               double a, double b, double c,
                                               double* __restrict__ lds) {
    const int t = threadIdx.x;
    __syncthreads();  // lds may still be read by a previous call
    lds[3 * t + 0] = a;
    lds[3 * t + 1] = b;
    lds[3 * t + 2] = c;
    __syncthreads();
    const long long rows = n - i0 < GS_BLOCK ? n - i0 : GS_BLOCK;
    for (int j = t; j < 3 * rows; j += GS_BLOCK) out[3 * i0 + j] = lds[j];
}

// compute_gv (src/cdiffraction.c:112-168).  TRIG: sin/cos per peak are
// read from so/co instead of being evaluated from omega.
template <bool TRIG>
__global__ __launch_bounds__(GS_BLOCK) void k_gv(const double* __restrict__ xl,
                                                 const double* __restrict__ omega,
                                                 const double* __restrict__ so_in,
                                                 const double* __restrict__ co_in,
                                                 gs_goniometer gp, double* __restrict__ gv, int n) {
    __shared__ double smat[9];
    __shared__ double sds;
    __shared__ double rows[3 * GS_BLOCK];
    gs_block_setup(gp, smat, &sds);
    const long long i0 = (long long)blockIdx.x * GS_BLOCK;
    const long long i = i0 + threadIdx.x;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (i < n) {
        double so, co;
        if (TRIG) {
            so = so_in[i];
            co = co_in[i];
        } else {
            gs_sincos(gs_omega_rad(omega[i], gp.omegasign), &so, &co);
        }
        double mat[9];
        for (int j = 0; j < 9; j++) mat[j] = smat[j];
        gs_gv_one(xl[3 * i + 0], xl[3 * i + 1], xl[3 * i + 2], so, co, mat, sds, gp.t[0],
                  gp.t[1], gp.t[2], g0, g1, g2);
    }
    gs_store_rows3(gv, i0, n, g0, g1, g2, rows);
}

// compute_xlylzl -> compute_gv fused: raw (s, f, omega) to g-vectors in one
// pass; the lab positions and per-peak sin/cos are optionally stored too.
// TRIG: sin/cos per peak are read from so_in/co_in (omega is not read).
template <bool TRIG>
__global__ __launch_bounds__(GS_BLOCK) void k_sf2gv(
    const double* __restrict__ s, const double* __restrict__ f, const double* __restrict__ omega,
    const double* __restrict__ so_in, const double* __restrict__ co_in, gs_detector det,
    gs_goniometer gp, double* __restrict__ xl_out, double* __restrict__ so_out,
    double* __restrict__ co_out, double* __restrict__ gv, int n) {
    __shared__ double smat[9];
    __shared__ double sds;
    __shared__ double rows[3 * GS_BLOCK];
    gs_block_setup(gp, smat, &sds);
    const long long i0 = (long long)blockIdx.x * GS_BLOCK;
    const long long i = i0 + threadIdx.x;
    double x0 = 0.0, x1 = 0.0, x2 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (i < n) {
        gs_xlylzl_one(s[i], f[i], det.p, det.r, det.dist, x0, x1, x2);
        double so, co;
        if (TRIG) {
            so = so_in[i];
            co = co_in[i];
        } else {
            gs_sincos(gs_omega_rad(omega[i], gp.omegasign), &so, &co);
        }
        double mat[9];
        for (int j = 0; j < 9; j++) mat[j] = smat[j];
        gs_gv_one(x0, x1, x2, so, co, mat, sds, gp.t[0], gp.t[1], gp.t[2], g0, g1, g2);
        if (so_out) {
            so_out[i] = so;
            co_out[i] = co;
        }
    }
    if (xl_out) gs_store_rows3(xl_out, i0, n, x0, x1, x2, rows);  // uniform branch
    gs_store_rows3(gv, i0, n, g0, g1, g2, rows);
}
