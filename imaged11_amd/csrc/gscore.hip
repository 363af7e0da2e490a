// libgscore: MI355X (gfx950) grain-scoring kernels behind a C ABI.
//
// Two families of entry points (declared in include/gscore.h):
//   * the reference's own C prototypes (score, score_and_refine, ...) on
//     host pointers: inputs are staged to device memory, the HIP kernels
//     run, outputs are copied back, the call returns when done;
//   * gs_* entry points on device pointers, enqueued on a caller's stream
//     (a hipStream_t passed as void*), with an int status.
// Errors: gs_* return a non-zero status; the reference prototypes, which
// have no status in the reference, record the error (gs_last_error) and
// return -1 from the int-returning ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sched.h>

#include "gs_kernels.h"
#include "gs_adjoint.h"
#include "gs_assign_gv.h"
#include "gs_index.h"
#include "gs_geometry.h"
#include "gs_grains.h"
#include "gs_pbp.h"
#include "../../include/gscore.h"

// Scratch for the entry point that allocates nothing: the block partials
// of the one-grain refinement (1024 blocks x 17 doubles = 136 KiB), one copy
// per GPU, held until the process ends.  Its use is explained beside
// scratch_ptr() below.
#define GS_SCRATCH_DOUBLES (1024 * GS_NPART)
static __device__ __align__(256) double scratch[GS_SCRATCH_DOUBLES];

namespace {

thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define GS_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err(GS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

#define GS_LAUNCHED()                                                                      \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return set_err(GS_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)

#define GS_CHECK(code)              \
    do {                            \
        int c_ = (code);            \
        if (c_ != GS_OK) return c_; \
    } while (0)

// Per-device state: a few int counters (default work queues) and the
// adjoint's default reduction workspace, allocated once per device.
struct DeviceCtx {
    bool ready = false;
    int cus = 0;
    int* counters = nullptr;  // GS_NCOUNTERS ints
    double* adj_ws = nullptr;  // the adjoint's workspace when the caller passes none
    long long adj_cap = 0;     // its size in doubles
};

constexpr int GS_MAX_PARTIAL_BLOCKS = 1024;
constexpr int GS_NCOUNTERS = 4096;
std::mutex g_ctx_mutex;
std::vector<DeviceCtx> g_ctx;

int get_ctx(DeviceCtx** out) {
    int dev = 0, n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_err(GS_ERR_NODEVICE, "no HIP device available (%s)",
                       e != hipSuccess ? hipGetErrorString(e) : "count is 0");
    GS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if ((int)g_ctx.size() < n) g_ctx.resize(n);
    DeviceCtx& c = g_ctx[dev];
    if (!c.ready) {
        GS_HIP(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, dev));
        GS_HIP(hipMalloc(&c.counters, sizeof(int) * GS_NCOUNTERS));
        c.ready = true;
    }
    *out = &c;
    return GS_OK;
}

// The device's adjoint workspace, grown to at least `doubles` (calls that
// overlap on one device must bring their own).
int adj_workspace(DeviceCtx* ctx, long long doubles, double** out) {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (ctx->adj_cap < doubles) {
        // an earlier launch may still be reading the old buffer
        if (ctx->adj_ws) {
            GS_HIP(hipDeviceSynchronize());
            GS_HIP(hipFree(ctx->adj_ws));
        }
        ctx->adj_ws = nullptr;
        ctx->adj_cap = 0;
        GS_HIP(hipMalloc(&ctx->adj_ws, sizeof(double) * (size_t)doubles));
        ctx->adj_cap = doubles;
    }
    *out = ctx->adj_ws;
    return GS_OK;
}

hipStream_t S(gs_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

int blocks_for(long long n, int per_block) {
    long long b = (n + per_block - 1) / per_block;
    return (int)std::max<long long>(1, b);
}

int fill_i32(int* p, int v, long long n, hipStream_t st) {
    if (n <= 0) return GS_OK;
    int nb = (int)std::min<long long>(blocks_for(n, GS_BLOCK), 4096);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(GS_BLOCK), 0, st, p, v, n);
    GS_LAUNCHED();
    return GS_OK;
}

int fill_f64(double* p, double v, long long n, hipStream_t st) {
    if (n <= 0) return GS_OK;
    int nb = (int)std::min<long long>(blocks_for(n, GS_BLOCK), 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(GS_BLOCK), 0, st, p, v, n);
    GS_LAUNCHED();
    return GS_OK;
}

// Block partials of the one-grain refinement: one buffer per host thread
// and device, so that calls from different host threads never share it
// (score_and_refine and refine_assigned are thread-safe in the reference).
struct PartialsTL {
    int dev = -1;
    double* buf = nullptr;
};
thread_local PartialsTL g_partials;

int get_partials(double** out) {
    int dev = 0;
    GS_HIP(hipGetDevice(&dev));
    if (g_partials.dev != dev || !g_partials.buf) {
        g_partials = PartialsTL();
        GS_HIP(hipMalloc(&g_partials.buf, sizeof(double) * GS_MAX_PARTIAL_BLOCKS * GS_NPART));
        g_partials.dev = dev;
    }
    *out = g_partials.buf;
    return GS_OK;
}

// Host-pointer staging: one growable device buffer per host thread.
struct Staging {
    int dev = -1;
    char* buf = nullptr;
    size_t cap = 0;
};
thread_local Staging g_stage;

int stage_reserve(size_t bytes, char** out) {
    int dev = 0;
    GS_HIP(hipGetDevice(&dev));
    if (g_stage.dev != dev) {  // buffers belong to one device
        g_stage = Staging();
        g_stage.dev = dev;
    }
    if (bytes > g_stage.cap) {
        if (g_stage.buf) GS_HIP(hipFree(g_stage.buf));
        g_stage.buf = nullptr;
        size_t cap = std::max<size_t>(bytes, 1 << 20);
        GS_HIP(hipMalloc(&g_stage.buf, cap));
        g_stage.cap = cap;
    }
    *out = g_stage.buf;
    return GS_OK;
}

// Reserve one staging buffer holding consecutive 256-byte aligned slices
// of the given sizes; the reservation is computed from the same list that
// is carved, so no slice can run past it.
size_t slice_bytes(size_t b) { return (b + 255) & ~size_t(255); }

int stage_slices(std::initializer_list<size_t> sizes, std::vector<char*>& out) {
    size_t need = 0;
    for (size_t b : sizes) need += slice_bytes(b);
    char* buf;
    GS_CHECK(stage_reserve(need, &buf));
    out.clear();
    size_t at = 0;
    for (size_t b : sizes) {
        out.push_back(buf + at);
        at += slice_bytes(b);
    }
    return GS_OK;
}

// Small host-pointer calls go through pinned, device-mapped host memory,
// one growable buffer per host thread: the kernels read the caller's data
// and write their results there directly, so a call costs no copy
// launches.  Used up to GS_PINNED_MAX bytes; larger inputs are copied to the
// device staging buffer instead.
constexpr size_t GS_PINNED_MAX = size_t(1) << 20;
struct Pinned {
    char* host = nullptr;
    char* dev = nullptr;  // the same memory as the device addresses it
    size_t cap = 0;
};
thread_local Pinned g_pinned;

int pinned_slices(std::initializer_list<size_t> sizes, std::vector<char*>& host,
                  std::vector<char*>& dev) {
    size_t need = 0;
    for (size_t b : sizes) need += slice_bytes(b);
    if (need > g_pinned.cap) {
        if (g_pinned.host) GS_HIP(hipHostFree(g_pinned.host));
        g_pinned = Pinned();
        void* h = nullptr;
        GS_HIP(hipHostMalloc(&h, need, hipHostMallocMapped | hipHostMallocPortable));
        void* d = nullptr;
        GS_HIP(hipHostGetDevicePointer(&d, h, 0));
        g_pinned.host = static_cast<char*>(h);
        g_pinned.dev = static_cast<char*>(d);
        g_pinned.cap = need;
    }
    host.clear();
    dev.clear();
    size_t at = 0;
    for (size_t b : sizes) {
        host.push_back(g_pinned.host + at);
        dev.push_back(g_pinned.dev + at);
        at += slice_bytes(b);
    }
    return GS_OK;
}

int check_ng(long long ng) {
    if (ng < 0) return set_err(GS_ERR_ARG, "negative array length %lld", ng);
    return GS_OK;
}

// ---- launchers shared by both families ------------------------------------

// Tiles of the all-pairs kernels: chunks of GS_BLOCK*P peaks times splits of
// at most GS_SPLIT_MAX grains (the LDS tile).  About 128 tiles per CU:
// with 4 resident blocks per CU that is ~32 rounds, so the partly filled
// last round costs ~3% (at 32 per CU it cost ~10%), while a split keeps
// at least 8 grains to amortise the tile's peak loads and window setup.
int pair_tiles(DeviceCtx* ctx, int K, int ng, int P, int* gsplit_out, int* nsplit_out,
               unsigned* nb_out, int split_max = GS_SPLIT_MAX, int tiles_per_cu = 128) {
    const long long nchunks = (ng + (long long)GS_BLOCK * P - 1) / ((long long)GS_BLOCK * P);
    const long long want = (long long)tiles_per_cu * ctx->cus;
    long long nsplit = std::max<long long>(1, (want + nchunks - 1) / nchunks);
    nsplit = std::min<long long>(nsplit, std::max(1, K / 8));
    nsplit = std::max<long long>(nsplit, (K + split_max - 1) / split_max);
    const int gsplit = (int)((K + nsplit - 1) / nsplit);
    nsplit = (K + gsplit - 1) / gsplit;
    const long long nb = nchunks * nsplit;
    if (nb > 0x7fffffffLL) return set_err(GS_ERR_ARG, "grid too large");
    *gsplit_out = gsplit;
    *nsplit_out = (int)nsplit;
    *nb_out = (unsigned)nb;
    return GS_OK;
}

int launch_score_many(const double* ubis, int K, const double* gv, int ng, double tol, int* counts,
                      int peaks_per_lane, int mode, hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0 || ng < 0) return set_err(GS_ERR_ARG, "negative size K=%d ng=%d", K, ng);
    GS_CHECK(fill_i32(counts, 0, K, st));
    if (K == 0 || ng == 0) return GS_OK;
    const double tolsq = tol * tol;
    // default tile height: 4 peaks per lane for the certified kernel (fewer
    // VGPRs, more resident waves), 8 for the exact one (measured best)
    int P = peaks_per_lane;
    if (P != 1 && P != 2 && P != 4 && P != 8 && P != 16) P = mode == GS_SCORE_EXACT ? 8 : 4;
    if (mode == GS_SCORE_SIEVE && P > 8) P = 8;
    int gsplit, nsplit;
    unsigned nb;
    if (mode == GS_SCORE_SIEVE)
        GS_CHECK(pair_tiles(ctx, K, ng, P, &gsplit, &nsplit, &nb, GS_SIEVE_SPLIT_MAX,
                            GS_SIEVE_TILES_PER_CU));
    else
        GS_CHECK(pair_tiles(ctx, K, ng, P, &gsplit, &nsplit, &nb));
    const dim3 grid(nb), block(GS_BLOCK);
    if (mode == GS_SCORE_SIEVE) {
        if (gsplit > GS_SIEVE_SPLIT_MAX) return set_err(GS_ERR_ARG, "sieve split %d too large", gsplit);
        switch (P) {
            case 1: hipLaunchKernelGGL(k_score_many_sieve<1>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 2: hipLaunchKernelGGL(k_score_many_sieve<2>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 8: hipLaunchKernelGGL(k_score_many_sieve<8>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            default: hipLaunchKernelGGL(k_score_many_sieve<4>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
        }
    } else if (mode == GS_SCORE_EXACT) {
        switch (P) {
            case 1: hipLaunchKernelGGL(k_score_many<1>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 2: hipLaunchKernelGGL(k_score_many<2>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 4: hipLaunchKernelGGL(k_score_many<4>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 16: hipLaunchKernelGGL(k_score_many<16>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            default: hipLaunchKernelGGL(k_score_many<8>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
        }
    } else {  // GS_SCORE_CERTIFIED
        switch (P) {
            case 1: hipLaunchKernelGGL(k_score_many_fast<1>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 2: hipLaunchKernelGGL(k_score_many_fast<2>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 4: hipLaunchKernelGGL(k_score_many_fast<4>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            case 16: hipLaunchKernelGGL(k_score_many_fast<16>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
            default: hipLaunchKernelGGL(k_score_many_fast<8>, grid, block, 0, st, ubis, K, gsplit, nsplit, gv, ng, tolsq, counts); break;
        }
    }
    GS_LAUNCHED();
    return GS_OK;
}

int launch_assign_many(const double* ubis, const int* glabels, int K, const double* gv, int ng,
                       double tol, double* drlv2, int* labels, int* counts, int mode,
                       hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0 || ng < 0) return set_err(GS_ERR_ARG, "negative size K=%d ng=%d", K, ng);
    GS_CHECK(fill_i32(counts, 0, K, st));
    if (K == 0 || ng == 0) return GS_OK;
    const double tolsq = tol * tol;
    constexpr int P = 2;
    const long long nb = (ng + (long long)GS_BLOCK * P - 1) / ((long long)GS_BLOCK * P);
    if (mode == GS_SCORE_EXACT)
        hipLaunchKernelGGL(k_assign_many<P>, dim3((unsigned)nb), dim3(GS_BLOCK), 0, st, ubis,
                           glabels, K, gv, ng, tolsq, drlv2, labels, counts);
    else
        hipLaunchKernelGGL(k_assign_many_fast<P>, dim3((unsigned)nb), dim3(GS_BLOCK), 0, st, ubis,
                           glabels, K, gv, ng, tolsq, drlv2, labels, counts);
    GS_LAUNCHED();
    return GS_OK;
}

// Blocks of the one-grain refinement: one per 4·GS_BLOCK peaks, at most
// GS_MAX_PARTIAL_BLOCKS (each block strides over the peaks), so the block
// partials never need more than GS_MAX_PARTIAL_BLOCKS · GS_NPART doubles.
int refine_one_blocks(int ng) { return std::min(GS_MAX_PARTIAL_BLOCKS, blocks_for(ng, GS_BLOCK * 4)); }

// The launches of the one-grain refinement with a given block-partials
// buffer (GS_MAX_PARTIAL_BLOCKS · GS_NPART doubles; unused by one block).
int refine_one_launch(int mode, double* ubi, const double* gv, int ng, double tol,
                      const int* labels, int label, int* n_out, double* mean_out,
                      double* partials, hipStream_t st) {
    const int nb = refine_one_blocks(ng);
    const double tolsq = tol * tol;
    if (nb == 1) {  // one block: it finishes the job itself (same bits, one launch)
        if (mode == 0)
            hipLaunchKernelGGL((k_refine_partials<0, true>), dim3(1), dim3(GS_BLOCK), 0, st, ubi,
                               gv, ng, tolsq, labels, label, nullptr, n_out, mean_out);
        else
            hipLaunchKernelGGL((k_refine_partials<1, true>), dim3(1), dim3(GS_BLOCK), 0, st, ubi,
                               gv, ng, tolsq, labels, label, nullptr, n_out, mean_out);
        GS_LAUNCHED();
        return GS_OK;
    }
    if (mode == 0)
        hipLaunchKernelGGL((k_refine_partials<0, false>), dim3(nb), dim3(GS_BLOCK), 0, st, ubi, gv,
                           ng, tolsq, labels, label, partials, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_refine_partials<1, false>), dim3(nb), dim3(GS_BLOCK), 0, st, ubi, gv,
                           ng, tolsq, labels, label, partials, nullptr, nullptr);
    GS_LAUNCHED();
    hipLaunchKernelGGL(k_refine_finalize, dim3(1), dim3(GS_WAVE), 0, st, partials, nb, ubi,
                       n_out, mean_out);
    GS_LAUNCHED();
    return GS_OK;
}

int launch_refine_one(int mode, double* ubi, const double* gv, int ng, double tol,
                      const int* labels, int label, int* n_out, double* mean_out, hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(ng));
    double* partials = nullptr;
    if (refine_one_blocks(ng) > 1) GS_CHECK(get_partials(&partials));
    return refine_one_launch(mode, ubi, gv, ng, tol, labels, label, n_out, mean_out, partials, st);
}

// ---- scratch without allocation ----------------------------------------------
//
// A fixed device array, one copy per GPU, that holds the block partials of
// the one-grain refinement for gs_score_and_refine_dev_noalloc, whose
// callers cannot allocate (code under graph capture, serving loops).  It
// holds the partials of the largest grid that refinement launches, so no
// peak count is too large for it.  Host code never passes the array's name
// to a kernel: its device address, looked up once per device with
// hipGetSymbolAddress, is kept in g_scratch_ptr.
static_assert(GS_SCRATCH_DOUBLES == GS_MAX_PARTIAL_BLOCKS * GS_NPART,
              "scratch holds the partials of the largest one-grain refinement grid");

constexpr int GS_SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[GS_SCRATCH_MAX_DEVICES] = {};

// The current device's copy of `scratch` (looked up on the first call on
// that device).  Allocates nothing.
int scratch_ptr(double** out) {
    int dev = 0;
    GS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= GS_SCRATCH_MAX_DEVICES)
        return set_err(GS_ERR_ARG, "device %d: the scratch table holds %d devices", dev,
                       GS_SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(GS_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return GS_OK;
}

// ---- host libm trig: the parity mode ----------------------------------------
//
// The reference evaluates sin(RAD * omega[i] * omegasign) and cos(...) per
// peak (src/cdiffraction.c:146-147, RAD = PI / 180.0 unparenthesised), and
// sin/cos(wedge * RAD), sin/cos(chi * RAD) once (:119-120, :130-131).  gcc -O2
// compiles each such pair into ONE glibc sincos() call of the same argument.
// glibc's sincos is not correctly rounded everywhere and differs from any
// device algorithm in the last bit for a few tenths of a percent of
// arguments, so the parity mode calls that same sincos here, on the host,
// with the same expression (this translation unit is compiled with
// -ffp-contract=off; PI/180.0 folds to the same double as in gcc), and the
// kernels take the results as input.

// Host threads for n independent peaks: the CPUs this process may run on,
// or OMP_NUM_THREADS when set (as the reference's OpenMP loop would use),
// at most one per 16384 peaks.
int host_threads(long long n) {
    long long want = 0;
    if (const char* env = std::getenv("OMP_NUM_THREADS")) want = std::atoll(env);
    if (want <= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        want = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : 1;
    }
    want = std::min<long long>(want, 256);
    want = std::min<long long>(want, n / 16384 + 1);
    return (int)std::max<long long>(1, want);
}

void sincos_omega_rows(const double* omega, double omegasign, double* so, double* co,
                       long long a, long long b) {
    for (long long i = a; i < b; i++) sincos(GS_PI / 180.0 * omega[i] * omegasign, so + i, co + i);
}

// so[i], co[i] = glibc sincos(((PI/180.0)*omega[i])*omegasign), the peaks
// split into contiguous ranges over host threads (each value is computed
// the same way whichever thread computes it).
void sincos_omega_libm(const double* omega, double omegasign, double* so, double* co,
                       long long n) {
    const int nt = host_threads(n);
    if (nt <= 1) {
        sincos_omega_rows(omega, omegasign, so, co, 0, n);
        return;
    }
    std::vector<std::thread> pool;
    long long next = n / nt;  // first row no thread has taken; this thread takes [0, n/nt)
    try {
        for (int t = 1; t < nt; t++) {
            const long long a = n * t / nt, b = n * (t + 1) / nt;
            pool.emplace_back(sincos_omega_rows, omega, omegasign, so, co, a, b);
            next = b;
        }
    } catch (...) {  // no more threads could start: this thread takes the rest
    }
    sincos_omega_rows(omega, omegasign, so, co, 0, n / nt);
    if (next < n) sincos_omega_rows(omega, omegasign, so, co, next, n);
    for (auto& th : pool) th.join();
}

// The wedge/chi matrix of compute_gv from glibc's sincos (src/cdiffraction.c:119-142).
void wedge_chi_matrix_libm(double wedge, double chi, double* mat) {
    double sw, cw, sc, cc;
    sincos(gs_wedge_rad(wedge), &sw, &cw);
    sincos(gs_wedge_rad(chi), &sc, &cc);
    gs_wedge_chi_from_trig(sw, cw, sc, cc, mat);
}

gs_detector make_detector(const double* p, const double* r, const double* dist) {
    gs_detector d;
    for (int i = 0; i < 4; i++) d.p[i] = p[i];
    for (int i = 0; i < 9; i++) d.r[i] = r[i];
    for (int i = 0; i < 3; i++) d.dist[i] = dist[i];
    return d;
}

// The goniometer of a launch.  Without a given matrix the wedge/chi
// matrix is formed here, on the host, by the same correctly rounded
// sincos the device would run (gs_sincos_v compiles for both, and a
// correctly rounded result is unique), so every block reads it from its
// arguments instead of one thread per block computing two sincos while
// the others wait.
gs_goniometer make_gonio(double omegasign, double wvln, double wedge, double chi, const double* t,
                         const double* mat) {
    gs_goniometer g;
    g.omegasign = omegasign;
    g.wvln = wvln;
    g.wedge = wedge;
    g.chi = chi;
    for (int i = 0; i < 3; i++) g.t[i] = t ? t[i] : 0.0;
    g.mat_given = 1;
    if (mat)
        for (int i = 0; i < 9; i++) g.mat[i] = mat[i];
    else
        gs_wedge_chi_matrix(wedge, chi, g.mat);
    return g;
}

int launch_xlylzl(const double* s, const double* f, const double* p, const double* r,
                  const double* dist, double* xl, int n, hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_xlylzl, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st, s, f,
                       make_detector(p, r, dist), xl, n);
    GS_LAUNCHED();
    return GS_OK;
}

int launch_gv(const double* xl, const double* omega, const double* so, const double* co,
              const gs_goniometer& gp, double* gv, int n, hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    if (so)
        hipLaunchKernelGGL(k_gv<true>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st, xl,
                           omega, so, co, gp, gv, n);
    else
        hipLaunchKernelGGL(k_gv<false>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st, xl,
                           omega, so, co, gp, gv, n);
    GS_LAUNCHED();
    return GS_OK;
}

int launch_geometry(const double* xl, const double* omega, const double* so, const double* co,
                    const gs_goniometer& gp, double* out, int n, hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    if (so)
        hipLaunchKernelGGL(k_geometry<1>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st,
                           xl, omega, so, co, gp, out, n);
    else
        hipLaunchKernelGGL(k_geometry<0>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st,
                           xl, omega, so, co, gp, out, n);
    GS_LAUNCHED();
    return GS_OK;
}

int launch_xlylzl_xpos(const double* s, const double* f, const double* p, const double* r,
                       const double* dist, const double* xpos, double* xl, int n,
                       hipStream_t st) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_xlylzl_xpos, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, st, s, f,
                       make_detector(p, r, dist), xpos, xl, n);
    GS_LAUNCHED();
    return GS_OK;
}

int check_refine_args(int K, int ng, const int* idx, long long n_idx, int n_iter) {
    if (K < 0 || n_iter < 0 || ng < 0 || n_idx < 0)
        return set_err(GS_ERR_ARG, "negative K, n_iter, ng or n_idx");
    if (!idx && n_idx > ng)
        return set_err(GS_ERR_ARG, "without idx the rows (ng = %d) must cover the n_idx = %lld "
                       "list entries", ng, n_idx);
    return GS_OK;
}

int launch_refine_lists(DeviceCtx* ctx, double* ubis, int K, const double* gv, int ng,
                        const int* idx, long long n_idx, const long long* off, const int* order,
                        double tol, int n_iter, int* n_out, double* mean_out, int lanes_per_grain,
                        int waves_per_cu, const int* tiers, int* queue, hipStream_t st) {
    const double tolsq = tol * tol;
    int lpg = lanes_per_grain;
    if (lpg != 8 && lpg != 16 && lpg != 32 && lpg != 64) lpg = 16;
    int wpc = waves_per_cu > 0 ? waves_per_cu : 8;
    const int gpw = GS_WAVE / lpg;
    // work items: at most one per grain with tiers (counted on the device)
    long long waves_needed = tiers ? K : (K + gpw - 1) / gpw;
    long long waves = std::min<long long>(waves_needed, (long long)wpc * ctx->cus);
    int nb = (int)std::max<long long>(1, (waves + GS_BLOCK / GS_WAVE - 1) / (GS_BLOCK / GS_WAVE));
    const dim3 grid(nb), block(GS_BLOCK);
#define GS_REFINE_LAUNCH(L, R)                                                                  \
    hipLaunchKernelGGL((k_refine_lists<L, R>), grid, block, 0, st, ubis, K, gv, ng, idx, n_idx, off, \
                       order, tolsq, n_iter, n_out, mean_out, tiers, queue)
    const bool rows = idx == nullptr;  // entry q is row q of gv
    switch (lpg) {
        case 8: if (rows) GS_REFINE_LAUNCH(8, true); else GS_REFINE_LAUNCH(8, false); break;
        case 32: if (rows) GS_REFINE_LAUNCH(32, true); else GS_REFINE_LAUNCH(32, false); break;
        case 64: if (rows) GS_REFINE_LAUNCH(64, true); else GS_REFINE_LAUNCH(64, false); break;
        default: if (rows) GS_REFINE_LAUNCH(16, true); else GS_REFINE_LAUNCH(16, false); break;
    }
#undef GS_REFINE_LAUNCH
    GS_LAUNCHED();
    return GS_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gs_last_error(void) { return g_err.c_str(); }

int gs_version(void) { return GS_VERSION; }

int gs_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gs_init(void) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    double* sp;
    return scratch_ptr(&sp);  // the no-allocation entry point's scratch on this device
}

// ---- reference prototypes on host pointers ---------------------------------

int verify_rounding(int n) {
    DeviceCtx* ctx;
    if (get_ctx(&ctx) != GS_OK) return -1;
    char* buf;
    if (stage_reserve(256, &buf) != GS_OK) return -1;
    int* bad = reinterpret_cast<int*>(buf);
    hipStream_t st = nullptr;
    if (fill_i32(bad, 0, 1, st) != GS_OK) return -1;
    hipLaunchKernelGGL(k_verify_rounding, dim3(3), dim3(GS_BLOCK), 0, st, n, bad);
    int h = -1;
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_err(GS_ERR_HIP, "verify_rounding: HIP error");
        return -1;
    }
    return h;
}

int score(double ubi[3][3], double gv[][3], double tol, int ng) {
    if (check_ng(ng) != GS_OK) return -1;
    if ((size_t)ng * 24 <= GS_PINNED_MAX) {
        // inputs read in place from pinned memory; the count is accumulated
        // with atomics, so it lives in device memory and comes back by copy
        std::vector<char*> h, d, sl;
        if (pinned_slices({72, 4, (size_t)ng * 24}, h, d) != GS_OK) return -1;
        if (stage_slices({4}, sl) != GS_OK) return -1;
        std::memcpy(h[0], ubi, 72);
        if (ng > 0) std::memcpy(h[2], gv, (size_t)ng * 24);
        int* dn = reinterpret_cast<int*>(sl[0]);
        hipStream_t st = nullptr;
        if (launch_score_many(reinterpret_cast<double*>(d[0]), 1,
                              reinterpret_cast<double*>(d[2]), ng, tol, dn, 0,
                              GS_SCORE_CERTIFIED, st) != GS_OK)
            return -1;
        if (hipMemcpyAsync(h[1], dn, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_err(GS_ERR_HIP, "score: copy back failed");
            return -1;
        }
        int n;
        std::memcpy(&n, h[1], 4);
        return n;
    }
    std::vector<char*> sl;
    if (stage_slices({72, (size_t)ng * 24, 4}, sl) != GS_OK) return -1;
    double* du = reinterpret_cast<double*>(sl[0]);
    double* dg = reinterpret_cast<double*>(sl[1]);
    int* dn = reinterpret_cast<int*>(sl[2]);
    hipStream_t st = nullptr;
    int h = -1;
    if (hipMemcpyAsync(du, ubi, 72, hipMemcpyHostToDevice, st) != hipSuccess ||
        (ng > 0 && hipMemcpyAsync(dg, gv, (size_t)ng * 24, hipMemcpyHostToDevice, st) != hipSuccess)) {
        set_err(GS_ERR_HIP, "score: copy to device failed");
        return -1;
    }
    if (launch_score_many(du, 1, dg, ng, tol, dn, 0, GS_SCORE_CERTIFIED, st) != GS_OK) return -1;
    if (hipMemcpyAsync(&h, dn, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_err(GS_ERR_HIP, "score: copy back failed");
        return -1;
    }
    return h;
}

static int refine_host(int mode, double* ubi, const double* gv, const int* labels, int label,
                       double tol, int* n_out, double* mean_out, int ng) {
    GS_CHECK(check_ng(ng));
    const size_t gv_bytes = (size_t)ng * 24, lab_bytes = mode ? (size_t)ng * 4 : 0;
    if (gv_bytes + lab_bytes <= GS_PINNED_MAX) {
        std::vector<char*> h, d;
        GS_CHECK(pinned_slices({72, 8, 4, gv_bytes, lab_bytes}, h, d));
        std::memcpy(h[0], ubi, 72);
        if (ng > 0) std::memcpy(h[3], gv, gv_bytes);
        if (mode && ng > 0) std::memcpy(h[4], labels, lab_bytes);
        hipStream_t st = nullptr;
        GS_CHECK(launch_refine_one(mode, reinterpret_cast<double*>(d[0]),
                                   reinterpret_cast<const double*>(d[3]), ng, tol,
                                   mode ? reinterpret_cast<const int*>(d[4]) : nullptr, label,
                                   reinterpret_cast<int*>(d[2]), reinterpret_cast<double*>(d[1]),
                                   st));
        GS_HIP(hipStreamSynchronize(st));
        std::memcpy(ubi, h[0], 72);
        std::memcpy(mean_out, h[1], 8);
        std::memcpy(n_out, h[2], 4);
        return GS_OK;
    }
    std::vector<char*> sl;
    GS_CHECK(stage_slices({72, (size_t)ng * 24, 8, 4, mode ? (size_t)ng * 4 : 0}, sl));
    double* du = reinterpret_cast<double*>(sl[0]);
    double* dg = reinterpret_cast<double*>(sl[1]);
    double* dm = reinterpret_cast<double*>(sl[2]);
    int* dn = reinterpret_cast<int*>(sl[3]);
    int* dl = mode ? reinterpret_cast<int*>(sl[4]) : nullptr;
    hipStream_t st = nullptr;
    GS_HIP(hipMemcpyAsync(du, ubi, 72, hipMemcpyHostToDevice, st));
    if (ng > 0) GS_HIP(hipMemcpyAsync(dg, gv, (size_t)ng * 24, hipMemcpyHostToDevice, st));
    if (mode && ng > 0) GS_HIP(hipMemcpyAsync(dl, labels, (size_t)ng * 4, hipMemcpyHostToDevice, st));
    GS_CHECK(launch_refine_one(mode, du, dg, ng, tol, dl, label, dn, dm, st));
    GS_HIP(hipMemcpyAsync(ubi, du, 72, hipMemcpyDeviceToHost, st));
    GS_HIP(hipMemcpyAsync(n_out, dn, 4, hipMemcpyDeviceToHost, st));
    GS_HIP(hipMemcpyAsync(mean_out, dm, 8, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

void score_and_refine(double ubi[3][3], double gv[][3], double tol, int* n_arg,
                      double* sumdrlv2_arg, int ng) {
    if (refine_host(0, &ubi[0][0], &gv[0][0], nullptr, 0, tol, n_arg, sumdrlv2_arg, ng) != GS_OK) {
        *n_arg = -1;
        *sumdrlv2_arg = NAN;
    }
}

void refine_assigned(double ubi[3][3], double gv[][3], int* labels, int label, int* npk,
                     double* sumdrlv2, int ng) {
    if (refine_host(1, &ubi[0][0], &gv[0][0], labels, label, 0.0, npk, sumdrlv2, ng) != GS_OK) {
        *npk = -1;
        *sumdrlv2 = NAN;
    }
}

int score_and_assign(double ubi[3][3], double gv[][3], double tol, double* drlv2, int* labels,
                     int label, int ng) {
    if (check_ng(ng) != GS_OK) return -1;
    std::vector<char*> sl;
    if (stage_slices({72, (size_t)ng * 24, (size_t)ng * 8, (size_t)ng * 4, 8}, sl) != GS_OK) return -1;
    double* du = reinterpret_cast<double*>(sl[0]);
    double* dg = reinterpret_cast<double*>(sl[1]);
    double* dd = reinterpret_cast<double*>(sl[2]);
    int* dl = reinterpret_cast<int*>(sl[3]);
    int* dn = reinterpret_cast<int*>(sl[4]);  // [0] count, [1] the label
    hipStream_t st = nullptr;
    int h = -1;
    bool ok = hipMemcpyAsync(du, ubi, 72, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(dn + 1, &label, 4, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok && ng > 0)
        ok = hipMemcpyAsync(dg, gv, (size_t)ng * 24, hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(dd, drlv2, (size_t)ng * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(dl, labels, (size_t)ng * 4, hipMemcpyHostToDevice, st) == hipSuccess;
    if (!ok) {
        set_err(GS_ERR_HIP, "score_and_assign: copy to device failed");
        return -1;
    }
    if (launch_assign_many(du, dn + 1, 1, dg, ng, tol, dd, dl, dn, GS_SCORE_CERTIFIED, st) != GS_OK)
        return -1;
    ok = hipMemcpyAsync(&h, dn, 4, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok && ng > 0)
        ok = hipMemcpyAsync(drlv2, dd, (size_t)ng * 8, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipMemcpyAsync(labels, dl, (size_t)ng * 4, hipMemcpyDeviceToHost, st) == hipSuccess;
    if (!ok || hipStreamSynchronize(st) != hipSuccess) {
        set_err(GS_ERR_HIP, "score_and_assign: copy back failed");
        return -1;
    }
    return h;
}

// compute_gv / compute_geometry on host arrays: sin/cos of every omega and
// the wedge/chi matrix from the host's glibc, as the reference computes
// them, the rest on the GPU (so the outputs are the reference's bits).
// `width` doubles of output per peak: 3 (gv) or 6 (tth, eta, ds, g).
static void host_gv_call(const char* what, int width, double xlylzl[][3], double omega[],
                         double omegasign, double wvln, double wedge, double chi, double t[3],
                         double* out, int n) {
    if (check_ng(n) != GS_OK || n == 0) return;
    std::vector<char*> sl, h, d;
    const size_t nb8 = (size_t)n * 8;
    if (stage_slices({nb8 * 3, nb8, nb8, nb8 * width}, sl) != GS_OK) return;
    if (pinned_slices({nb8, nb8}, h, d) != GS_OK) return;
    double* dx = reinterpret_cast<double*>(sl[0]);
    double* dso = reinterpret_cast<double*>(sl[1]);
    double* dco = reinterpret_cast<double*>(sl[2]);
    double* dout = reinterpret_cast<double*>(sl[3]);
    double* hso = reinterpret_cast<double*>(h[0]);
    double* hco = reinterpret_cast<double*>(h[1]);
    hipStream_t st = nullptr;
    if (hipMemcpyAsync(dx, xlylzl, nb8 * 3, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err(GS_ERR_HIP, "%s: copy to device failed", what);
        return;
    }
    sincos_omega_libm(omega, omegasign, hso, hco, n);
    if (hipMemcpyAsync(dso, hso, nb8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dco, hco, nb8, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err(GS_ERR_HIP, "%s: copy to device failed", what);
        return;
    }
    double mat[9];
    wedge_chi_matrix_libm(wedge, chi, mat);
    gs_goniometer gp = make_gonio(omegasign, wvln, wedge, chi, t, mat);
    if ((width == 3 ? launch_gv(dx, nullptr, dso, dco, gp, dout, n, st)
                    : launch_geometry(dx, nullptr, dso, dco, gp, dout, n, st)) != GS_OK)
        return;
    if (hipMemcpyAsync(out, dout, nb8 * width, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        set_err(GS_ERR_HIP, "%s: copy back failed", what);
}

void compute_gv(double xlylzl[][3], double omega[], double omegasign, double wvln, double wedge,
                double chi, double t[3], double gv[][3], int n) {
    host_gv_call("compute_gv", 3, xlylzl, omega, omegasign, wvln, wedge, chi, t, &gv[0][0], n);
}

void compute_geometry(double xlylzl[][3], double omega[], double omegasign, double wvln,
                      double wedge, double chi, double t[3], double out[][6], int n) {
    host_gv_call("compute_geometry", 6, xlylzl, omega, omegasign, wvln, wedge, chi, t,
                 &out[0][0], n);
}

void compute_xlylzl_xpos_variable(double s[], double f[], double p[4], double r[9],
                                  double dist[3], double xpos[], double xlylzl[][3], int n) {
    if (check_ng(n) != GS_OK || n == 0) return;
    std::vector<char*> sl;
    if (stage_slices({(size_t)n * 8, (size_t)n * 8, (size_t)n * 8, (size_t)n * 24}, sl) != GS_OK)
        return;
    double* ds = reinterpret_cast<double*>(sl[0]);
    double* df = reinterpret_cast<double*>(sl[1]);
    double* dp = reinterpret_cast<double*>(sl[2]);
    double* dx = reinterpret_cast<double*>(sl[3]);
    hipStream_t st = nullptr;
    if (hipMemcpyAsync(ds, s, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(df, f, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dp, xpos, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err(GS_ERR_HIP, "compute_xlylzl_xpos_variable: copy to device failed");
        return;
    }
    if (launch_xlylzl_xpos(ds, df, p, r, dist, dp, dx, n, st) != GS_OK) return;
    if (hipMemcpyAsync(xlylzl, dx, (size_t)n * 24, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        set_err(GS_ERR_HIP, "compute_xlylzl_xpos_variable: copy back failed");
}

void compute_xlylzl(double s[], double f[], double p[4], double r[9], double dist[3],
                    double xlylzl[][3], int n) {
    if (check_ng(n) != GS_OK || n == 0) return;
    std::vector<char*> sl;
    if (stage_slices({(size_t)n * 8, (size_t)n * 8, (size_t)n * 24}, sl) != GS_OK) return;
    double* ds = reinterpret_cast<double*>(sl[0]);
    double* df = reinterpret_cast<double*>(sl[1]);
    double* dx = reinterpret_cast<double*>(sl[2]);
    hipStream_t st = nullptr;
    if (hipMemcpyAsync(ds, s, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(df, f, (size_t)n * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err(GS_ERR_HIP, "compute_xlylzl: copy to device failed");
        return;
    }
    if (launch_xlylzl(ds, df, p, r, dist, dx, n, st) != GS_OK) return;
    if (hipMemcpyAsync(xlylzl, dx, (size_t)n * 24, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        set_err(GS_ERR_HIP, "compute_xlylzl: copy back failed");
}

// ---- device-pointer entry points -------------------------------------------

// The device sincos algorithm (gs_sincos) compiled for the host CPU: lets
// the accuracy of the exact same code be checked without a GPU.  No entry
// point of the library calls it.
void gs_selftest_sincos_host(const double* x, double* s, double* c, int n) {
    for (int i = 0; i < n; i++) gs_sincos(x[i], &s[i], &c[i]);
}

// The series path of that algorithm alone (what its table path falls back
// to), compiled for the host; *n_table (when not NULL) = how many of the n
// arguments the table path settles.  Checking aid only.
void gs_selftest_sincos_series_host(const double* x, double* s, double* c, int n, int* n_table) {
    int settled = 0;
    for (int i = 0; i < n; i++) {
        const gs_sc v = gs_sincos_series_v(x[i]);
        s[i] = v.s;
        c[i] = v.c;
        if (fabs(x[i]) < 823549.6 && x[i] != 0.0) {
            int q;
            gs_sc w;
            settled += gs_sincos_table(gs_reduce_pio2(x[i], &q), &w) ? 1 : 0;
        }
    }
    if (n_table) *n_table = settled;
}

// The certified score decision (gs_drlv2_fma + gs_cert_window) compiled
// for the host, next to the reference arithmetic (gs_drlv2): for each row
// of gv, out[k] = 1 if `drlv2 < tol^2` is certain, 0 if certainly false,
// 2 if undecided, plus 4 when the reference arithmetic says true.  gmax is
// the bound on |gv| the window is computed for.  Checking aid only.
void gs_selftest_certified_fract_host(const double* ubi, const double* gv, int n, double tol,
                                      double gmax, int* out) {
    const gs_ubi u = gs_load_ubi(ubi);
    const double tolsq = tol * tol;
    const double e = gs_cert_window_fract(u, gmax, tolsq);
    const double mid = tolsq - 0.75;
    const double lo = mid - e, hi = mid + e;
    for (int k = 0; k < n; k++) {
        const double* g = gv + 3 * (size_t)k;
        double s = gs_drlv2_fract(u, g[0], g[1], g[2]);
        int d = s < lo ? 1 : (s < hi ? 2 : 0);
        if (gs_drlv2(u, g[0], g[1], g[2]) < tolsq) d |= 4;
        out[k] = d;
    }
}

void gs_selftest_certified_host(const double* ubi, const double* gv, int n, double tol,
                                double gmax, int* out) {
    const gs_ubi u = gs_load_ubi(ubi);
    const double tolsq = tol * tol;
    const double e = gs_cert_window(u, gmax, tolsq);
    const double lo = tolsq - e, hi = tolsq + e;
    for (int k = 0; k < n; k++) {
        const double* g = gv + 3 * (size_t)k;
        double s = gs_drlv2_fma(u, g[0], g[1], g[2]);
        int d = s < lo ? 1 : (s < hi ? 2 : 0);
        if (gs_drlv2(u, g[0], g[1], g[2]) < tolsq) d |= 4;
        out[k] = d;
    }
}

// The fast-path decision of gs_gv_assign_many for single pairs, on the
// host, next to the reference arithmetic: row k has its own peak (xl, so,
// co), UBI and translation.  out[k] = 1 if `drlv2 < tol^2` is decided by the
// fast path as certainly false, 2 if left to the exact route; +4 when the
// reference's compute_gv + score arithmetic gives drlv2 < tol^2.  Checking
// aid only.
void gs_selftest_scan_host(const double* xl, const double* so, const double* co,
                           const double* ubis, const double* trans, int n, const double* mat,
                           double wvln, double tol, int* out) {
    const double ds = 1. / wvln;
    const double tolsq = tol * tol;
    for (int k = 0; k < n; k++) {
        const double* x = xl + 3 * (size_t)k;
        const double* t = trans + 3 * (size_t)k;
        const gs_ubi u = gs_load_ubi(ubis + 9 * (size_t)k);
        double a[3], c[3];
        gs_scan_consts(x[0], x[1], x[2], so[k], co[k], mat, ds, a, c);
        const double n2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        const double tt = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        const double hi = (tolsq - 0.75) + gs_scan_window(u, ds, tolsq, tt, n2);
        const double sf = gs_scan_fast_drlv2(u, a, c, t[0], t[1], t[2], ds);  // drlv2 - 3/4
        int dcs = (sf >= hi) ? 1 : 2;
        double g0, g1, g2;
        gs_gv_one(x[0], x[1], x[2], so[k], co[k], mat, ds, t[0], t[1], t[2], g0, g1, g2);
        if (gs_drlv2(u, g0, g1, g2) < tolsq) dcs |= 4;
        out[k] = dcs;
    }
}

void gs_selftest_gn_ubis_host(const double* nsums, const double* ubis, int K, double* out) {
    for (int g = 0; g < K; g++)
        gs_gn_ubi(nsums + GS_NSUMS * (size_t)g, ubis ? ubis + 9 * (size_t)g : nullptr,
                  out + 9 * (size_t)g);
}

int gs_clear_error(void) {
    g_err.clear();
    return GS_OK;
}

int gs_score_many(const double* ubis, int K, const double* gv, int ng, double tol, int* counts,
                  int peaks_per_lane, int mode, gs_stream_t stream) {
    if (mode != GS_SCORE_CERTIFIED && mode != GS_SCORE_EXACT && mode != GS_SCORE_SIEVE)
        return set_err(GS_ERR_ARG, "gs_score_many: unknown mode %d", mode);
    return launch_score_many(ubis, K, gv, ng, tol, counts, peaks_per_lane, mode, S(stream));
}

int gs_assign_many(const double* ubis, const int* grain_labels, int K, const double* gv, int ng,
                   double tol, double* drlv2, int* labels, int* counts, int mode,
                   gs_stream_t stream) {
    if (mode != GS_SCORE_CERTIFIED && mode != GS_SCORE_EXACT)
        return set_err(GS_ERR_ARG, "gs_assign_many: unknown mode %d", mode);
    return launch_assign_many(ubis, grain_labels, K, gv, ng, tol, drlv2, labels, counts, mode,
                              S(stream));
}

int gs_score_and_refine_dev(double* ubi, const double* gv, int ng, double tol, int* n_out,
                            double* mean_out, gs_stream_t stream) {
    return launch_refine_one(0, ubi, gv, ng, tol, nullptr, 0, n_out, mean_out, S(stream));
}

int gs_refine_assigned_dev(double* ubi, const double* gv, int ng, const int* labels, int label,
                           int* n_out, double* mean_out, gs_stream_t stream) {
    return launch_refine_one(1, ubi, gv, ng, 0.0, labels, label, n_out, mean_out, S(stream));
}

// gs_score_and_refine_dev without any allocation: the block partials live
// in the library's fixed `scratch` array of the current device instead of
// a buffer allocated on a thread's first call.  Same kernels, same grid,
// so the same bits as gs_score_and_refine_dev.  Nothing in the scratch has
// to start at zero (every block writes its partials before the finishing
// launch reads them); calls that overlap on one GPU share the array and
// race.
int gs_score_and_refine_dev_noalloc(double* ubi, const double* gv, int ng, double tol,
                                    int* n_out, double* mean_out, gs_stream_t stream) {
    if (ng < 0) return set_err(GS_ERR_ARG, "gs_score_and_refine_dev_noalloc: negative ng %d", ng);
    if (!ubi || !n_out || !mean_out || (ng > 0 && !gv))
        return set_err(GS_ERR_ARG, "gs_score_and_refine_dev_noalloc: null pointer");
    const int nb = refine_one_blocks(ng);
    if ((long long)nb * GS_NPART > GS_SCRATCH_DOUBLES)
        return set_err(GS_ERR_ARG, "gs_score_and_refine_dev_noalloc: %d block partials exceed "
                       "the scratch of %d doubles", nb, GS_SCRATCH_DOUBLES);
    double* partials;
    GS_CHECK(scratch_ptr(&partials));
    return refine_one_launch(0, ubi, gv, ng, tol, nullptr, 0, n_out, mean_out, partials,
                             S(stream));
}

// The UBI and the two results on the host, the peaks on the device: the
// UBI and results pass through the thread's pinned mapped buffer, so a call
// is one launch and one wait on the stream.
static int refine_host_ubi(int mode, double* ubi, const double* gv, int ng, double tol,
                           const int* labels, int label, int* n_out, double* mean_out,
                           hipStream_t st) {
    GS_CHECK(check_ng(ng));
    if (!ubi || !n_out || !mean_out) return set_err(GS_ERR_ARG, "null ubi or result pointer");
    std::vector<char*> h, d;
    GS_CHECK(pinned_slices({72, 8, 4}, h, d));
    std::memcpy(h[0], ubi, 72);
    GS_CHECK(launch_refine_one(mode, reinterpret_cast<double*>(d[0]), gv, ng, tol, labels, label,
                               reinterpret_cast<int*>(d[2]), reinterpret_cast<double*>(d[1]), st));
    GS_HIP(hipStreamSynchronize(st));
    std::memcpy(ubi, h[0], 72);
    std::memcpy(mean_out, h[1], 8);
    std::memcpy(n_out, h[2], 4);
    return GS_OK;
}

int gs_score_and_refine_host_ubi(double* ubi, const double* gv, int ng, double tol, int* n_out,
                                 double* mean_out, gs_stream_t stream) {
    return refine_host_ubi(0, ubi, gv, ng, tol, nullptr, 0, n_out, mean_out, S(stream));
}

int gs_refine_assigned_host_ubi(double* ubi, const double* gv, int ng, const int* labels,
                                int label, int* n_out, double* mean_out, gs_stream_t stream) {
    return refine_host_ubi(1, ubi, gv, ng, 0.0, labels, label, n_out, mean_out, S(stream));
}

int gs_compute_xlylzl_dev(const double* s, const double* f, const double p[4], const double r[9],
                          const double dist[3], double* xlylzl, int n, gs_stream_t stream) {
    return launch_xlylzl(s, f, p, r, dist, xlylzl, n, S(stream));
}

int gs_compute_gv_dev(const double* xlylzl, const double* omega, double omegasign, double wvln,
                      double wedge, double chi, const double t[3], double* gv, int n,
                      gs_stream_t stream) {
    gs_goniometer gp = make_gonio(omegasign, wvln, wedge, chi, t, nullptr);
    return launch_gv(xlylzl, omega, nullptr, nullptr, gp, gv, n, S(stream));
}

int gs_compute_gv_trig_dev(const double* xlylzl, const double* so, const double* co,
                           const double mat[9], double wvln, const double t[3], double* gv, int n,
                           gs_stream_t stream) {
    if (!so || !co || !mat) return set_err(GS_ERR_ARG, "gs_compute_gv_trig_dev: null so/co/mat");
    gs_goniometer gp = make_gonio(1.0, wvln, 0.0, 0.0, t, mat);
    return launch_gv(xlylzl, nullptr, so, co, gp, gv, n, S(stream));
}

int gs_sincos_omega(const double* omega, double omegasign, double* so, double* co, int n,
                    gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_sincos_omega, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       omega, omegasign, so, co, n);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_sincos_omega_host(const double* omega, double omegasign, double* so, double* co, int n) {
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    if (!omega || !so || !co) return set_err(GS_ERR_ARG, "gs_sincos_omega_host: null array");
    sincos_omega_libm(omega, omegasign, so, co, n);
    return GS_OK;
}

int gs_wedge_chi_matrix_host(double wedge, double chi, double* mat) {
    if (!mat) return set_err(GS_ERR_ARG, "gs_wedge_chi_matrix_host: null mat");
    wedge_chi_matrix_libm(wedge, chi, mat);
    return GS_OK;
}

// Device omega -> device so/co through the host's glibc: chunks of peaks
// go down to pinned host memory, their sincos is taken there, and the
// results go back up.  Returns when so/co are written.
int gs_sincos_omega_parity(const double* omega, double omegasign, double* so, double* co, int n,
                           gs_stream_t stream) {
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    if (!omega || !so || !co) return set_err(GS_ERR_ARG, "gs_sincos_omega_parity: null array");
    hipStream_t st = S(stream);
    constexpr long long CHUNK = 1LL << 21;  // 48 MiB of pinned memory at most
    const long long c = std::min<long long>(CHUNK, n);
    std::vector<char*> h, d;
    GS_CHECK(pinned_slices({(size_t)c * 8, (size_t)c * 8, (size_t)c * 8}, h, d));
    double* hw = reinterpret_cast<double*>(h[0]);
    double* hs = reinterpret_cast<double*>(h[1]);
    double* hc = reinterpret_cast<double*>(h[2]);
    for (long long a = 0; a < n; a += c) {
        const long long m = std::min<long long>(c, n - a);
        GS_HIP(hipMemcpyAsync(hw, omega + a, m * 8, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));  // also: the previous chunk's uploads are done
        sincos_omega_libm(hw, omegasign, hs, hc, m);
        GS_HIP(hipMemcpyAsync(so + a, hs, m * 8, hipMemcpyHostToDevice, st));
        GS_HIP(hipMemcpyAsync(co + a, hc, m * 8, hipMemcpyHostToDevice, st));
    }
    GS_HIP(hipStreamSynchronize(st));
    return GS_OK;
}

int gs_sf2gv_trig(const double* s, const double* f, const double* so, const double* co,
                  const double p[4], const double r[9], const double dist[3], const double mat[9],
                  double wvln, const double t[3], double* xlylzl_out, double* gv, int n,
                  gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (!so || !co || !mat) return set_err(GS_ERR_ARG, "gs_sf2gv_trig: null so/co/mat");
    if (n == 0) return GS_OK;
    gs_goniometer gp = make_gonio(1.0, wvln, 0.0, 0.0, t, mat);
    hipLaunchKernelGGL(k_sf2gv<true>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       s, f, nullptr, so, co, make_detector(p, r, dist), gp, xlylzl_out, nullptr,
                       nullptr, gv, n);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_sf2gv(const double* s, const double* f, const double* omega, const double p[4],
             const double r[9], const double dist[3], double omegasign, double wvln, double wedge,
             double chi, const double t[3], double* xlylzl_out, double* so_out, double* co_out,
             double* gv, int n, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if ((so_out == nullptr) != (co_out == nullptr))
        return set_err(GS_ERR_ARG, "gs_sf2gv: so_out and co_out must both be given or both NULL");
    if (n == 0) return GS_OK;
    gs_goniometer gp = make_gonio(omegasign, wvln, wedge, chi, t, nullptr);
    hipLaunchKernelGGL(k_sf2gv<false>, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0,
                       S(stream), s, f, omega, nullptr, nullptr, make_detector(p, r, dist), gp,
                       xlylzl_out, so_out, co_out, gv, n);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_refine_lists(double* ubis, int K, const double* gv, int ng, const int* idx, long long n_idx,
                    const long long* off, const int* order, double tol, int n_iter, int* n_out,
                    double* mean_out, int lanes_per_grain, int waves_per_cu, int* queue_ws,
                    gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_refine_args(K, ng, idx, n_idx, n_iter));
    if (K == 0 || n_iter == 0) return GS_OK;
    hipStream_t st = S(stream);
    int* queue = queue_ws ? queue_ws : ctx->counters;
    GS_CHECK(fill_i32(queue, 0, 1, st));
    return launch_refine_lists(ctx, ubis, K, gv, ng, idx, n_idx, off, order, tol, n_iter, n_out,
                               mean_out, lanes_per_grain, waves_per_cu, nullptr, queue, st);
}

// The tier counts; queue (may be NULL) is cleared by the same launch.
int launch_count_tiers(const long long* off, int K, long long n_idx, long long long_min,
                       long long mid_min, int* tiers, int* queue, hipStream_t st) {
    hipLaunchKernelGGL(k_count_tiers, dim3(1), dim3(GS_TIER_THREADS), 0, st, off, K, n_idx,
                       long_min, mid_min, tiers, queue);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_refine_lists_long(double* ubis, int K, const double* gv, int ng, const int* idx,
                         long long n_idx, const long long* off, const int* order, double tol,
                         int n_iter, int* n_out, double* mean_out, int lanes_per_grain,
                         int waves_per_cu, long long long_min, long long mid_min, int* queue_ws,
                         gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_refine_args(K, ng, idx, n_idx, n_iter));
    if (K == 0 || n_iter == 0) return GS_OK;
    hipStream_t st = S(stream);
    // queue_ws[0] the work-queue counter, queue_ws[1..2] the tier counts
    // (the library's per-device counters 0 and 2..3 without a workspace)
    int* queue = queue_ws ? queue_ws : ctx->counters;
    int* tiers_ws = queue_ws ? queue_ws + 1 : ctx->counters + 2;
    const int* tiers = nullptr;
    if (long_min > 0 || mid_min > 0) {  // the counting launch also clears the queue
        GS_CHECK(launch_count_tiers(off, K, n_idx, long_min, mid_min, tiers_ws, queue, st));
        tiers = tiers_ws;
    } else {
        GS_CHECK(fill_i32(queue, 0, 1, st));
    }
    return launch_refine_lists(ctx, ubis, K, gv, ng, idx, n_idx, off, order, tol, n_iter, n_out,
                               mean_out, lanes_per_grain, waves_per_cu, tiers, queue, st);
}

int gs_gather_rows(const double* gv, int ng, const int* idx, long long n_idx, double* rows,
                   gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (ng < 0 || n_idx < 0) return set_err(GS_ERR_ARG, "gs_gather_rows: negative ng or n_idx");
    if (n_idx == 0) return GS_OK;
    if (!idx || !rows) return set_err(GS_ERR_ARG, "gs_gather_rows: null idx or rows");
    hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(n_idx, GS_BLOCK)), dim3(GS_BLOCK), 0,
                       S(stream), gv, ng, idx, n_idx, rows);
    GS_LAUNCHED();
    return GS_OK;
}

// A captured multi-iteration refinement: one persistent refine launch per
// iteration, each on its own work-queue counter, all counters cleared by a
// fill kernel at the head of the graph.
struct gs_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int gs_refine_graph_create_long(gs_graph_t* out, double* ubis, int K, const double* gv, int ng,
                                const int* idx, long long n_idx, const long long* off,
                                const int* order, double tol, int n_iter, int* n_out,
                                double* mean_out, int lanes_per_grain, int waves_per_cu,
                                long long long_min, long long mid_min, int* queue_ws,
                                gs_stream_t stream) {
    if (!out) return set_err(GS_ERR_ARG, "gs_refine_graph_create: null output handle");
    *out = nullptr;
    if (!queue_ws) return set_err(GS_ERR_ARG, "gs_refine_graph_create: queue_ws (n_iter ints) needed");
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_refine_args(K, ng, idx, n_idx, n_iter));
    if (K == 0 || n_iter == 0) return set_err(GS_ERR_ARG, "gs_refine_graph_create: empty problem");
    hipStream_t st = S(stream);
    if (st == nullptr) return set_err(GS_ERR_ARG, "gs_refine_graph_create: capture needs a non-null stream");
    GS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    // queue_ws[0..n_iter) the per-iteration counters, queue_ws[n_iter..+1]
    // the tier counts
    int rc = fill_i32(queue_ws, 0, n_iter, st);
    const int* tiers = nullptr;
    if (rc == GS_OK && (long_min > 0 || mid_min > 0)) {
        rc = launch_count_tiers(off, K, n_idx, long_min, mid_min, queue_ws + n_iter, nullptr,
                                st);
        tiers = queue_ws + n_iter;
    }
    for (int it = 0; rc == GS_OK && it < n_iter; it++)
        rc = launch_refine_lists(ctx, ubis, K, gv, ng, idx, n_idx, off, order, tol, 1,
                                 n_out + (size_t)it * K, mean_out + (size_t)it * K,
                                 lanes_per_grain, waves_per_cu, tiers, queue_ws + it, st);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc != GS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_err(GS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_err(GS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    gs_graph_s* g = new gs_graph_s;
    g->graph = graph;
    g->exec = exec;
    *out = g;
    return GS_OK;
}

int gs_refine_graph_create(gs_graph_t* out, double* ubis, int K, const double* gv, int ng,
                           const int* idx, long long n_idx, const long long* off, const int* order,
                           double tol, int n_iter, int* n_out, double* mean_out,
                           int lanes_per_grain, int waves_per_cu, int* queue_ws,
                           gs_stream_t stream) {
    return gs_refine_graph_create_long(out, ubis, K, gv, ng, idx, n_idx, off, order, tol, n_iter,
                                       n_out, mean_out, lanes_per_grain, waves_per_cu, 0, 0,
                                       queue_ws, stream);
}

int gs_graph_launch(gs_graph_t g, gs_stream_t stream) {
    if (!g || !g->exec) return set_err(GS_ERR_ARG, "gs_graph_launch: invalid graph");
    GS_HIP(hipGraphLaunch(g->exec, S(stream)));
    return GS_OK;
}

int gs_graph_destroy(gs_graph_t g) {
    if (!g) return GS_OK;
    hipError_t e1 = g->exec ? hipGraphExecDestroy(g->exec) : hipSuccess;
    hipError_t e2 = g->graph ? hipGraphDestroy(g->graph) : hipSuccess;
    delete g;
    if (e1 != hipSuccess || e2 != hipSuccess)
        return set_err(GS_ERR_HIP, "gs_graph_destroy: %s",
                       hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    return GS_OK;
}

// ---- adjoint of the joint objective ------------------------------------------

int gs_joint_hkl(const double* sc, const double* fc, const double* so, const double* co, int npk,
                 const double* geo, const double* ubis, const double* trans, int K, const int* idx,
                 const int* owner, long long n_idx, int* hkl, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (npk < 0 || K < 0 || n_idx < 0) return set_err(GS_ERR_ARG, "gs_joint_hkl: negative size");
    if (!idx && n_idx > npk)
        return set_err(GS_ERR_ARG, "gs_joint_hkl: without idx the per-entry arrays (npk = %d) "
                       "must cover the n_idx = %lld entries", npk, n_idx);
    if (n_idx == 0) return GS_OK;
    if (idx)
        hipLaunchKernelGGL(k_joint_hkl<false>, dim3(blocks_for(n_idx, GS_BLOCK)), dim3(GS_BLOCK), 0,
                           S(stream), sc, fc, so, co, npk, geo, ubis, trans, K, idx, owner, n_idx, hkl);
    else
        hipLaunchKernelGGL(k_joint_hkl<true>, dim3(blocks_for(n_idx, GS_BLOCK)), dim3(GS_BLOCK), 0,
                           S(stream), sc, fc, so, co, npk, geo, ubis, trans, K, idx, owner, n_idx, hkl);
    GS_LAUNCHED();
    return GS_OK;
}

long long gs_joint_workspace_doubles(long long n_pieces) {
    return n_pieces < 0 ? 0 : gs_adj_ws_doubles(n_pieces);
}

int gs_joint_schedule(const long long* off, int K, long long n_idx, int piece_len,
                      long long* piece_off, int* piece_grain, int* slots, long long cap,
                      long long* n_pieces) {
    if (K < 0 || n_idx < 0) return set_err(GS_ERR_ARG, "gs_joint_schedule: negative size");
    if (piece_len < 1) return set_err(GS_ERR_ARG, "gs_joint_schedule: piece_len %d < 1", piece_len);
    if (!piece_off || !n_pieces || (K > 0 && !off))
        return set_err(GS_ERR_ARG, "gs_joint_schedule: null pointer");
    // list bounds clipped as the kernels clip them
    auto bounds = [&](int g, long long& o0, long long& o1) {
        o1 = off[g + 1] < 0 ? 0 : (off[g + 1] > n_idx ? n_idx : off[g + 1]);
        o0 = off[g] < 0 ? 0 : (off[g] > o1 ? o1 : off[g]);
    };
    long long n = 0, longest = 0;
    piece_off[0] = 0;
    for (int g = 0; g < K; g++) {
        long long o0, o1;
        bounds(g, o0, o1);
        n += (o1 - o0 + piece_len - 1) / piece_len;
        piece_off[g + 1] = n;
        longest = std::max(longest, o1 - o0);
    }
    *n_pieces = n;
    if (!piece_grain && !slots) return GS_OK;  // a query of the count
    if (!piece_grain || !slots) return set_err(GS_ERR_ARG, "gs_joint_schedule: null pointer");
    if (n > cap)
        return set_err(GS_ERR_ARG, "gs_joint_schedule: %lld pieces exceed the capacity %lld", n,
                       cap);
    if (n > INT_MAX) return set_err(GS_ERR_ARG, "gs_joint_schedule: %lld pieces", n);
    // launch order: by length, longest first, ties in piece order (a
    // counting sort over the lengths 1 .. cap, cap = the longest piece)
    const long long cap_len = std::min<long long>(piece_len, longest);
    std::vector<long long> start((size_t)cap_len + 2, 0);
    for (int g = 0; g < K; g++) {
        long long o0, o1;
        bounds(g, o0, o1);
        for (long long p = piece_off[g]; p < piece_off[g + 1]; p++) {
            piece_grain[p] = g;
            const long long q0 = o0 + (p - piece_off[g]) * piece_len;
            const long long len = std::min<long long>(piece_len, o1 - q0);
            start[(size_t)(cap_len - len) + 1]++;
        }
    }
    for (long long b = 1; b <= cap_len; b++) start[(size_t)b] += start[(size_t)b - 1];
    for (int g = 0; g < K; g++) {
        long long o0, o1;
        bounds(g, o0, o1);
        for (long long p = piece_off[g]; p < piece_off[g + 1]; p++) {
            const long long q0 = o0 + (p - piece_off[g]) * piece_len;
            const long long len = std::min<long long>(piece_len, o1 - q0);
            slots[start[(size_t)(cap_len - len)]++] = (int)p;
        }
    }
    return GS_OK;
}

int gs_joint_adjoint(const double* sc, const double* fc, const double* so, const double* co,
                     int npk, const double* geo, const double* ubis, const double* trans, int K,
                     const int* idx, long long n_idx, const long long* off, const int* hkl,
                     const long long* piece_off, const int* piece_grain, const int* slots,
                     long long n_pieces, int piece_len, int gspace, double* loss_g,
                     double* grad_ubi, double* grad_t, double* geo_part, double* nsums,
                     double* out, int lanes_per_piece, double* ws, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (npk < 0 || K < 0 || n_idx < 0 || n_pieces < 0)
        return set_err(GS_ERR_ARG, "gs_joint_adjoint: negative size");
    if (!idx && n_idx > npk)
        return set_err(GS_ERR_ARG, "gs_joint_adjoint: without idx the per-entry arrays (npk = %d) "
                       "must cover the n_idx = %lld entries", npk, n_idx);
    if (piece_len < 1) return set_err(GS_ERR_ARG, "gs_joint_adjoint: piece_len %d < 1", piece_len);
    if (!geo || !out || !piece_off || (K > 0 && (!off || !loss_g || !grad_ubi || !grad_t)) ||
        (n_pieces > 0 && (!piece_grain || !slots || !hkl)))
        return set_err(GS_ERR_ARG, "gs_joint_adjoint: null pointer");
    if (n_pieces > INT_MAX) return set_err(GS_ERR_ARG, "gs_joint_adjoint: %lld pieces", n_pieces);
    hipStream_t st = S(stream);
    int lpg = lanes_per_piece;
    if (lpg != 16 && lpg != 32 && lpg != 64) lpg = 64;
    if (!ws) GS_CHECK(adj_workspace(ctx, gs_adj_ws_doubles(n_pieces), &ws));
    double* part = ws;
    double* blk = ws + n_pieces * GS_ADJ_STRIDE;
    const long long nb = gs_adj_blocks(n_pieces, lpg);
    if (nb > INT_MAX) return set_err(GS_ERR_ARG, "gs_joint_adjoint: too many pieces");
    const dim3 grid((unsigned)nb), block(GS_BLOCK);
#define GS_ADJ_LAUNCH(L, GSP, NS)                                                                 \
    if (idx)                                                                                      \
        hipLaunchKernelGGL((k_joint_pieces<L, GSP, NS, false>), grid, block, 0, st, sc, fc, so,   \
                           co, npk, geo, ubis, trans, K, idx, n_idx, off, hkl, piece_off,         \
                           piece_grain, slots, n_pieces, piece_len, part, blk);                   \
    else                                                                                          \
        hipLaunchKernelGGL((k_joint_pieces<L, GSP, NS, true>), grid, block, 0, st, sc, fc, so,    \
                           co, npk, geo, ubis, trans, K, idx, n_idx, off, hkl, piece_off,         \
                           piece_grain, slots, n_pieces, piece_len, part, blk)
#define GS_ADJ_LPG(GSP, NS)                         \
    switch (lpg) {                                  \
        case 16: GS_ADJ_LAUNCH(16, GSP, NS); break; \
        case 32: GS_ADJ_LAUNCH(32, GSP, NS); break; \
        default: GS_ADJ_LAUNCH(64, GSP, NS); break; \
    }
    if (nb > 0) {
        if (gspace) {
            if (nsums) { GS_ADJ_LPG(true, true) } else { GS_ADJ_LPG(true, false) }
        } else {
            if (nsums) { GS_ADJ_LPG(false, true) } else { GS_ADJ_LPG(false, false) }
        }
        GS_LAUNCHED();
    }
#undef GS_ADJ_LPG
#undef GS_ADJ_LAUNCH
    const dim3 fgrid((unsigned)(gs_fin_grain_blocks(K) + 1)), fblock(GS_FIN_THREADS);
#define GS_FIN_LAUNCH(GSP, NS)                                                                  \
    hipLaunchKernelGGL((k_joint_finish<GSP, NS>), fgrid, fblock, 0, st, part, piece_off,         \
                       n_pieces, K, ubis, geo, loss_g, grad_ubi, grad_t, geo_part, nsums, blk, nb, \
                       out)
    if (gspace) {
        if (nsums) { GS_FIN_LAUNCH(true, true); } else { GS_FIN_LAUNCH(true, false); }
    } else {
        if (nsums) { GS_FIN_LAUNCH(false, true); } else { GS_FIN_LAUNCH(false, false); }
    }
#undef GS_FIN_LAUNCH
    GS_LAUNCHED();
    return GS_OK;
}

int gs_gauss_newton_ubis(const double* nsums, const double* ubis, int K, double* out,
                         gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0) return set_err(GS_ERR_ARG, "gs_gauss_newton_ubis: K = %d < 0", K);
    if (K == 0) return GS_OK;
    if (!nsums || !out) return set_err(GS_ERR_ARG, "gs_gauss_newton_ubis: null pointer");
    hipLaunchKernelGGL(k_gn_ubis, dim3(blocks_for(K, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       nsums, ubis, K, out);
    GS_LAUNCHED();
    return GS_OK;
}

// ---- compute_gv per grain fused with score_and_assign -------------------------

int gs_gv_assign_many(const double* xlylzl, const double* so, const double* co, int ng,
                      double wvln, double wedge, double chi, const double* mat,
                      const double* ubis, const double* trans, const int* grain_labels, int K,
                      double tol, double* drlv2, int* labels, int* counts, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0 || ng < 0) return set_err(GS_ERR_ARG, "negative size K=%d ng=%d", K, ng);
    hipStream_t st = S(stream);
    GS_CHECK(fill_i32(counts, 0, K, st));
    if (K == 0 || ng == 0) return GS_OK;
    gs_goniometer gp = make_gonio(1.0, wvln, wedge, chi, nullptr, mat);
    constexpr int P = 2;
    const long long nb = (ng + (long long)GS_BLOCK * P - 1) / ((long long)GS_BLOCK * P);
    hipLaunchKernelGGL(k_gv_assign_many<P>, dim3((unsigned)nb), dim3(GS_BLOCK), 0, st, xlylzl, so,
                       co, ng, gp, ubis, trans, grain_labels, K, tol * tol, drlv2, labels, counts);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_gv_lists(const double* xlylzl, const double* so, const double* co, int ng, double wvln,
                double wedge, double chi, const double* mat, const double* trans, int K,
                const int* idx, const int* owner, long long n_idx, double* out,
                gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0 || ng < 0 || n_idx < 0) return set_err(GS_ERR_ARG, "gs_gv_lists: negative size");
    if (n_idx == 0) return GS_OK;
    gs_goniometer gp = make_gonio(1.0, wvln, wedge, chi, nullptr, mat);
    hipLaunchKernelGGL(k_gv_lists, dim3(blocks_for(n_idx, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       xlylzl, so, co, ng, gp, trans, K, idx, owner, n_idx, out);
    GS_LAUNCHED();
    return GS_OK;
}

// ---- refinegrains driver -------------------------------------------------------

int gs_grain_gv(const double* xlylzl, const double* so, const double* co, const double* omega,
                int ng, double omegasign, double wvln, double wedge, double chi,
                const double* trans, const double* mats, int R, const double* rb,
                const double* pre, double slop, const int* idx, const int* owner, long long M,
                double* tth, double* eta, double* omega_calc, double* gv, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (R < 0 || ng < 0 || M < 0) return set_err(GS_ERR_ARG, "gs_grain_gv: negative size");
    if (M == 0) return GS_OK;
    if (!xlylzl || !so || !co || !trans || !idx || !owner || !gv)
        return set_err(GS_ERR_ARG, "gs_grain_gv: NULL array");
    if (mats && (!omega || !rb))
        return set_err(GS_ERR_ARG, "gs_grain_gv: omega-float needs omega and rb");
    gs_goniometer gp = make_gonio(omegasign, wvln, wedge, chi, nullptr, nullptr);
    gs_omfloat of;
    for (int i = 0; i < 3; i++) of.rb[i] = rb ? rb[i] : 0.0;
    for (int i = 0; i < 9; i++) of.pre[i] = pre ? pre[i] : 0.0;
    of.has_pre = pre != nullptr;
    of.slop = slop;
    const dim3 grid(blocks_for(M, GS_BLOCK));
    if (mats)
        hipLaunchKernelGGL(k_grain_gv<1>, grid, dim3(GS_BLOCK), 0, S(stream), xlylzl, so, co, omega,
                           ng, gp, trans, mats, R, of, idx, owner, M, tth, eta, omega_calc, gv);
    else
        hipLaunchKernelGGL(k_grain_gv<0>, grid, dim3(GS_BLOCK), 0, S(stream), xlylzl, so, co, omega,
                           ng, gp, trans, mats, R, of, idx, owner, M, tth, eta, omega_calc, gv);
    GS_LAUNCHED();
    return GS_OK;
}

// the kernel's per-entry arithmetic run on the host (no device needed):
// host arrays, same arguments and outputs as gs_grain_gv
int gs_grain_gv_host(const double* xlylzl, const double* so, const double* co,
                     const double* omega, int ng, double omegasign, double wvln, double wedge,
                     double chi, const double* trans, const double* mats, int R, const double* rb,
                     const double* pre, double slop, const int* idx, const int* owner, long long M,
                     double* tth, double* eta, double* omega_calc, double* gv) {
    if (R < 0 || ng < 0 || M < 0) return set_err(GS_ERR_ARG, "gs_grain_gv_host: negative size");
    if (mats && (!omega || !rb))
        return set_err(GS_ERR_ARG, "gs_grain_gv_host: omega-float needs omega and rb");
    gs_omfloat of;
    for (int i = 0; i < 3; i++) of.rb[i] = rb ? rb[i] : 0.0;
    for (int i = 0; i < 9; i++) of.pre[i] = pre ? pre[i] : 0.0;
    of.has_pre = pre != nullptr;
    of.slop = slop;
    double mat[9];
    gs_wedge_chi_matrix(wedge, chi, mat);
    const double ds = 1. / wvln;
    const double nan = std::nan("");
    for (long long q = 0; q < M; q++) {
        double out[6] = {nan, nan, nan, nan, nan, nan};
        const int k = idx[q], r = owner[q];
        if (k >= 0 && k < ng && r >= 0 && r < R)
            gs_grain_gv_one(of, mat, ds, xlylzl[3 * (size_t)k], xlylzl[3 * (size_t)k + 1],
                            xlylzl[3 * (size_t)k + 2], so[k], co[k],
                            mats ? omega[k] * omegasign : 0.0, trans + 3 * (size_t)r,
                            mats ? mats + 9 * (size_t)r : nullptr, out);
        if (tth) tth[q] = out[0];
        if (eta) eta[q] = out[1];
        if (mats && omega_calc) omega_calc[q] = out[2];
        for (int j = 0; j < 3; j++) gv[3 * q + j] = out[3 + j];
    }
    return GS_OK;
}

// ---- point-by-point refinement ------------------------------------------------

int gs_pbp_assign(const double* so, const double* co, const double* dty, const double* gv,
                  int ng, const double* pts, const int* ptask, int P, const double* ubis, int T,
                  double y0, double ystep, double tol, int* counts, const long long* task_off,
                  int* cursor, int* out_peak, int* out_hkl, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (ng < 0 || P < 0 || T < 0) return set_err(GS_ERR_ARG, "gs_pbp_assign: negative size");
    if (!counts) return set_err(GS_ERR_ARG, "gs_pbp_assign: counts is NULL");
    const bool fill = task_off != nullptr;
    if (fill && (!cursor || !out_peak || !out_hkl))
        return set_err(GS_ERR_ARG, "gs_pbp_assign: the fill pass needs cursor and outputs");
    hipStream_t st = S(stream);
    if (!fill) GS_CHECK(fill_i32(counts, 0, T, st));
    else GS_CHECK(fill_i32(cursor, 0, T, st));
    if (ng == 0 || P == 0 || T == 0) return GS_OK;
    if (!so || !co || !dty || !gv || !pts || !ptask || !ubis)
        return set_err(GS_ERR_ARG, "gs_pbp_assign: NULL array");
    const long long ny = (P + GS_PBP_POINTS - 1) / GS_PBP_POINTS;
    if (ny > 65535) return set_err(GS_ERR_ARG, "gs_pbp_assign: too many points (%d)", P);
    const dim3 grid(blocks_for(ng, GS_BLOCK), (unsigned)ny);
    if (fill)
        hipLaunchKernelGGL(k_pbp_assign<1>, grid, dim3(GS_BLOCK), 0, st, so, co, dty, gv, ng, pts,
                           ptask, P, ubis, y0, ystep, tol * tol, counts, task_off, cursor,
                           out_peak, out_hkl);
    else
        hipLaunchKernelGGL(k_pbp_assign<0>, grid, dim3(GS_BLOCK), 0, st, so, co, dty, gv, ng, pts,
                           ptask, P, ubis, y0, ystep, tol * tol, counts, task_off, cursor,
                           out_peak, out_hkl);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_pbp_merge_sums(const long long* perm, const long long* peak, long long npk,
                      const double* si, const double* sc, const double* fc, const double* om,
                      const double* dty, const double* xpos, const double* eta,
                      const long long* goff, int G, double* out, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (G < 0 || npk < 0) return set_err(GS_ERR_ARG, "gs_pbp_merge_sums: negative size");
    if (G == 0) return GS_OK;
    gs_merge_cols c{si, sc, fc, om, dty, xpos, eta};
    hipLaunchKernelGGL(k_pbp_merge_sums, dim3(blocks_for(G, GS_BLOCK)), dim3(GS_BLOCK), 0,
                       S(stream), perm, peak, c, goff, G, npk, out);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_pbp_ubi_fit(const double* design, const long long* off, const int* fit,
                   const double* ubi_in, const double B0[9], int T, double* ubifit, int* rank,
                   int* worth, double* ubi_out, double* ub0, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (T < 0) return set_err(GS_ERR_ARG, "gs_pbp_ubi_fit: negative task count");
    if (T == 0) return GS_OK;
    if (!B0) return set_err(GS_ERR_ARG, "gs_pbp_ubi_fit: null B0");
    gs_mat9 b;
    for (int i = 0; i < 9; i++) b.m[i] = B0[i];
    hipLaunchKernelGGL(k_pbp_ubi_fit, dim3(blocks_for(T, 64)), dim3(64), 0, S(stream), design, off,
                       fit, ubi_in, b, T, ubifit, rank, worth, ubi_out, ub0);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_pbp_strain_fit(const double* design, const long long* off, const int* worth, int T,
                      double* eps, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (T < 0) return set_err(GS_ERR_ARG, "gs_pbp_strain_fit: negative task count");
    if (T == 0) return GS_OK;
    hipLaunchKernelGGL(k_pbp_strain_fit, dim3(blocks_for(T, 64)), dim3(64), 0, S(stream), design,
                       off, worth, T, eps);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_pbp_rows_dot(const double* ubis, const long long* task, long long R, const double* g,
                    double tolsq, double* h, long long* hkl, unsigned char* ok,
                    gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (R < 0) return set_err(GS_ERR_ARG, "gs_pbp_rows_dot: negative row count");
    if (R == 0) return GS_OK;
    if (!ubis || !g || (!h && !hkl) || (hkl && !ok) || (ok && !hkl))
        return set_err(GS_ERR_ARG, "gs_pbp_rows_dot: NULL array");
    hipLaunchKernelGGL(k_pbp_rows_dot, dim3(blocks_for(R, GS_BLOCK)), dim3(GS_BLOCK), 0,
                       S(stream), ubis, task, R, g, tolsq, h, hkl, ok);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_segment_sum(const double* v, int C, const long long* off, int nseg, double* out,
                   gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (C < 0 || nseg < 0) return set_err(GS_ERR_ARG, "gs_segment_sum: negative size");
    if ((long long)C * nseg == 0) return GS_OK;
    if (!v || !off || !out) return set_err(GS_ERR_ARG, "gs_segment_sum: NULL array");
    hipLaunchKernelGGL(k_segment_sum<false>, dim3(blocks_for((long long)C * nseg, GS_BLOCK)),
                       dim3(GS_BLOCK), 0, S(stream), v, C, off, nseg, out);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_segment_max(const double* v, int C, const long long* off, int nseg, double* out,
                   gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (C < 0 || nseg < 0) return set_err(GS_ERR_ARG, "gs_segment_max: negative size");
    if ((long long)C * nseg == 0) return GS_OK;
    if (!v || !off || !out) return set_err(GS_ERR_ARG, "gs_segment_max: NULL array");
    hipLaunchKernelGGL(k_segment_sum<true>, dim3(blocks_for((long long)C * nseg, GS_BLOCK)),
                       dim3(GS_BLOCK), 0, S(stream), v, C, off, nseg, out);
    GS_LAUNCHED();
    return GS_OK;
}

// ---- column-file geometry ----------------------------------------------------

int gs_compute_geometry_dev(const double* xlylzl, const double* omega, double omegasign,
                            double wvln, double wedge, double chi, const double t[3], double* out,
                            int n, gs_stream_t stream) {
    gs_goniometer gp = make_gonio(omegasign, wvln, wedge, chi, t, nullptr);
    return launch_geometry(xlylzl, omega, nullptr, nullptr, gp, out, n, S(stream));
}

int gs_compute_geometry_trig_dev(const double* xlylzl, const double* so, const double* co,
                                 const double mat[9], double wvln, const double t[3], double* out,
                                 int n, gs_stream_t stream) {
    if (!so || !co || !mat) return set_err(GS_ERR_ARG, "gs_compute_geometry_trig_dev: NULL");
    gs_goniometer gp = make_gonio(1.0, wvln, 0.0, 0.0, t, mat);
    return launch_geometry(xlylzl, nullptr, so, co, gp, out, n, S(stream));
}

int gs_compute_xlylzl_xpos_dev(const double* s, const double* f, const double p[4],
                               const double r[9], const double dist[3], const double* xpos,
                               double* xlylzl, int n, gs_stream_t stream) {
    return launch_xlylzl_xpos(s, f, p, r, dist, xpos, xlylzl, n, S(stream));
}

int gs_sf2geometry(const double* s, const double* f, const double* omega, const double p[4],
                   const double r[9], const double dist[3], double omegasign, double wvln,
                   double wedge, double chi, const double t[3], double* xlylzl_out, double* out,
                   int n, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(n));
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_sf2geometry, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       s, f, omega, make_detector(p, r, dist),
                       make_gonio(omegasign, wvln, wedge, chi, t, nullptr), xlylzl_out, out, n);
    GS_LAUNCHED();
    return GS_OK;
}

// ---- indexer ---------------------------------------------------------------

void closest(double x[], double v[], int* ibest, double* best, int nx, int nv) {
    if (nx < 0 || nv < 0) {
        set_err(GS_ERR_ARG, "closest: negative length");
        return;
    }
    std::vector<char*> sl;
    if (stage_slices({(size_t)nx * 8, (size_t)nv * 8, 4, 8}, sl) != GS_OK) return;
    double* dx = reinterpret_cast<double*>(sl[0]);
    double* dv = reinterpret_cast<double*>(sl[1]);
    int* di = reinterpret_cast<int*>(sl[2]);
    double* db = reinterpret_cast<double*>(sl[3]);
    hipStream_t st = nullptr;
    if ((nx > 0 && hipMemcpyAsync(dx, x, (size_t)nx * 8, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (nv > 0 && hipMemcpyAsync(dv, v, (size_t)nv * 8, hipMemcpyHostToDevice, st) != hipSuccess)) {
        set_err(GS_ERR_HIP, "closest: copy to device failed");
        return;
    }
    hipLaunchKernelGGL(k_closest, dim3(1), dim3(GS_BLOCK), 0, st, dx, nx, dv, nv, di, db);
    if (hipGetLastError() != hipSuccess) {
        set_err(GS_ERR_HIP, "closest: kernel launch failed");
        return;
    }
    int hi = 0;
    double hb = 0.0;
    if (hipMemcpyAsync(&hi, di, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&hb, db, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_err(GS_ERR_HIP, "closest: copy back failed");
        return;
    }
    *ibest = hi;
    *best = hb;
}

void quickorient(double UBI[9], double BT[9]) {
    std::vector<char*> sl;
    if (stage_slices({72, 72}, sl) != GS_OK) return;
    double* du = reinterpret_cast<double*>(sl[0]);
    double* db = reinterpret_cast<double*>(sl[1]);
    hipStream_t st = nullptr;
    if (hipMemcpyAsync(du, UBI, 72, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(db, BT, 72, hipMemcpyHostToDevice, st) != hipSuccess) {
        set_err(GS_ERR_HIP, "quickorient: copy to device failed");
        return;
    }
    hipLaunchKernelGGL(k_quickorient, dim3(1), dim3(64), 0, st, du, db);
    if (hipGetLastError() != hipSuccess) {
        set_err(GS_ERR_HIP, "quickorient: kernel launch failed");
        return;
    }
    if (hipMemcpyAsync(UBI, du, 72, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        set_err(GS_ERR_HIP, "quickorient: copy back failed");
}

int gs_gv_ds(const double* gv, int ng, double* ds, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    GS_CHECK(check_ng(ng));
    if (ng == 0) return GS_OK;
    hipLaunchKernelGGL(k_gv_ds, dim3(blocks_for(ng, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream), gv,
                       ng, ds);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_ring_assign(const double* ds, int n, const double* ringds, int nr, double tol, int* ra,
                   gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (n < 0 || nr < 0) return set_err(GS_ERR_ARG, "gs_ring_assign: negative size");
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_ring_assign, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       ds, n, ringds, nr, tol, ra);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_unit_rows(const double* gv, int ng, const int* idx, int n, double* out,
                 gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (n < 0 || ng < 0) return set_err(GS_ERR_ARG, "gs_unit_rows: negative size");
    if (n == 0) return GS_OK;
    hipLaunchKernelGGL(k_unit_rows, dim3(blocks_for(n, GS_BLOCK)), dim3(GS_BLOCK), 0, S(stream),
                       gv, ng, idx, n, out);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_find_hits(const double* n1, int N1, const double* n2, int N2, const double* cs, int nv,
                 int* ibest, double* best, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (N1 < 0 || N2 < 0 || nv < 0) return set_err(GS_ERR_ARG, "gs_find_hits: negative size");
    if (nv > GS_MAX_COSES)
        return set_err(GS_ERR_ARG, "gs_find_hits: %d cosines (at most %d)", nv, GS_MAX_COSES);
    if (N1 == 0) return GS_OK;
    const int nb = (N1 + GS_FIND_ROWS - 1) / GS_FIND_ROWS;
    hipLaunchKernelGGL(k_find_hits, dim3(nb), dim3(GS_BLOCK), 0, S(stream), n1, N1, n2, N2, cs, nv,
                       ibest, best);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_find_all(const double* n1, int N1, const double* n2, int N2, const double* cs, int nv,
                double mtol, unsigned long long* counter, long long* keys, long long cap,
                gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (N1 < 0 || N2 < 0 || nv < 0 || cap < 0)
        return set_err(GS_ERR_ARG, "gs_find_all: negative size");
    if (nv > GS_MAX_COSES)
        return set_err(GS_ERR_ARG, "gs_find_all: %d cosines (at most %d)", nv, GS_MAX_COSES);
    if (!counter) return set_err(GS_ERR_ARG, "gs_find_all: counter is NULL");
    GS_CHECK(fill_i32(reinterpret_cast<int*>(counter), 0, 2, S(stream)));
    if (N1 == 0) return GS_OK;
    hipLaunchKernelGGL(k_find_all, dim3(N1), dim3(GS_BLOCK), 0, S(stream), n1, N1, n2, N2, cs, nv,
                       mtol, counter, keys, cap);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_orient_hits(const double* gv, int ng, const int* hi, const int* hj, int nh,
                   const double* c2ab, int nc, const double* bt, const int* bsel, double* ubis,
                   int* choice, double* costheta, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (ng < 0 || nh < 0) return set_err(GS_ERR_ARG, "gs_orient_hits: negative size");
    if (nc < 1) return set_err(GS_ERR_ARG, "gs_orient_hits: empty cosine table");
    if (nh == 0) return GS_OK;
    hipLaunchKernelGGL(k_orient_hits, dim3(blocks_for(nh, GS_BLOCK)), dim3(GS_BLOCK), 0,
                       S(stream), gv, ng, hi, hj, nh, c2ab, nc, bt, bsel, ubis, choice,
                       costheta);
    GS_LAUNCHED();
    return GS_OK;
}

int gs_select_many(const double* ubis, int K, const double* gv, int ng, double thr,
                   const long long* off, int* cursor, int* idx, gs_stream_t stream) {
    DeviceCtx* ctx;
    GS_CHECK(get_ctx(&ctx));
    if (K < 0 || ng < 0) return set_err(GS_ERR_ARG, "gs_select_many: negative size");
    if (idx && !off) return set_err(GS_ERR_ARG, "gs_select_many: idx needs off");
    GS_CHECK(fill_i32(cursor, 0, K, S(stream)));
    if (K == 0 || ng == 0) return GS_OK;
    constexpr int P = 4;
    int gsplit, nsplit;
    unsigned nb;
    if (thr <= GS_SIEVE_THR_MAX) {  // few pairs pass row 0: the sieved selection
        GS_CHECK(pair_tiles(ctx, K, ng, P, &gsplit, &nsplit, &nb, GS_SIEVE_SPLIT_MAX,
                            GS_SIEVE_TILES_PER_CU));
        if (gsplit > GS_SIEVE_SPLIT_MAX) return set_err(GS_ERR_ARG, "sieve split %d too large", gsplit);
        hipLaunchKernelGGL((k_score_many_sieve<P, true>), dim3(nb), dim3(GS_BLOCK), 0, S(stream),
                           ubis, K, gsplit, nsplit, gv, ng, thr, nullptr, off, cursor, idx);
        GS_LAUNCHED();
        return GS_OK;
    }
    GS_CHECK(pair_tiles(ctx, K, ng, P, &gsplit, &nsplit, &nb));
    hipLaunchKernelGGL(k_select_many<P>, dim3(nb), dim3(GS_BLOCK), 0, S(stream), ubis, K, gsplit,
                       nsplit, gv, ng, thr, off, cursor, idx);
    GS_LAUNCHED();
    return GS_OK;
}

// One iteration of B Nelder-Mead simplexes stepped in lockstep, split at
// the points where the objective is called (ImageD11/simplex.py:115-227,
// the reference's one-problem loop; the batched driver around these calls
// is SimplexBatch in imaged11_amd/simplex.py).  S (B, n+3, n): the n+1
// vertices, the centroid and the reflected point of every problem; E
// (B, n+1) their errors.  Comparisons follow the reference's `<` / `>`
// (a NaN never wins), sums run in vertex order, and no product is fused
// into an addition, so every value is the one the reference computes.
//
// gs_simplex_order: per active problem (done[b] == 0) the highest, lowest
// and second-highest vertices, the convergence test (standard deviation of
// the errors / n <= epsilon: done[b] = 1), lowest/iters bookkeeping, and
// for the problems that go on (a[b] = 1) the centroid of all vertices but
// the highest into S[b, n+1] and the reflected point into S[b, n+2].
int gs_simplex_order(int B, int n, double* S, const double* E, unsigned char* done,
                     long long* lowest, long long* iters, int it, double epsilon, double kR,
                     long long* hi, long long* lo, long long* second, unsigned char* a,
                     int* n_active) {
    if (B < 0 || n < 1 || n > 64 || !S || !E || !done || !lowest || !iters || !hi || !lo || !second || !a ||
        !n_active)
        return set_err(GS_ERR_ARG, "gs_simplex_order: bad arguments");
    const int nv = n + 1, ns = n + 3;
    int na = 0;
    for (int b = 0; b < B; b++) {
        a[b] = 0;
        if (done[b]) continue;
        const double* e = E + (size_t)b * nv;
        long long h = 0, l = 0;
        for (int v = 1; v < nv; v++) {
            if (e[v] > e[h]) h = v;
            if (e[v] < e[l]) l = v;
        }
        long long sec = l;
        for (int v = 0; v < nv; v++)
            if (v != h && v != sec && e[v] > e[sec]) sec = v;
        double tot = 0.0;
        for (int v = 0; v < nv; v++) tot = tot + e[v];
        const double mean = tot / (double)nv;
        double ss = 0.0;
        for (int v = 0; v < nv; v++) {
            const double d = e[v] - mean;
            ss = ss + d * d;
        }
        hi[b] = h;
        lo[b] = l;
        second[b] = sec;
        lowest[b] = l;
        iters[b] = it;
        if (std::sqrt(ss / (double)n) <= epsilon) {
            done[b] = 1;
            continue;
        }
        a[b] = 1;
        na++;
        double* sb = S + (size_t)b * ns * n;
        double* cen = sb + (size_t)(n + 1) * n;
        double* refl = sb + (size_t)(n + 2) * n;
        const double* sh = sb + (size_t)h * n;
        for (int j = 0; j < n; j++) {
            double c = 0.0;
            for (int v = 0; v < nv; v++)
                if (v != h) c = c + sb[(size_t)v * n + j];
            cen[j] = c / (double)n;
        }
        for (int j = 0; j < n; j++) {
            const double x = kR * sh[j];
            const double y = (1.0 - kR) * cen[j];
            refl[j] = x + y;
        }
    }
    *n_active = na;
    return GS_OK;
}

// gs_simplex_reflected: fr[b] = the error at the reflected point of every
// active problem.  The reflected point replaces the highest vertex when it
// is lower; then expansion (fr <= the lowest error) or contraction (fr >=
// the second-highest error) is chosen and its point written to trial[b]
// (step[b] = 1, expand[b] = 1 for an expansion).
int gs_simplex_reflected(int B, int n, double* S, double* E, const unsigned char* a,
                         const long long* hi, const long long* lo, const long long* second,
                         const double* fr, double kE, double kC, double* trial,
                         unsigned char* expand, unsigned char* step, int* n_step) {
    if (B < 0 || n < 1 || n > 64 || !S || !E || !a || !hi || !lo || !second || !fr || !trial || !expand ||
        !step || !n_step)
        return set_err(GS_ERR_ARG, "gs_simplex_reflected: bad arguments");
    const int nv = n + 1, ns = n + 3;
    int nt = 0;
    for (int b = 0; b < B; b++) {
        expand[b] = 0;
        step[b] = 0;
        if (!a[b]) continue;
        double* e = E + (size_t)b * nv;
        double* sb = S + (size_t)b * ns * n;
        const double* cen = sb + (size_t)(n + 1) * n;
        const double* refl = sb + (size_t)(n + 2) * n;
        const long long h = hi[b];
        const double f = fr[b];
        if (f < e[h]) {
            e[h] = f;
            for (int j = 0; j < n; j++) sb[(size_t)h * n + j] = refl[j];
        }
        const bool ex = f <= e[lo[b]];
        const bool co = !ex && f >= e[second[b]];
        double* t = trial + (size_t)b * n;
        for (int j = 0; j < n; j++) {
            if (ex) {
                const double x = kE * refl[j];
                const double y = (1.0 - kE) * cen[j];
                t[j] = x + y;
            } else {
                const double x = kC * sb[(size_t)h * n + j];
                const double y = (1.0 - kC) * cen[j];
                t[j] = x + y;
            }
        }
        expand[b] = ex;
        step[b] = ex || co;
        nt += step[b];
    }
    *n_step = nt;
    return GS_OK;
}

// gs_simplex_trial: fb[b] = the error at trial[b] of every problem with
// step[b].  The trial point replaces the highest vertex when it is lower;
// a contraction that fails shrinks every vertex but the lowest halfway
// towards it (shrink[b] = 1); the caller then evaluates those vertices in
// vertex order.
int gs_simplex_trial(int B, int n, double* S, double* E, const unsigned char* step,
                     const unsigned char* expand, const long long* hi, const long long* lo,
                     const double* fb, const double* trial, unsigned char* shrink,
                     int* n_shrink) {
    if (B < 0 || n < 1 || n > 64 || !S || !E || !step || !expand || !hi || !lo || !fb || !trial ||
        !shrink || !n_shrink)
        return set_err(GS_ERR_ARG, "gs_simplex_trial: bad arguments");
    const int nv = n + 1, ns = n + 3;
    int nk = 0;
    for (int b = 0; b < B; b++) {
        shrink[b] = 0;
        if (!step[b]) continue;
        double* e = E + (size_t)b * nv;
        double* sb = S + (size_t)b * ns * n;
        const long long h = hi[b], l = lo[b];
        const double* t = trial + (size_t)b * n;
        const bool take = fb[b] < e[h];
        if (take) {
            e[h] = fb[b];
            for (int j = 0; j < n; j++) sb[(size_t)h * n + j] = t[j];
        }
        if (expand[b] || take) continue;
        shrink[b] = 1;
        nk++;
        double base[64];
        for (int j = 0; j < n; j++) base[j] = sb[(size_t)l * n + j];
        for (int v = 0; v < nv; v++) {
            if (v == l) continue;
            for (int j = 0; j < n; j++)
                sb[(size_t)v * n + j] = 0.5 * (sb[(size_t)v * n + j] + base[j]);
        }
    }
    *n_shrink = nk;
    return GS_OK;
}

// The decision loop of indexer.scorethem (ImageD11/indexing.py:778-838) on
// host arrays, over follow-ups computed beforehand for every hit: visited
// from the last hit to the first, a hit is skipped when either peak is
// already assigned or i == j, or when its trial score is <= minpks; its
// follow-up's peak list then becomes a grain if the fraction of those peaks
// not yet assigned exceeds `uniqueness` (0/0 is NaN: never).  Sequential by
// definition (each accept marks peaks the later hits test), so it runs on
// the host, in native code.
int gs_index_decide(int nhits, const int* hit_i, const int* hit_j, const int* npk, int minpks,
                    const int* cand, const int* lists, const long long* list_off, int ncand,
                    double uniqueness, int max_grains, int label_base, int* ga, int npeaks,
                    int* accepted, int* n_accepted, int* stop_at, int* nuniq) {
    if (nhits < 0 || npeaks < 0 || ncand < 0)
        return set_err(GS_ERR_ARG, "gs_index_decide: negative size");
    if (!n_accepted || !stop_at || !nuniq)
        return set_err(GS_ERR_ARG, "gs_index_decide: null output");
    int h = nhits, ng = 0, nu = 0;
    while (h > 0 && ng < max_grains) {
        h--;
        const int i = hit_i[h], j = hit_j[h];
        if (i < 0 || i >= npeaks || j < 0 || j >= npeaks)
            return set_err(GS_ERR_ARG, "gs_index_decide: hit %d names peak %d / %d outside [0, %d)",
                           h, i, j, npeaks);
        if (ga[i] > -1 || ga[j] > -1 || i == j) continue;
        if (npk[h] <= minpks) continue;
        const int c = cand[h];
        if (c < 0 || c >= ncand)
            return set_err(GS_ERR_ARG, "gs_index_decide: hit %d has no follow-up", h);
        const long long a = list_off[c], b = list_off[c + 1];
        long long free = 0;
        for (long long q = a; q < b; q++) {
            const int k = lists[q];
            if (k < 0 || k >= npeaks)
                return set_err(GS_ERR_ARG, "gs_index_decide: peak %d outside [0, %d)", k, npeaks);
            free += ga[k] == -1;
        }
        const double uniq = (double)free * 1.0 / (double)(b - a);
        if (uniq > uniqueness) {
            const int label = label_base + ng + 1;
            for (long long q = a; q < b; q++) ga[lists[q]] = label;
            accepted[ng] = h;
            ng++;
        } else {
            nu++;
        }
    }
    *n_accepted = ng;
    *stop_at = h;
    *nuniq = nu;
    return GS_OK;
}

}  // extern "C"
