/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C restatement of ImageD11's grain-scoring path, used as the
 * checker for the HIP kernels in imaged11_amd/.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it; the product path
 * never calls into this file.
 *
 * Every function follows the reference operation by operation, because the
 * parity bar for counts, labels and drlv2 is bit-exactness: the same
 * left-to-right FP64 expression order, no FMA contraction (build with
 * -ffp-contract=off, no -march), the same MAGIC rounding, the same
 * macro-expansion order for the degree->radian conversions.
 *
 * Pinning (what checks each function independently of its own text):
 *   - oracle/_ref/libref.so, the reference C compiled from its sources,
 *     bit-for-bit on seeded inputs (tests/test_oracle_vs_ref.py) and through
 *     the committed golden vectors (tests/golden/);
 *   - closed forms: verify_rounding == 0, ties-to-even on half integers,
 *     exact lattices give drlv2 == 0, singular normal matrices leave the
 *     UBI unchanged, the 3x3 inverse against numpy.linalg.inv;
 *   - an independent numpy formula chain for compute_gv (tth/eta route of
 *     transform.compute_tth_eta_from_xyz + compute_g_vectors) and for
 *     compute_xlylzl (transform.compute_xyz_lab), as in the reference's
 *     test/test_cImageD11.py;
 *   - the ken_simul fixtures (simAl.gve, test.ubi, test.idx) for
 *     fight_over_peaks / refine_assigned peak counts;
 *   - central finite differences for the joint objective used by the
 *     adjoint (gs_oracle_joint_loss has no reference counterpart; it is
 *     pinned by brute-force recomputation through the other oracle
 *     functions and by the normal equations of score_and_refine);
 *   - indexer functions (closest, quickorient, find, orient): closest and
 *     quickorient bit-for-bit against the reference C; the two 3-term dot
 *     orders (explicit fma calls, the only FMAs in this file) bit-for-bit
 *     against numpy's np.dot; find against the reference's own loop run
 *     with numpy and the reference C closest; orient against the numpy
 *     unitcell.orient; the whole search against the reference's own
 *     test (test/ken_simul/testken.py: 20 grains from simAl.gve) and the
 *     test.idx peak counts.
 *
 * Row-major flat arrays: ubi[9] is ubi[i][j] at 3*i+j, gv[3*k+c].
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1.5 * 2^52: (x + M) - M is round-to-nearest-even for |x| < 2^51.
 * Follows src/closest.c:84-85 (conv_double_to_int_fast). */
#define ORC_MAGIC 6755399441055744.0

static inline double orc_round(double x) {
    /* two literal IEEE adds: the build never reassociates (no -ffast-math) */
    return (x + ORC_MAGIC) - ORC_MAGIC;
}

/* src/closest.c:78-79 (conv_double_to_int_safe) */
static double orc_round_safe(double x) { return floor(x + 0.5); }

/* src/closest.c:100-119 */
int orc_verify_rounding(int n) {
    int bad = 0;
    for (int pass = 0; pass < 2; pass++) {
        int base = pass == 0 ? n : -n;
        for (int i = -100; i < 200; i++) {
            double v = base + i * 50.0;
            int a = (int)orc_round(v);
            int b = (int)orc_round_safe(v);
            bad += (a != b);
        }
    }
    return bad;
}

/* src/closest.c:37-76 (inverse3x3): in place, returns 0 or -1 if det == 0 */
int orc_inverse3x3(double A[9]) {
#define E(i, j) A[3 * (i) + (j)]
    double det = E(0, 0) * (E(2, 2) * E(1, 1) - E(2, 1) * E(1, 2)) -
                 E(1, 0) * (E(2, 2) * E(0, 1) - E(2, 1) * E(0, 2)) +
                 E(2, 0) * (E(1, 2) * E(0, 1) - E(1, 1) * E(0, 2));
    if (!(det != 0.))
        return -1;
    double B[9];
    B[0] = (E(2, 2) * E(1, 1) - E(2, 1) * E(1, 2)) / det;
    B[1] = -(E(2, 2) * E(0, 1) - E(2, 1) * E(0, 2)) / det;
    B[2] = (E(1, 2) * E(0, 1) - E(1, 1) * E(0, 2)) / det;
    B[3] = -(E(2, 2) * E(1, 0) - E(2, 0) * E(1, 2)) / det;
    B[4] = (E(2, 2) * E(0, 0) - E(2, 0) * E(0, 2)) / det;
    B[5] = -(E(1, 2) * E(0, 0) - E(1, 0) * E(0, 2)) / det;
    B[6] = (E(2, 1) * E(1, 0) - E(2, 0) * E(1, 1)) / det;
    B[7] = -(E(2, 1) * E(0, 0) - E(2, 0) * E(0, 1)) / det;
    B[8] = (E(1, 1) * E(0, 0) - E(1, 0) * E(0, 1)) / det;
#undef E
    memcpy(A, B, sizeof(B));
    return 0;
}

/* h_i = (u_i0*g0 + u_i1*g1) + u_i2*g2, the expression of src/closest.c:236 */
static inline void orc_hkl(const double *u, const double *g, double h[3]) {
    for (int i = 0; i < 3; i++)
        h[i] = u[3 * i + 0] * g[0] + u[3 * i + 1] * g[1] + u[3 * i + 2] * g[2];
}

/* sum of squared distances to the nearest integers, src/closest.c:237-242 */
static inline double orc_drlv2(const double h[3], double ih[3]) {
    double t[3];
    for (int i = 0; i < 3; i++) {
        ih[i] = orc_round(h[i]);
        t[i] = h[i] - ih[i];
    }
    return t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
}

/* src/closest.c:227-248 */
int orc_score(const double *ubi, const double *gv, double tol, int ng) {
    double atol = tol * tol, h[3], ih[3];
    int n = 0;
    for (int k = 0; k < ng; k++) {
        orc_hkl(ubi, gv + 3 * k, h);
        if (orc_drlv2(h, ih) < atol)
            n++;
    }
    return n;
}

/* Least-squares solve shared by score_and_refine and refine_assigned:
 * UB = R.H^-1, UBI = UB^-1, written back only if both inverses exist.
 * src/closest.c:320-336 and :458-473. */
static void orc_solve(double *ubi, const double R[9], double H[9]) {
    double UB[9] = {0};
    int k = orc_inverse3x3(H);
    if (k == 0) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int l = 0; l < 3; l++)
                    UB[3 * i + j] = UB[3 * i + j] + R[3 * i + l] * H[3 * l + j];
    }
    if (k == 0 && orc_inverse3x3(UB) == 0)
        memcpy(ubi, UB, sizeof(UB));
}

/* accumulate one kept peak into R and H, src/closest.c:310-316 */
static inline void orc_accumulate(double R[9], double H[9], const double ih[3],
                                  const double *g) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            R[3 * i + j] = R[3 * i + j] + ih[j] * g[i];
            H[3 * i + j] = H[3 * i + j] + ih[j] * ih[i];
        }
}

/* src/closest.c:267-345 */
void orc_score_and_refine(double *ubi, const double *gv, double tol, int *n_out,
                          double *mean_out, int ng) {
    double R[9] = {0}, H[9] = {0}, h[3], ih[3];
    double tolsq = tol * tol, sumdrlv2 = 0.;
    int n = 0;
    for (int k = 0; k < ng; k++) {
        orc_hkl(ubi, gv + 3 * k, h);
        double s = orc_drlv2(h, ih);
        if (s < tolsq) {
            n++;
            sumdrlv2 += s;
            orc_accumulate(R, H, ih, gv + 3 * k);
        }
    }
    orc_solve(ubi, R, H);
    if (n > 0)
        sumdrlv2 /= n;
    *n_out = n;
    *mean_out = sumdrlv2;
}

/* src/closest.c:366-397 */
int orc_score_and_assign(const double *ubi, const double *gv, double tol,
                         double *drlv2, int *labels, int label, int ng) {
    double tolsq = tol * tol;
    int n = 0;
    for (int k = 0; k < ng; k++) {
        double h[3], ih[3];
        orc_hkl(ubi, gv + 3 * k, h);
        double s = orc_drlv2(h, ih);
        if (s < tolsq && s < drlv2[k]) {
            labels[k] = label;
            drlv2[k] = s;
            n++;
        } else if (labels[k] == label) {
            labels[k] = -1;
        }
    }
    return n;
}

/* src/closest.c:416-474 */
void orc_refine_assigned(double *ubi, const double *gv, const int *labels,
                         int label, int *npk, double *mean_out, int ng) {
    double R[9] = {0}, H[9] = {0}, h[3], ih[3], tot = 0.;
    int n = 0;
    for (int k = 0; k < ng; k++) {
        if (labels[k] != label)
            continue;
        n++;
        orc_hkl(ubi, gv + 3 * k, h);
        tot += orc_drlv2(h, ih);
        orc_accumulate(R, H, ih, gv + 3 * k);
    }
    *npk = n;
    *mean_out = n > 0 ? tot / n : 0.;
    orc_solve(ubi, R, H);
}

/* ---------------------------------------------------------------------
 * Diffraction geometry, src/cdiffraction.c.  PI/180.0 is written without
 * parentheses in the reference macro, so sin(wedge*RAD) is
 * sin((wedge*PI)/180.0) while sin(RAD*omega*sign) is
 * sin(((PI/180.0)*omega)*sign).  src/cdiffraction.c:7-9, :119-147.
 */
#define ORC_PI 3.14159265358979323846

/* Each sin(x), cos(x) pair of the reference (src/cdiffraction.c:119-120,
 * :130-131, :146-147) is compiled by gcc -O2 into ONE glibc sincos(x) call
 * (visible in the disassembly of the reference built by oracle/Makefile).
 * glibc's sincos and its separate sin/cos disagree in the last bit for
 * about 0.1% of arguments, so the oracle calls sincos explicitly instead of
 * depending on that optimisation. */
static inline void orc_sincos(double x, double *s, double *c) { sincos(x, s, c); }

/* the combined wedge/chi matrix, src/cdiffraction.c:119-142 with the
 * matmat macro of src/cdiffraction.h:20-29 (c = b.a in row-major terms) */
void orc_wedge_chi_matrix(double wedge, double chi, double mat[9]) {
    double sw, cw, sc, cc;
    orc_sincos(wedge * ORC_PI / 180.0, &sw, &cw);
    orc_sincos(chi * ORC_PI / 180.0, &sc, &cc);
    double w[9] = {cw, 0.0, -sw, 0., 1.0, 0., sw, 0.0, cw};
    double c[9] = {1., 0.0, 0., 0., cc, -sc, 0., sc, cc};
    for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++)
            mat[3 * col + row] = c[row] * w[3 * col + 0] +
                                 c[row + 3] * w[3 * col + 1] +
                                 c[row + 6] * w[3 * col + 2];
}

/* sin/cos of the omega angle exactly as the reference forms the argument */
void orc_sincos_omega(const double *omega, double omegasign, double *so,
                      double *co, int n) {
    for (int i = 0; i < n; i++)
        orc_sincos(ORC_PI / 180.0 * omega[i] * omegasign, so + i, co + i);
}

/* one peak of compute_gv given its sin/cos, src/cdiffraction.c:149-166 */
static inline void orc_gv_one(const double *xl, double so, double co,
                              const double mat[9], double ds, const double t[3],
                              double *g) {
    double u[3], o[3], d[3], k[3], v[3];
    u[0] = co * t[0] - so * t[1];
    u[1] = so * t[0] + co * t[1];
    u[2] = t[2];
    o[0] = mat[0] * u[0] + mat[1] * u[1] + mat[2] * u[2];
    o[1] = mat[3] * u[0] + mat[4] * u[1] + mat[5] * u[2];
    o[2] = mat[6] * u[0] + mat[7] * u[1] + mat[8] * u[2];
    d[0] = xl[0] - o[0];
    d[1] = xl[1] - o[1];
    d[2] = xl[2] - o[2];
    double modyz = 1. / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    k[0] = ds * (d[0] * modyz - 1.);
    k[1] = ds * d[1] * modyz;
    k[2] = ds * d[2] * modyz;
    v[0] = mat[0] * k[0] + mat[3] * k[1] + mat[6] * k[2];
    v[1] = mat[1] * k[0] + mat[4] * k[1] + mat[7] * k[2];
    v[2] = mat[2] * k[0] + mat[5] * k[1] + mat[8] * k[2];
    g[0] = co * v[0] + so * v[1];
    g[1] = -so * v[0] + co * v[1];
    g[2] = v[2];
}

/* src/cdiffraction.c:112-168 */
void orc_compute_gv(const double *xlylzl, const double *omega, double omegasign,
                    double wvln, double wedge, double chi, const double t[3],
                    double *gv, int n) {
    double mat[9];
    orc_wedge_chi_matrix(wedge, chi, mat);
    for (int i = 0; i < n; i++) {
        double so, co;
        orc_sincos(ORC_PI / 180.0 * omega[i] * omegasign, &so, &co);
        double ds = 1. / wvln;
        orc_gv_one(xlylzl + 3 * i, so, co, mat, ds, t, gv + 3 * i);
    }
}

/* compute_gv with the per-peak sin/cos and the wedge/chi matrix supplied
 * by the caller: the same arithmetic as orc_compute_gv after the libm
 * calls.  Used to pin the device arithmetic bit-for-bit independently of
 * how sin/cos were evaluated. */
void orc_compute_gv_trig(const double *xlylzl, const double *so,
                         const double *co, const double mat[9], double wvln,
                         const double t[3], double *gv, int n) {
    double ds = 1. / wvln;
    for (int i = 0; i < n; i++)
        orc_gv_one(xlylzl + 3 * i, so[i], co[i], mat, ds, t, gv + 3 * i);
}

/* src/cdiffraction.c:190-224 */
void orc_compute_xlylzl(const double *s, const double *f, const double p[4],
                        const double r[9], const double dist[3], double *xlylzl,
                        int n) {
    for (int i = 0; i < n; i++) {
        double v1 = (f[i] - p[1]) * p[3];
        double v2 = (s[i] - p[0]) * p[2];
        for (int j = 0; j < 3; j++)
            xlylzl[3 * i + j] = r[3 * j + 1] * v1 + r[3 * j + 2] * v2 + dist[j];
    }
}

/* src/cdiffraction.c:245-277 */
void orc_compute_xlylzl_xpos_variable(const double *s, const double *f, const double p[4],
                                      const double r[9], const double dist[3],
                                      const double *xpos, double *xlylzl, int n) {
    for (int i = 0; i < n; i++) {
        double v1 = (f[i] - p[1]) * p[3];
        double v2 = (s[i] - p[0]) * p[2];
        xlylzl[3 * i + 0] = r[1] * v1 + r[2] * v2 + dist[0] - xpos[i];
        xlylzl[3 * i + 1] = r[4] * v1 + r[5] * v2 + dist[1];
        xlylzl[3 * i + 2] = r[7] * v1 + r[8] * v2 + dist[2];
    }
}

/* one peak of compute_geometry after the libm sin/cos
 * (src/cdiffraction.c:62-92); the DEG macro expands to 180.0 / PI. */
static void orc_geometry_one(const double *xl, double so, double co, const double mat[9],
                             double wvln, const double t[3], double *out) {
    double u[3] = {co * t[0] - so * t[1], so * t[0] + co * t[1], t[2]}, o[3], d[3], k[3], v[3];
    for (int r = 0; r < 3; r++) o[r] = mat[3 * r] * u[0] + mat[3 * r + 1] * u[1] + mat[3 * r + 2] * u[2];
    for (int r = 0; r < 3; r++) d[r] = xl[r] - o[r];
    double modyz = 1. / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    out[0] = 180.0 / ORC_PI * atan2(sqrt(d[1] * d[1] + d[2] * d[2]), d[0]);
    double ds = 1. / wvln;
    k[0] = ds * (d[0] * modyz - 1.);
    k[1] = ds * d[1] * modyz;
    k[2] = ds * d[2] * modyz;
    out[1] = 180.0 / ORC_PI * atan2(-d[1], d[2]);
    out[2] = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    for (int c = 0; c < 3; c++) v[c] = mat[c] * k[0] + mat[3 + c] * k[1] + mat[6 + c] * k[2];
    out[3] = co * v[0] + so * v[1];
    out[4] = -so * v[0] + co * v[1];
    out[5] = v[2];
}

/* src/cdiffraction.c:31-94 (compute_geometry); out is (n, 6) */
void orc_compute_geometry(const double *xlylzl, const double *omega, double omegasign,
                          double wvln, double wedge, double chi, const double t[3], double *out,
                          int n) {
    double mat[9];
    orc_wedge_chi_matrix(wedge, chi, mat);
    for (int i = 0; i < n; i++) {
        double so, co;
        orc_sincos(ORC_PI / 180.0 * omega[i] * omegasign, &so, &co);
        orc_geometry_one(xlylzl + 3 * i, so, co, mat, wvln, t, out + 6 * i);
    }
}

/* compute_geometry with the per-peak sin/cos and wedge/chi matrix given */
void orc_compute_geometry_trig(const double *xlylzl, const double *so, const double *co,
                               const double mat[9], double wvln, const double t[3], double *out,
                               int n) {
    for (int i = 0; i < n; i++) orc_geometry_one(xlylzl + 3 * i, so[i], co[i], mat, wvln, t, out + 6 * i);
}

/* ---------------------------------------------------------------------
 * Loop drivers: the Python callers' per-grain loops, in C so that the
 * oracle can check full-size batched results in seconds.
 */

/* indexer.score over many UBIs (ImageD11/indexing.py:1090-1097), one
 * score() per grain; grains are independent so they run in parallel. */
void orc_score_many(const double *ubis, int K, const double *gv, int ng,
                    double tol, int *counts, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int g = 0; g < K; g++)
        counts[g] = orc_score(ubis + 9 * g, gv, tol, ng);
}

/* refinegrains.assignlabels / indexer.fight_over_peaks loop
 * (ImageD11/refinegrains.py:718-734, ImageD11/indexing.py:885-889):
 * score_and_assign for each grain in order, sharing drlv2/labels.  The
 * per-peak state machine only depends on that peak, so peaks may be
 * split across threads without changing any result. */
void orc_assign_many(const double *ubis, const int *grain_labels, int K,
                     const double *gv, int ng, double tol, double *drlv2,
                     int *labels, int *counts, int nthreads) {
    double tolsq = tol * tol;
    for (int g = 0; g < K; g++)
        counts[g] = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        int *local = calloc((size_t)(K > 0 ? K : 1), sizeof(int));
#ifdef _OPENMP
#pragma omp for schedule(static, 4096)
#endif
        for (int k = 0; k < ng; k++) {
            double d = drlv2[k];
            int lab = labels[k];
            for (int g = 0; g < K; g++) {
                double h[3], ih[3];
                orc_hkl(ubis + 9 * g, gv + 3 * k, h);
                double s = orc_drlv2(h, ih);
                if (s < tolsq && s < d) {
                    lab = grain_labels[g];
                    d = s;
                    local[g]++;
                } else if (lab == grain_labels[g]) {
                    lab = -1;
                }
            }
            drlv2[k] = d;
            labels[k] = lab;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int g = 0; g < K; g++)
            counts[g] += local[g];
        free(local);
    }
}

/* refinegrains.assignlabels with a per-grain translation: compute_gv for
 * the grain's origin, then score_and_assign (ImageD11/refinegrains.py:718-734).
 * so/co/mat are the per-peak trig values and wedge/chi matrix. */
void orc_gv_assign_many(const double *xlylzl, const double *so, const double *co,
                        const double mat[9], double wvln, const double *ubis,
                        const double *trans, const int *grain_labels, int K,
                        int ng, double tol, double *drlv2, int *labels,
                        int *counts, int nthreads) {
    double tolsq = tol * tol, ds = 1. / wvln;
    for (int g = 0; g < K; g++)
        counts[g] = 0;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        int *local = calloc((size_t)(K > 0 ? K : 1), sizeof(int));
#ifdef _OPENMP
#pragma omp for schedule(static, 1024)
#endif
        for (int k = 0; k < ng; k++) {
            double d = drlv2[k];
            int lab = labels[k];
            for (int g = 0; g < K; g++) {
                double gvec[3], h[3], ih[3];
                orc_gv_one(xlylzl + 3 * k, so[k], co[k], mat, ds, trans + 3 * g,
                           gvec);
                orc_hkl(ubis + 9 * g, gvec, h);
                double s = orc_drlv2(h, ih);
                if (s < tolsq && s < d) {
                    lab = grain_labels[g];
                    d = s;
                    local[g]++;
                } else if (lab == grain_labels[g]) {
                    lab = -1;
                }
            }
            drlv2[k] = d;
            labels[k] = lab;
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int g = 0; g < K; g++)
            counts[g] += local[g];
        free(local);
    }
}

/* Batched score_and_refine over per-grain peak lists (CSR: the peaks of
 * grain g are gv[idx[off[g]] .. idx[off[g+1]-1]]), n_iter calls per grain
 * as refinegrains.refine does (ImageD11/refinegrains.py:459-498, two
 * calls).  n_out/mean_out are [n_iter][K]. */
void orc_refine_lists(double *ubis, int K, const double *gv, const int *idx,
                      const long long *off, double tol, int n_iter, int *n_out,
                      double *mean_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *buf = NULL;
        long long cap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int g = 0; g < K; g++) {
            long long m = off[g + 1] - off[g];
            if (m > cap) {
                free(buf);
                cap = m;
                buf = malloc(sizeof(double) * 3 * (size_t)(cap > 0 ? cap : 1));
            }
            for (long long i = 0; i < m; i++)
                memcpy(buf + 3 * i, gv + 3 * (size_t)idx[off[g] + i],
                       3 * sizeof(double));
            for (int it = 0; it < n_iter; it++)
                orc_score_and_refine(ubis + 9 * g, buf, tol,
                                     n_out + (size_t)it * K + g,
                                     mean_out + (size_t)it * K + g, (int)m);
        }
        free(buf);
    }
}

/* ---------------------------------------------------------------------
 * Joint objective for the adjoint (no reference counterpart): for grain g
 * with translation t_g and fixed peak list / fixed integer hkl,
 *   L = sum_g sum_{k in A_g} || UBI_g . gv_k(geometry, t_g) - hkl_k ||^2
 * with gv_k = compute_gv(compute_xlylzl(s_k, f_k)).  Plain double, used by
 * finite-difference tests of the device gradient.  mat is the wedge/chi
 * matrix so that wedge and chi derivatives can be taken through it.
 */
double orc_joint_loss(const double *sc, const double *fc, const double *so,
                      const double *co, const double p[4], const double r[9],
                      const double dist[3], double wvln, const double mat[9],
                      const double *ubis, const double *trans, int K,
                      const int *idx, const long long *off, const int *hkl) {
    double loss = 0.;
    double ds = 1. / wvln;
    for (int g = 0; g < K; g++) {
        double part = 0.;
        for (long long q = off[g]; q < off[g + 1]; q++) {
            int k = idx[q];
            double xl[3], gvec[3], h[3];
            orc_compute_xlylzl(sc + k, fc + k, p, r, dist, xl, 1);
            orc_gv_one(xl, so[k], co[k], mat, ds, trans + 3 * g, gvec);
            orc_hkl(ubis + 9 * g, gvec, h);
            for (int i = 0; i < 3; i++) {
                double e = h[i] - (double)hkl[3 * q + i];
                part += e * e;
            }
        }
        loss += part;
    }
    return loss;
}

/* ---- indexer ----------------------------------------------------------- */

/* src/closest.c:188-207: scan x-major, strict '<', start at 99 / index 0. */
void orc_closest(const double *x, const double *v, int *ibest, double *best, int nx, int nv) {
    double b = 99.0;
    int ib = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nv; j++)
            if (fabs(x[i] - v[j]) < b) {
                b = fabs(x[i] - v[j]);
                ib = i;
            }
    *ibest = ib;
    *best = b;
}

/* src/cdiffraction.c:293-328 (Busing & Levy from two observed g-vectors). */
void orc_quickorient(double *u, const double *bt) {
    double m[9];
    m[6] = u[1] * u[5] - u[2] * u[4];
    m[7] = u[2] * u[3] - u[0] * u[5];
    m[8] = u[0] * u[4] - u[1] * u[3];
    double s = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int c = 0; c < 3; c++) m[c] = u[c] / s;
    s = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
    for (int c = 6; c < 9; c++) m[c] /= s;
    m[3] = m[1] * m[8] - m[2] * m[7];
    m[4] = m[2] * m[6] - m[0] * m[8];
    m[5] = m[0] * m[7] - m[1] * m[6];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            u[3 * r + c] = bt[3 * r + 0] * m[c] + bt[3 * r + 1] * m[3 + c] + bt[3 * r + 2] * m[6 + c];
}

/* unit rows of gv[idx] as indexer.find builds n1, n2
 * (ImageD11/indexing.py:696-706): |g| = sqrt((x*x + y*y) + z*z), then
 * each component divided by it. */
void orc_unit_rows(const double *gv, const int *idx, int n, double *out) {
    for (int r = 0; r < n; r++) {
        const double *g = gv + 3 * (long)idx[r];
        double m = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        for (int c = 0; c < 3; c++) out[3 * r + c] = g[c] / m;
    }
}

/* The indexer's 3-term dot products as numpy 2.2 / OpenBLAS 0.3.29
 * (x86-64 Haswell/SkylakeX kernels) evaluates them: dgemv (np.dot of a
 * matrix row and a vector) fma(x2,y2, fma(x0,y0, x1*y1)); ddot (np.dot of
 * two vectors) fma(x2,y2, fma(x1,y1, x0*y0)).  Pinned against numpy in
 * tests/test_indexer_cpu.py. */
static double orc_dot3_gemv(const double *x, const double *y) {
    return fma(x[2], y[2], fma(x[0], y[0], x[1] * y[1]));
}
static double orc_dot3_ddot(const double *x, const double *y) {
    return fma(x[2], y[2], fma(x[1], y[1], x[0] * y[0]));
}

/* out[r] = x[r] . y[r] in the dgemv (mode 0) or ddot (mode 1) order, for
 * pinning the two orders against numpy. */
void orc_dot3_rows(const double *x, const double *y, int n, int mode, double *out) {
    for (int r = 0; r < n; r++)
        out[r] = mode == 0 ? orc_dot3_gemv(x + 3 * (long)r, y + 3 * (long)r)
                           : orc_dot3_ddot(x + 3 * (long)r, y + 3 * (long)r);
}

/* indexer.find with cosine_tol > 0 (ImageD11/indexing.py:710-716): for
 * every row i of n1, costheta[j] = np.dot(n2, n1[i])[j] and
 * closest(costheta, cs). */
void orc_find_hits(const double *n1, int N1, const double *n2, int N2, const double *cs, int nv,
                   int *ibest, double *best) {
    double *ct = (double *)malloc(sizeof(double) * (N2 > 0 ? N2 : 1));
    for (int i = 0; i < N1; i++) {
        const double *a = n1 + 3 * (long)i;
        for (int j = 0; j < N2; j++) {
            const double *b = n2 + 3 * (long)j;
            ct[j] = orc_dot3_gemv(b, a);
        }
        orc_closest(ct, cs, ibest + i, best + i, N2, nv);
    }
    free(ct);
}

/* ImageD11/indexing.py:717-723 (cosine_tol < 0): keys (i*nv + m)*N2 + j of
 * every |cs[m] - costheta_ij| < mtol in i, m, j order; returns the count
 * (keys written while < cap). */
long long orc_find_all(const double *n1, int N1, const double *n2, int N2, const double *cs,
                       int nv, double mtol, long long *keys, long long cap) {
    long long n = 0;
    double *ct = (double *)malloc(sizeof(double) * (N2 > 0 ? N2 : 1));
    for (int i = 0; i < N1; i++) {
        const double *a = n1 + 3 * (long)i;
        for (int j = 0; j < N2; j++) {
            const double *b = n2 + 3 * (long)j;
            ct[j] = orc_dot3_gemv(b, a);
        }
        for (int m = 0; m < nv; m++)
            for (int j = 0; j < N2; j++)
                if (fabs(cs[m] - ct[j]) < mtol) {
                    if (n < cap) keys[n] = ((long long)i * nv + m) * N2 + j;
                    n++;
                }
    }
    free(ct);
    return n;
}

/* unitcell.orient, crange <= 0 (ImageD11/unitcell.py:617-630, 649-653),
 * for the peak pairs (hi[h], hj[h]). */
void orc_orient_hits(const double *gv, const int *hi, const int *hj, int nh, const double *c2ab,
                     int nc, const double *bt, double *ubis, int *choice) {
    for (int h = 0; h < nh; h++) {
        const double *g1 = gv + 3 * (long)hi[h], *g2 = gv + 3 * (long)hj[h];
        double dot = orc_dot3_ddot(g1, g2);
        double s1 = g1[0] * g1[0] + g1[1] * g1[1] + g1[2] * g1[2];
        double s2 = g2[0] * g2[0] + g2[1] * g2[1] + g2[2] * g2[2];
        double ct = dot / sqrt(s1 * s2);
        int i = 0; /* searchsorted(c2ab, ct, side='left'); NaN sorts last */
        if (isnan(ct))
            i = nc;
        else
            while (i < nc && c2ab[i] < ct) i++;
        int b = i;
        if (i > 0 && (i == nc || fabs(ct - c2ab[i - 1]) < fabs(ct - c2ab[i]))) b = i - 1;
        double *u = ubis + 9 * (long)h;
        for (int c = 0; c < 3; c++) {
            u[c] = g1[c];
            u[3 + c] = g2[c];
            u[6 + c] = 0.0;
        }
        orc_quickorient(u, bt + 9 * (long)b);
        if (choice) choice[h] = b;
    }
}

int orc_openmp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 0;
#endif
}
