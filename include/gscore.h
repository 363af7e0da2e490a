/*
 * gscore.h — C ABI of libgscore, the MI355X (gfx950) grain-scoring path.
 *
 * Two families of entry points:
 *
 *  1. The reference's own prototypes, on host pointers, so that existing
 *     bindings (the f2py signatures of src/_cImageD11.pyf:106-170,551-583,
 *     or a ctypes/cffi stub) can load this library in place of the
 *     ImageD11._cImageD11 extension.  Arrays are row-major float64 / int32
 *     exactly as the reference expects; inputs are staged to the GPU, the
 *     kernels run, outputs are copied back before the call returns.  The
 *     reference has no error codes; on a HIP error these record a message
 *     (gs_last_error) and return -1 (int results), set *n = -1 (refine), or
 *     leave the output untouched (geometry).
 *
 *  2. gs_* entry points on device pointers, enqueued on a caller's stream
 *     (a hipStream_t passed as void*, NULL = the null stream), returning a
 *     gs_status.  Small parameter arrays (p, r, dist, t, mat) are host
 *     pointers read at launch time.
 *
 * All sizes are C int as in the reference (n < 2^31).
 */
#ifndef GSCORE_H
#define GSCORE_H

#ifdef __cplusplus
extern "C" {
#endif

#define GS_VERSION 1

typedef void *gs_stream_t;

enum gs_status {
    GS_OK = 0,
    GS_ERR_ARG = 1,      /* invalid argument (sizes, null pointers)     */
    GS_ERR_HIP = 2,      /* a HIP runtime call or kernel launch failed  */
    GS_ERR_NODEVICE = 3  /* no HIP device is visible                    */
};

/* ---- library state -------------------------------------------------- */
const char *gs_last_error(void);  /* message of the last failure on this thread */
int gs_clear_error(void);
int gs_version(void);
int gs_device_count(void);
int gs_init(void);               /* allocate the per-device state now */

/* The device sin/cos algorithm evaluated on the host CPU, for checking its
 * accuracy without a GPU; no other entry point calls it. */
void gs_selftest_sincos_host(const double *x, double *s, double *c, int n);
/* The same algorithm's series path alone (the path its table path falls
 * back to), on the host; *n_table = how many arguments the table path
 * settles.  Checking aid only. */
void gs_selftest_sincos_series_host(const double *x, double *s, double *c, int n, int *n_table);

/* The certified score decision evaluated on the host next to the
 * reference arithmetic, for checking the error bound without a GPU:
 * out[k] = 1 certain true, 0 certain false, 2 undecided; +4 when the
 * reference arithmetic gives true.  No other entry point calls it. */
void gs_selftest_certified_host(const double *ubi, const double *gv, int n, double tol,
                                double gmax, int *out);
/* the same for the score kernel's fract form (drlv2 - 3/4 against
 * tol^2 - 3/4 +- its window). */
void gs_selftest_certified_fract_host(const double *ubi, const double *gv, int n, double tol,
                                      double gmax, int *out);

/* The fast-path decision of gs_gv_assign_many evaluated on the host for
 * single pairs (row k: peak xl/so/co, UBI ubis[9k..], translation
 * trans[3k..]), next to the reference arithmetic: out[k] = 1 certainly
 * fails the tolerance, 2 left to the exact route; +4 when the reference
 * arithmetic passes it.  No other entry point calls it. */
void gs_selftest_scan_host(const double *xl, const double *so, const double *co,
                           const double *ubis, const double *trans, int n, const double *mat,
                           double wvln, double tol, int *out);

/* gs_gauss_newton_ubis's per-grain update evaluated on the host, for
 * checking it without a GPU.  No other entry point calls it. */
void gs_selftest_gn_ubis_host(const double *nsums, const double *ubis, int K, double *out);

/* ---- 1. reference prototypes (host pointers) ------------------------ */

/* src/closest.c:100-119 — MAGIC rounding vs floor(x+0.5) on 600 points,
 * evaluated on the GPU; returns the number of mismatches (0). */
int verify_rounding(int n);

/* src/closest.c:227-248 — peaks with sum((h - round h)^2) < tol^2,
 * h = ubi . gv[k]. */
int score(double ubi[3][3], double gv[][3], double tol, int ng);

/* src/closest.c:267-345 — score, then UB = R.H^-1 over the kept peaks,
 * ubi := UB^-1 unless singular; *n = kept peaks, *sumdrlv2 = their mean
 * drlv2 under the input ubi. */
void score_and_refine(double ubi[3][3], double gv[][3], double tol, int *n_arg,
                      double *sumdrlv2_arg, int ng);

/* src/closest.c:366-397 — peaks fitting better than drlv2[k] take
 * `label`; a peak that had `label` and no longer fits gets -1. */
int score_and_assign(double ubi[3][3], double gv[][3], double tol, double *drlv2,
                     int *labels, int label, int ng);

/* src/closest.c:416-474 — refine on the peaks with labels[k] == label. */
void refine_assigned(double ubi[3][3], double gv[][3], int *labels, int label, int *npk,
                     double *sumdrlv2, int ng);

/* src/cdiffraction.c:112-168 — g-vectors from lab positions and omega.
 * sin/cos of omega, wedge and chi come from the host's glibc sincos() with
 * the reference's expressions (gcc -O2 turns each sin/cos pair of the
 * reference into that one call), the rest runs on the GPU: the output is
 * the reference's, bit for bit. */
void compute_gv(double xlylzl[][3], double omega[], double omegasign, double wvln,
                double wedge, double chi, double t[3], double gv[][3], int n);

/* src/cdiffraction.c:190-224 — lab positions from detector (s, f). */
void compute_xlylzl(double s[], double f[], double p[4], double r[9], double dist[3],
                    double xlylzl[][3], int n);

/* ---- 2. device-pointer entry points --------------------------------- */

/* score modes: all give the reference's counts for every input.
 * CERTIFIED evaluates h with FMA and decides each pair from a proven error
 * bound, recomputing the few undecidable pairs with the reference's own
 * arithmetic; EXACT uses the reference's arithmetic for every pair. */
#define GS_SCORE_CERTIFIED 0
#define GS_SCORE_EXACT 1
/* gs_score_many only: row 0 of every pair first; the pairs it cannot
 * exclude (|h0 - round h0| near or below tol) take the CERTIFIED test. */
#define GS_SCORE_SIEVE 2

/* score for K UBIs (K,3,3) against ng g-vectors -> counts[K] (int32);
 * the loop indexer.score runs over trial UBIs (ImageD11/indexing.py:1090).
 * peaks_per_lane: 1, 2, 4, 8 or 16 (tile height; 0 = default: 4 certified,
 * 8 exact; the sieve takes at most 8). */
int gs_score_many(const double *ubis, int K, const double *gv, int ng, double tol,
                  int *counts, int peaks_per_lane, int mode, gs_stream_t stream);

/* score_and_assign for K grains in order, label grain_labels[g] for grain
 * g, on shared drlv2/labels (ImageD11/refinegrains.py:718-734,
 * ImageD11/indexing.py:885-889); counts[g] = value of the g-th call.
 * mode GS_SCORE_CERTIFIED or GS_SCORE_EXACT, same results. */
int gs_assign_many(const double *ubis, const int *grain_labels, int K, const double *gv,
                   int ng, double tol, double *drlv2, int *labels, int *counts, int mode,
                   gs_stream_t stream);

/* one-grain score_and_refine / refine_assigned on device memory; ubi (9
 * doubles) is updated in place, n_out/mean_out are device scalars. */
int gs_score_and_refine_dev(double *ubi, const double *gv, int ng, double tol, int *n_out,
                            double *mean_out, gs_stream_t stream);
int gs_refine_assigned_dev(double *ubi, const double *gv, int ng, const int *labels,
                           int label, int *n_out, double *mean_out, gs_stream_t stream);
/* gs_score_and_refine_dev for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It calls no hipMalloc,
 * hipFree or allocating routine: the block partials (at most 1024 blocks x
 * 17 doubles; every ng >= 0 fits, since the grid is capped at 1024
 * blocks that stride over the peaks) live in a fixed 136 KiB device array
 * of the library, one per GPU, held until the process ends and looked up
 * by gs_init() or the first call on a device.  Same results, bit for bit,
 * as gs_score_and_refine_dev.  Calls that overlap on one GPU share that
 * array and race: serialise them (one stream, or events).  Rejects a
 * negative ng or a null pointer with GS_ERR_ARG. */
int gs_score_and_refine_dev_noalloc(double *ubi, const double *gv, int ng, double tol,
                                    int *n_out, double *mean_out, gs_stream_t stream);
/* The same two calls with ubi (9 doubles, updated in place), n_out and
 * mean_out on the HOST and gv (and labels) on the device: the calling
 * pattern of a per-grain loop over device-resident peaks.  Waits for the
 * stream before returning. */
int gs_score_and_refine_host_ubi(double *ubi, const double *gv, int ng, double tol, int *n_out,
                                 double *mean_out, gs_stream_t stream);
int gs_refine_assigned_host_ubi(double *ubi, const double *gv, int ng, const int *labels,
                                int label, int *n_out, double *mean_out, gs_stream_t stream);

/* compute_xlylzl / compute_gv on device memory. */
int gs_compute_xlylzl_dev(const double *s, const double *f, const double p[4],
                          const double r[9], const double dist[3], double *xlylzl, int n,
                          gs_stream_t stream);
int gs_compute_gv_dev(const double *xlylzl, const double *omega, double omegasign,
                      double wvln, double wedge, double chi, const double t[3], double *gv,
                      int n, gs_stream_t stream);

/* compute_gv with per-peak sin/cos of omega and the wedge/chi matrix given
 * (the arithmetic of src/cdiffraction.c:149-166 after the libm calls). */
int gs_compute_gv_trig_dev(const double *xlylzl, const double *so, const double *co,
                           const double mat[9], double wvln, const double t[3], double *gv,
                           int n, gs_stream_t stream);

/* sin and cos of ((PI/180)*omega)*omegasign per peak, on the device (the
 * fast mode: the device's correctly rounded sincos, which differs from
 * glibc's in the last bit for a few tenths of a percent of arguments). */
int gs_sincos_omega(const double *omega, double omegasign, double *so, double *co, int n,
                    gs_stream_t stream);

/* The parity mode's trig: glibc sincos() on the host, with the reference's
 * expressions (src/cdiffraction.c:146-147 for omega, :119-120 and :130-131
 * for wedge and chi; gcc -O2 compiles each sin/cos pair there into one
 * sincos call), so sin/cos, and every g-vector computed from them on the
 * GPU, are the reference's bits.  The peaks are split over the CPUs the
 * process may run on (OMP_NUM_THREADS when set).
 * gs_sincos_omega_host: host arrays.
 * gs_wedge_chi_matrix_host: the 3x3 matrix mat (row-major) of compute_gv
 *   (src/cdiffraction.c:121-142) for wedge, chi in degrees.
 * gs_sincos_omega_parity: DEVICE omega in, DEVICE so/co out, through
 *   pinned host memory on `stream`; returns when so/co are written. */
int gs_sincos_omega_host(const double *omega, double omegasign, double *so, double *co, int n);
int gs_wedge_chi_matrix_host(double wedge, double chi, double *mat);
int gs_sincos_omega_parity(const double *omega, double omegasign, double *so, double *co, int n,
                           gs_stream_t stream);

/* compute_xlylzl -> compute_gv fused (ImageD11/transform.py:729-738,
 * Ctransform.sf2gv); xlylzl_out and so_out/co_out may be NULL. */
int gs_sf2gv(const double *s, const double *f, const double *omega, const double p[4],
             const double r[9], const double dist[3], double omegasign, double wvln,
             double wedge, double chi, const double t[3], double *xlylzl_out, double *so_out,
             double *co_out, double *gv, int n, gs_stream_t stream);

/* gs_sf2gv with the per-peak sin/cos (so, co) and the wedge/chi matrix mat
 * (9 host doubles) given, e.g. by the parity mode above. */
int gs_sf2gv_trig(const double *s, const double *f, const double *so, const double *co,
                  const double p[4], const double r[9], const double dist[3],
                  const double mat[9], double wvln, const double t[3], double *xlylzl_out,
                  double *gv, int n, gs_stream_t stream);

/* Batched score_and_refine: grain g refines on the peaks gv[idx[off[g]]]
 * .. gv[idx[off[g+1]-1]], n_iter successive calls (refinegrains.refine,
 * ImageD11/refinegrains.py:459-498, makes two) inside ONE persistent
 * launch.  ubis (K,3,3) updated in place; n_out/mean_out are [n_iter][K].
 * `order` (K ints) is the order in which the persistent waves take grains
 * (longest list first balances best).  lanes_per_grain 8/16/32/64 and
 * waves_per_cu tune the launch (0 = defaults).  queue_ws: one device int
 * used as the work-queue counter (NULL = a per-device counter of the
 * library, which calls overlapping on one device would share).  Entries
 * out of range are ignored, never dereferenced: grains outside [0, K) in
 * order, offsets outside [0, n_idx], peak indices outside [0, ng).
 * idx may be NULL: list entry q is then row q of gv (rows gathered once by
 * gs_gather_rows, or lists already stored contiguously), ng >= n_idx.
 * Every iteration then reads contiguous rows instead of one cache line
 * per entry; the sums see the same values in the same order, so the
 * results are bit-identical to the indexed call. */
int gs_refine_lists(double *ubis, int K, const double *gv, int ng, const int *idx,
                    long long n_idx, const long long *off, const int *order, double tol,
                    int n_iter, int *n_out, double *mean_out, int lanes_per_grain,
                    int waves_per_cu, int *queue_ws, gs_stream_t stream);

/* The same with lanes in proportion to list length: the grains whose list
 * has at least long_min entries (counted on the device) are refined by a
 * whole wave each, those with at least mid_min by 32 lanes each, the others
 * by lanes_per_grain lanes each.  With `order` longest list first, no list
 * sets the makespan.  queue_ws: 3 device ints (the counter and the two
 * counts; NULL = the library's per-device counters).  long_min = mid_min =
 * 0 is gs_refine_lists; 0 switches a tier off.  A grain's sums are reduced
 * over its lanes, so a grain in another tier can differ from the
 * lanes_per_grain launch by rounding. */
int gs_refine_lists_long(double *ubis, int K, const double *gv, int ng, const int *idx,
                         long long n_idx, const long long *off, const int *order, double tol,
                         int n_iter, int *n_out, double *mean_out, int lanes_per_grain,
                         int waves_per_cu, long long long_min, long long mid_min, int *queue_ws,
                         gs_stream_t stream);

/* rows[q] = gv[idx[q]] (3 doubles) for q < n_idx, a NaN row where idx[q] is
 * outside [0, ng): the list rows of gs_refine_lists packed once, in list
 * order.  A NaN row is never kept, as an out-of-range index is never kept.
 * No counterpart in the reference (score_and_refine takes each grain's
 * rows as its gv argument, src/closest.c:267-345). */
int gs_gather_rows(const double *gv, int ng, const int *idx, long long n_idx, double *rows,
                   gs_stream_t stream);

/* The same refinement as a hipGraph: n_iter single-iteration persistent
 * launches (each a full score_and_refine pass over all grains) captured
 * once on `stream` and replayed with gs_graph_launch.  queue_ws must hold
 * n_iter device ints; all buffers must stay allocated while the graph is
 * used.  Replays start from whatever ubis holds.  idx may be NULL as for
 * gs_refine_lists (then gather the rows again before a replay whenever gv
 * has changed). */
typedef struct gs_graph_s *gs_graph_t;
int gs_refine_graph_create(gs_graph_t *out, double *ubis, int K, const double *gv, int ng,
                           const int *idx, long long n_idx, const long long *off,
                           const int *order, double tol, int n_iter, int *n_out,
                           double *mean_out, int lanes_per_grain, int waves_per_cu,
                           int *queue_ws, gs_stream_t stream);
/* The graph of gs_refine_lists_long: queue_ws holds n_iter + 2 ints. */
int gs_refine_graph_create_long(gs_graph_t *out, double *ubis, int K, const double *gv, int ng,
                                const int *idx, long long n_idx, const long long *off,
                                const int *order, double tol, int n_iter, int *n_out,
                                double *mean_out, int lanes_per_grain, int waves_per_cu,
                                long long long_min, long long mid_min, int *queue_ws,
                                gs_stream_t stream);
int gs_graph_launch(gs_graph_t graph, gs_stream_t stream);
int gs_graph_destroy(gs_graph_t graph);

/* ---- adjoint of the joint refinement objective ----------------------
 * No counterpart in the reference (which fits geometry by Nelder-Mead on
 * refinegrains.gof, ImageD11/refinegrains.py:517-624).  The objective is
 *   L = sum_g sum_{q in list g} || UBI_g . G_q - hkl_q ||^2
 * with G_q = compute_gv(compute_xlylzl(sc[k], fc[k]), omega_k) for grain g's
 * translation, k = idx[q], and fixed integer hkl_q.  so/co are the
 * per-peak sin/cos of omega (gs_sincos_omega).  geo (26 device doubles):
 * p[4] (z_center, y_center, z_size, y_size), r[9], dist[3], wavelength,
 * mat[9] (wedge/chi matrix).
 *
 * gs_joint_hkl: hkl[3q..] = round(UBI_g . G_q) for every list entry, owner[q]
 * = the grain of entry q.
 * gs_joint_schedule (host arrays): the pieces of a pass.  Every grain's
 * list [off[g], off[g+1]) (clipped to [0, n_idx]) is cut into pieces of at
 * most piece_len consecutive entries, numbered in grain order: grain g
 * owns pieces piece_off[g] .. piece_off[g+1]-1 (piece_off: K+1 entries),
 * piece_grain[p] = the grain of piece p, and slots[] = the pieces by
 * length, longest first (ties in piece order).  *n_pieces = the count;
 * with piece_grain and slots both NULL only piece_off and the count are
 * written (a query), otherwise both need room for the count (cap).
 * gs_joint_workspace_doubles(n_pieces): the doubles of workspace that
 * gs_joint_adjoint needs for n_pieces pieces (any lane count).
 * gs_joint_adjoint: one forward + reverse pass over the pieces of a
 * schedule (piece_len as given to gs_joint_schedule), lanes_per_piece 16,
 * 32 or 64 (anything else: 64) lanes per piece; then per grain loss_g[K],
 * grad_ubi[K*9], grad_t[K*3], and when not NULL geo_part[K*26] (dL/dgeo of
 * grain g, slot 16 = dL/d(1/wavelength)) and nsums[K*15] = sum h G^T and
 * sum G G^T (00 01 02 11 12 22) for a Gauss-Newton UBI update; out[27] =
 * dL/dgeo (26, slot 16 is dL/dwavelength) and L.  Every sum is taken in an
 * order fixed by the schedule (deterministic).  ws: the workspace, or NULL
 * for the library's per-device buffer (calls overlapping on one device must
 * pass their own).  The pair's forward pass (compute_xlylzl, compute_gv)
 * uses FMA and a reciprocal square root, so G is within a few ulps of
 * compute_gv's (gs_joint_hkl keeps compute_gv's own arithmetic).
 * gspace != 0: the g-space objective 1/2 sum ||UB.hkl - G||^2 (UB =
 * UBI^-1), grad_ubi then holds dE/dUB = UB.H - R with score_and_refine's R
 * and H (src/closest.c:310-316).
 * idx may be NULL in gs_joint_hkl and gs_joint_adjoint: sc, fc, so and co
 * are then per list entry, already in list order (entry q is element q,
 * npk >= n_idx), and the pass streams them instead of gathering four cache
 * lines per entry. */
int gs_joint_hkl(const double *sc, const double *fc, const double *so, const double *co,
                 int npk, const double *geo, const double *ubis, const double *trans, int K,
                 const int *idx, const int *owner, long long n_idx, int *hkl,
                 gs_stream_t stream);
int gs_joint_schedule(const long long *off, int K, long long n_idx, int piece_len,
                      long long *piece_off, int *piece_grain, int *slots, long long cap,
                      long long *n_pieces);
long long gs_joint_workspace_doubles(long long n_pieces);
int gs_joint_adjoint(const double *sc, const double *fc, const double *so, const double *co,
                     int npk, const double *geo, const double *ubis, const double *trans, int K,
                     const int *idx, long long n_idx, const long long *off, const int *hkl,
                     const long long *piece_off, const int *piece_grain, const int *slots,
                     long long n_pieces, int piece_len, int gspace, double *loss_g,
                     double *grad_ubi, double *grad_t, double *geo_part, double *nsums,
                     double *out, int lanes_per_piece, double *ws, gs_stream_t stream);

/* gs_gauss_newton_ubis: the Gauss-Newton UBI step of a joint refinement,
 * after a gs_joint_adjoint call that wrote nsums.  Per grain out[9g..] =
 * (sum h G^T)(sum G G^T)^-1 from nsums[15g..] (the least-squares UBI for
 * the fixed hkl and the current g-vectors), by the cofactor inverse; a
 * grain whose sum G G^T has determinant 0 keeps ubis[9g..] (NaN when ubis
 * is NULL), as score_and_refine keeps a UBI it cannot solve for
 * (src/closest.c:320-336, which fits UB in g-space instead).  out may be
 * ubis.  Device pointers, stream-ordered. */
int gs_gauss_newton_ubis(const double *nsums, const double *ubis, int K, double *out,
                         gs_stream_t stream);

/* ---- compute_gv per grain + score_and_assign, all grains in order ----
 * The loop of refinegrains.assignlabels (ImageD11/refinegrains.py:
 * 718-734): for g = 0..K-1, compute_gv(xlylzl, omega, ..., trans[g]) then
 * score_and_assign(ubis[g], gv, tol, drlv2, labels, grain_labels[g]).
 * so/co: per-peak sin/cos of omega*omegasign (gs_sincos_omega); mat: the
 * wedge/chi matrix (9 host doubles) or NULL to build it from wedge, chi.
 * counts[g] = the value the g-th score_and_assign returns. */
int gs_gv_assign_many(const double *xlylzl, const double *so, const double *co, int ng,
                      double wvln, double wedge, double chi, const double *mat,
                      const double *ubis, const double *trans, const int *grain_labels, int K,
                      double tol, double *drlv2, int *labels, int *counts, gs_stream_t stream);

/* out[q] = compute_gv of peak idx[q] for the translation of grain owner[q]
 * (n_idx rows of 3); rows with an out-of-range index are NaN. */
int gs_gv_lists(const double *xlylzl, const double *so, const double *co, int ng, double wvln,
                double wedge, double chi, const double *mat, const double *trans, int K,
                const int *idx, const int *owner, long long n_idx, double *out,
                gs_stream_t stream);

/* ---- refinegrains driver ---- */

/* refinegrains.compute_gv (ImageD11/refinegrains.py:366-456) for every list
 * entry q at once: peak idx[q] seen by request owner[q] (a grain at
 * translation trans[owner], R requests).  Device pointers; so/co are the
 * per-peak sin/cos of ((PI/180)*omega)*omegasign, omega the raw angles in
 * degrees (read only with mats).
 *   mats == NULL: observed geometry (tth, eta of
 *     transform.compute_tth_eta_from_xyz, g of compute_g_vectors at
 *     omega*omegasign);
 *   mats != NULL: the omega-float branch (OmFloat = True) with fitted UBIs
 *     mats (R,3,3): omega_calc from uncompute_g_vectors of m^-1.round(m.g),
 *     wrapped and clipped to +-slop, and g at omega_calc.  rb (3) and pre
 *     (9, NULL when wedge == chi == 0) are the host-side vectors
 *     post^T.(-1/wavelength,0,0) and chiwedge^T of uncompute_g_vectors
 *     (transform.py:529-564).
 * tth, eta, omega_calc may be NULL; gv is (M,3).  Out-of-range entries
 * give NaN. */
int gs_grain_gv(const double *xlylzl, const double *so, const double *co, const double *omega,
                int ng, double omegasign, double wvln, double wedge, double chi,
                const double *trans, const double *mats, int R, const double *rb,
                const double *pre, double slop, const int *idx, const int *owner, long long M,
                double *tth, double *eta, double *omega_calc, double *gv, gs_stream_t stream);

/* gs_grain_gv's per-entry arithmetic on the host, host pointers (a CPU
 * self-test of the kernel's code; no device is touched). */
int gs_grain_gv_host(const double *xlylzl, const double *so, const double *co,
                     const double *omega, int ng, double omegasign, double wvln, double wedge,
                     double chi, const double *trans, const double *mats, int R, const double *rb,
                     const double *pre, double slop, const int *idx, const int *owner, long long M,
                     double *tth, double *eta, double *omega_calc, double *gv);

/* ---- point-by-point refinement ---- */

/* The per-voxel selection and first scoring of refine_map
 * (ImageD11/sinograms/point_by_point.py:1432-1545, get_voxel_idx :359-368)
 * for every (voxel, UBI) task at once.  Device pointers.  Peaks: so, co
 * (sin, cos of omega), dty, gv (ng,3).  Voxels: pts (P,2) = (xi0, yi0) and
 * ptask (P+1) = the voxel's range of tasks (tasks sorted by voxel); ubis
 * (T,3,3).  A peak is assigned to a task when |y0 - xi0 sin - yi0 cos -
 * dty| <= ystep and sum((h - round(h))^2) < tol^2, h = UBI.g, round half
 * to even.
 *   task_off == NULL: counting pass, counts[t] = assigned peaks of task t
 *     (zeroed here);
 *   task_off != NULL: filling pass with the counts of the counting pass and
 *     their exclusive prefix sums task_off (T+1): writes out_peak and
 *     out_hkl (3 ints) of each assigned peak at task_off[t] + a slot
 *     (cursor (T) is caller scratch, zeroed here); the order inside a task
 *     is unspecified. */
int gs_pbp_assign(const double *so, const double *co, const double *dty, const double *gv,
                  int ng, const double *pts, const int *ptask, int P, const double *ubis, int T,
                  double y0, double ystep, double tol, int *counts, const long long *task_off,
                  int *cursor, int *out_peak, int *out_hkl, gs_stream_t stream);

/* refine_map's merge (ImageD11/sinograms/point_by_point.py:338-356) for
 * all groups at once: out (G, 7) = per group g the sums of w, sc w, fc w,
 * omega w, dty w, xpos w, eta w over rows r in [goff[g], goff[g+1]) of the
 * grouped order, row r being peak p = peak[perm[r]] (npk peaks) with
 * w = si[p]; each sum from 0.0 in row order (np.bincount's order), each
 * product x * w.  Device pointers; perm and peak int64. */
int gs_pbp_merge_sums(const long long *perm, const long long *peak, long long npk,
                      const double *si, const double *sc, const double *fc, const double *om,
                      const double *dty, const double *xpos, const double *eta,
                      const long long *goff, int G, double *out, gs_stream_t stream);

/* The least-squares fits of refine_map, one thread per task t < T, as
 * np.linalg.lstsq solves them (minimum-norm solution, singular values of
 * the weighted design matrix <= eps * max(rows, columns) * the largest
 * taken as zero), rows of task t = [off[t], off[t+1]) of `design`:
 *   gs_pbp_ubi_fit: design (rows, 6) = (w g, w hkl) (weighted_lstsq_ubi_fit,
 *     ImageD11/sinograms/point_by_point.py:408-419) for tasks with
 *     fit[t] != 0; ubifit (T,3,3) (NaN when not fitted), rank (T) as lstsq
 *     returns it, worth (T) = rank 3, cond < 1e14, det > 0, matrix_rank 3
 *     (:1595-1597), ubi_out = ubifit where worth else ubi_in, ub0 (T,3,3) =
 *     U.B0 with U from ubifit's cell (:443-490, B0 9 host doubles) where
 *     worth, NaN elsewhere;
 *   gs_pbp_strain_fit: design (rows, 7) = (w M, w strain) (:1652-1656)
 *     for tasks with worth[t] != 0; eps (T, 6) = (sxx syy szz sxy sxz syz),
 *     NaN for the other tasks.
 * Non-finite rows give NaN (numpy's lstsq fails on them). */
int gs_pbp_ubi_fit(const double *design, const long long *off, const int *fit,
                   const double *ubi_in, const double B0[9], int T, double *ubifit, int *rank,
                   int *worth, double *ubi_out, double *ub0, gs_stream_t stream);
int gs_pbp_strain_fit(const double *design, const long long *off, const int *worth, int T,
                      double *eps, gs_stream_t stream);

/* Row products of refine_map's merged-peak scoring and strain fit
 * (ImageD11/sinograms/point_by_point.py:1581-1590,1611-1614,1639): per row
 * r < R, h[3r+i] = ((u_i0 g_0 + u_i1 g_1) + u_i2 g_2), every product and
 * sum rounded on its own, with u = ubis[9 task[r] ..] (or ubis[9r ..] when
 * task is NULL) and g = g[3r ..]; h may be NULL.  With hkl and ok given
 * (both or neither): hkl = h rounded half to even, ok[r] = ((e_0^2 +
 * e_1^2) + e_2^2) < tolsq with e = h - hkl.  Device pointers. */
int gs_pbp_rows_dot(const double *ubis, const long long *task, long long R, const double *g,
                    double tolsq, double *h, long long *hkl, unsigned char *ok,
                    gs_stream_t stream);

/* Per-segment column sums of a row-major (rows, C) array: out (S, C),
 * out[s][c] = the sum of rows off[s] .. off[s+1]-1 of column c, added one
 * row at a time from 0.0 (np.bincount's order: refine_map's merge,
 * ImageD11/sinograms/point_by_point.py:338-356).  Device pointers; off is
 * (S+1) int64. */
int gs_segment_sum(const double *v, int C, const long long *off, int S, double *out,
                   gs_stream_t stream);
/* The same layout, out[s][c] = the maximum of 0.0 and rows off[s] ..
 * off[s+1]-1 of column c (NaN if any is NaN): refine_map's per-task weight
 * maximum (:1650), the 0.0 being the zeros torch reduces into. */
int gs_segment_max(const double *v, int C, const long long *off, int S, double *out,
                   gs_stream_t stream);

/* ---- column-file geometry ---- */

/* compute_geometry (src/cdiffraction.c:31-94; f2py src/_cImageD11.pyf:534-549):
 * per peak out[k] = (tth, eta, ds, gx, gy, gz)
 * from lab positions and omega (columnfile.updateGeometry,
 * ImageD11/columnfile.py:469-504).  Host pointers. */
void compute_geometry(double xlylzl[][3], double omega[], double omegasign, double wvln,
                      double wedge, double chi, double t[3], double out[][6], int n);

/* compute_xlylzl_xpos_variable (src/cdiffraction.c:245-277; f2py
 * src/_cImageD11.pyf:585-601): compute_xlylzl
 * with xpos[k] subtracted from x (Ctransform.sf2gvlocal,
 * ImageD11/transform.py:740-751).  Host pointers. */
void compute_xlylzl_xpos_variable(double s[], double f[], double p[4], double r[9],
                                  double dist[3], double xpos[], double xlylzl[][3], int n);

/* the same two on device memory */
int gs_compute_geometry_dev(const double *xlylzl, const double *omega, double omegasign,
                            double wvln, double wedge, double chi, const double t[3], double *out,
                            int n, gs_stream_t stream);
int gs_compute_xlylzl_xpos_dev(const double *s, const double *f, const double p[4],
                               const double r[9], const double dist[3], const double *xpos,
                               double *xlylzl, int n, gs_stream_t stream);

/* compute_geometry with per-peak sin/cos of omega and the wedge/chi matrix
 * given (the arithmetic after the libm calls). */
int gs_compute_geometry_trig_dev(const double *xlylzl, const double *so, const double *co,
                                 const double mat[9], double wvln, const double t[3], double *out,
                                 int n, gs_stream_t stream);

/* compute_xlylzl -> compute_geometry fused, the updateGeometry pass of
 * ImageD11/columnfile.py:491-504 (Ctransform.sf2xyz + xyz2geometry);
 * xlylzl_out may be NULL. */
int gs_sf2geometry(const double *s, const double *f, const double *omega, const double p[4],
                   const double r[9], const double dist[3], double omegasign, double wvln,
                   double wedge, double chi, const double t[3], double *xlylzl_out, double *out,
                   int n, gs_stream_t stream);

/* ---- indexer: ring assignment, pair search, trial orientations ---- */

/* closest (src/closest.c:188-207; f2py src/_cImageD11.pyf:88-103): the index in x of
 * the smallest |x[i] - v[j]| (first in x-major order) and that value;
 * 0 and 99 when none is below 99.  Host pointers. */
void closest(double x[], double v[], int *ibest, double *best, int nx, int nv);

/* quickorient (src/cdiffraction.c:293-328; src/_cImageD11.pyf:603-613): UBI rows 0, 1 hold g1, g2 on
 * entry and the orientation BT . [t1; t2; t3] on return.  Host pointers. */
void quickorient(double UBI[9], double BT[9]);

/* ds[k] = |gv[k]| as np.sqrt((gv*gv).sum(axis=1))
 * (ImageD11/indexing.py:345). */
int gs_gv_ds(const double *gv, int ng, double *ds, gs_stream_t stream);

/* indexer.assigntorings (ImageD11/indexing.py:447-457): ra[k] = the ring
 * j with the smallest |ds[k] - ringds[j]| < tol (first on ties), else -1. */
int gs_ring_assign(const double *ds, int n, const double *ringds, int nr, double tol, int *ra,
                   gs_stream_t stream);

/* out[r] = gv[idx[r]] / |gv[idx[r]]| (indexer.find, ImageD11/indexing.py:
 * 696-706); NaN rows for out-of-range indices. */
int gs_unit_rows(const double *gv, int ng, const int *idx, int n, double *out,
                 gs_stream_t stream);

/* indexer.find, cosine_tol > 0 (ImageD11/indexing.py:710-716): for each
 * row i of n1 (N1 x 3 unit vectors of ring-1 peaks), closest(n2 . n1[i],
 * cs) over the N2 ring-2 unit vectors -> ibest[i] (row of n2), best[i].
 * nv <= 4096. */
int gs_find_hits(const double *n1, int N1, const double *n2, int N2, const double *cs, int nv,
                 int *ibest, double *best, gs_stream_t stream);

/* indexer.find, cosine_tol < 0 (ImageD11/indexing.py:717-723): counts the
 * (i, m, j) with |cs[m] - n2[j] . n1[i]| < mtol into *counter (device,
 * zeroed here); when keys is not NULL also writes key = (i*nv + m)*N2 + j
 * for the first cap of them, in no particular order (sort to get the
 * reference's i, m, j order).  nv <= 4096. */
int gs_find_all(const double *n1, int N1, const double *n2, int N2, const double *cs, int nv,
                double mtol, unsigned long long *counter, long long *keys, long long cap,
                gs_stream_t stream);

/* unitcell.orient + quickorient (ImageD11/unitcell.py:617-653) for nh peak
 * pairs (hi[h], hj[h]) with the sorted cosine table c2ab (nc entries) and
 * its BT matrices bt[9b..]: ubis[9h..] = the trial UBI.  bsel NULL: b is
 * the entry nearest the observed cosine (orient with crange <= 0, the
 * trial of indexer.scorethem, ImageD11/indexing.py:784-800); otherwise
 * b = bsel[h] (the crange > 0 variants).  choice[h] = b and costheta[h] =
 * the observed cosine when not NULL.  Feed ubis to gs_score_many to score
 * every trial in one launch. */
int gs_orient_hits(const double *gv, int ng, const int *hi, const int *hj, int nh,
                   const double *c2ab, int nc, const double *bt, const int *bsel, double *ubis,
                   int *choice, double *costheta, gs_stream_t stream);

/* The peaks each of K UBIs indexes with the reference arithmetic
 * (drlv2 < thr): thr = tol^2 gives the peaks score_and_refine fits
 * (src/closest.c:287-318), min(tol^2, 1) the mask of indexer.getind
 * (score_and_assign from drlv2 = 1, ImageD11/indexing.py:1060-1088).
 * cursor (K ints, zeroed here) receives the counts; with off (K+1 CSR
 * offsets from those counts) and idx, the peak indices are also written,
 * grain g's in idx[off[g]..off[g+1]) in no particular order.  For
 * thr <= 0.01 the one-row sieve of GS_SCORE_SIEVE runs first. */
int gs_select_many(const double *ubis, int K, const double *gv, int ng, double thr,
                   const long long *off, int *cursor, int *idx, gs_stream_t stream);

/* One iteration of B Nelder-Mead simplexes of n <= 64 variables stepped in
 * lockstep, on HOST arrays, split where the objective is called
 * (ImageD11/simplex.py:115-227, the reference's one-problem loop, which
 * refinegrains.refinepositions runs per grain).  S (B, n+3, n): vertices,
 * centroid, reflected point; E (B, n+1): vertex errors; flags are bytes.
 * gs_simplex_order: per problem not done, hi/lo/second vertex, the
 * convergence test (done), lowest/iters = lo/it, and for the problems that
 * go on (a = 1, *n_active of them) the centroid and the reflected point.
 * The caller evaluates the reflected points into fr, then
 * gs_simplex_reflected accepts a lower one and writes the expansion or
 * contraction point into trial (step = 1, expand for an expansion); the
 * caller evaluates them into fb, then gs_simplex_trial accepts a lower one
 * and shrinks the simplexes whose contraction failed (shrink = 1; the
 * caller evaluates their non-lowest vertices in vertex order).  No device
 * work; every value is the reference's, comparison and rounding alike. */
int gs_simplex_order(int B, int n, double *S, const double *E, unsigned char *done,
                     long long *lowest, long long *iters, int it, double epsilon, double kR,
                     long long *hi, long long *lo, long long *second, unsigned char *a,
                     int *n_active);
int gs_simplex_reflected(int B, int n, double *S, double *E, const unsigned char *a,
                         const long long *hi, const long long *lo, const long long *second,
                         const double *fr, double kE, double kC, double *trial,
                         unsigned char *expand, unsigned char *step, int *n_step);
int gs_simplex_trial(int B, int n, double *S, double *E, const unsigned char *step,
                     const unsigned char *expand, const long long *hi, const long long *lo,
                     const double *fb, const double *trial, unsigned char *shrink,
                     int *n_shrink);

/* The decision loop of indexer.scorethem (ImageD11/indexing.py:778-838) on
 * HOST arrays, once every hit's follow-up is known: hits h = nhits-1 .. 0
 * (hit_i/hit_j their two peaks, npk their trial scores, cand[h] the row of
 * their follow-up or -1); row c's peak list is lists[list_off[c] ..
 * list_off[c+1]).  ga (npeaks, -1 = unassigned) is updated in place: an
 * accepted hit's peaks get label_base + (its rank) + 1.  accepted (nhits
 * ints) receives the accepted hits in order, *n_accepted their number,
 * *stop_at the first hit not visited (the reference pops every visited
 * hit), *nuniq the hits rejected as not unique.  No device work. */
int gs_index_decide(int nhits, const int *hit_i, const int *hit_j, const int *npk, int minpks,
                    const int *cand, const int *lists, const long long *list_off, int ncand,
                    double uniqueness, int max_grains, int label_base, int *ga, int npeaks,
                    int *accepted, int *n_accepted, int *stop_at, int *nuniq);

#ifdef __cplusplus
}
#endif

#endif /* GSCORE_H */
