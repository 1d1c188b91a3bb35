/*
 * itomo -- MI355X (gfx950) parallel-beam projectors and the iterations built on them.
 *
 * C ABI of libitomo.so.  Every pointer named image/sino/x/b/R/C/work is a DEVICE
 * pointer (HBM, float32, C-contiguous) unless the comment says "host".  `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Calls are stream-ordered and
 * asynchronous unless stated; they allocate no device memory unless stated.
 *
 * Plan scratch.  A plan owns one device scratch buffer (the transposed image of the
 * distance-driven forward, FP64 partial sums, the ray-driven back projection).  The
 * projection and iteration calls that take a plan (itomo_forward/back, itomo_sirt_*,
 * itomo_mlem_*, itomo_sps_*) size it for (plan, nslices, launch shapes) on first use and
 * whenever a call needs more: that growth synchronizes `stream`, frees the old buffer and
 * allocates the new one.  Calls that need no more than the buffer holds allocate
 * nothing.  A plan must not be used from two streams at once.
 *
 * Array layouts are the reference's:
 *   image : npix x npix, row-major (img[row][col])
 *   sino  : nang x npix; for the distance-driven projector row nang-1-v holds view v of
 *           the fftshift-ed angle list (radon_dd.c:79, radon_distance_driven.pyx:27);
 *           for the ray-driven and pixel-driven projectors row v holds view v
 *           (radon_rd.c:90, radon_pd.c:61; the pixel-driven wrapper fftshifts the angles,
 *           radon_pixel_driven.pyx:30); for the slant-stacking projector row nang-1-v
 *           holds view v (radon_ss.c:111; its wrapper does not fftshift).
 * Batched calls take `nslices` images (stride npix*npix) and sinograms (stride nang*npix)
 * back to back.
 *
 * Status: 0 = ITOMO_OK, negative = error (itomo_last_error() gives the detail).
 */
#ifndef ITOMO_H
#define ITOMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ITOMO_OK 0
#define ITOMO_EINVAL (-1)     /* bad argument (size, pointer, flag) */
#define ITOMO_EANGLE (-2)     /* an angle the reference C cannot handle (see plan) */
#define ITOMO_EHIP (-3)       /* HIP runtime error */
#define ITOMO_ENOMEM (-4)     /* host allocation failed */

#define ITOMO_PROJ_DD 0       /* distance-driven, radon_dd.c */
#define ITOMO_PROJ_RD 1       /* ray-driven, radon_rd.c */
#define ITOMO_PROJ_PD 2       /* pixel-driven, radon_pd.c */
#define ITOMO_PROJ_SS 3       /* slant-stacking, radon_ss.c */

#define ITOMO_ARITH_FAST 0    /* float arithmetic of the reference as built with its own
                                 flags (-O3 -ffast-math, create_module.py:14) */
#define ITOMO_ARITH_IEEE 1    /* float arithmetic of the reference C as written (IEEE) */

typedef struct itomo_plan itomo_plan;

/* Library version string and the detail of the last error on this thread. */
const char *itomo_version(void);
const char *itomo_last_error(void);

/* ---------------------------------------------------------------------------------
 * Reference-signature entry points.
 * Replace `void radon_dd(float* image, int npix, float* angles, int nang, int oper,
 * float* sino)` (radon_dd.c:54) and `radon_rd` (radon_rd.c:76), which the Cython
 * modules bind at radon_distance_driven.pyx:8 and radon_ray_driven.pyx:8.
 * Same arguments and meaning: `angles` (HOST, radians, as the C receives them, i.e.
 * already fftshift-ed for dd), oper 0 = forward (accumulates into sino), 1 = back
 * (accumulates into image).  Adds a stream and a status return.  Synchronous with
 * respect to the host for the plan upload only; the projection itself is queued on
 * `stream`.  Allocates (and frees) a small per-call geometry table.
 * ------------------------------------------------------------------------------- */
int itomo_radon_dd(float *image, int npix, const float *angles, int nang, int oper,
                   float *sino, void *stream);
int itomo_radon_rd(float *image, int npix, const float *angles, int nang, int oper,
                   float *sino, void *stream);
/* Replaces `void radon_pd(float* image, int npix, float* angles, int nang, int oper,
 * int method, float* sino)` (radon_pd.c:20), bound at radon_pixel_driven.pyx:7.
 * method 0 = nearest bin, 1 = linear split between two bins (the facade always passes 1,
 * class_projectors_radon.py:62,75).  Method 0 follows the reference's unchecked
 * indexing, sino[v*npix + l]: a sample with l outside [0, npix) lands in the
 * neighbouring sinogram row; flat indices outside the array (where the reference writes
 * out of bounds) are skipped. */
int itomo_radon_pd(float *image, int npix, const float *angles, int nang, int oper,
                   int method, float *sino, void *stream);
/* Replaces `void radon_ss(float* image, int npix, float* angles, int nang, int oper,
 * int method, float* sino)` (radon_ss.c:93), bound at radon_slant_stacking.pyx:7.
 * method 0 = nearest pixel, 1 = linear between two pixels (the facade passes 1,
 * class_projectors_radon.py:68,81).  For lines that miss the image square (|t| close to
 * npix/2) the reference's ray_limits yields a step range outside the image and its loop
 * indexes outside the image rows (out-of-bounds reads and writes); here samples outside
 * the image are skipped, which equals the reference wherever its samples stay inside. */
int itomo_radon_ss(float *image, int npix, const float *angles, int nang, int oper,
                   int method, float *sino, void *stream);

/* ---------------------------------------------------------------------------------
 * Plans: per-geometry constants (per-view sin/cos in the reference's float
 * arithmetic, branch, output row) uploaded once.  `angles` is HOST memory, radians,
 * in the order the reference C receives them.
 * Distance-driven plans reject (ITOMO_EANGLE) views that sweep rows with sin <= 0 or
 * columns with cos <= 0: for those the reference's sweep indexes bin -1 (undefined
 * behaviour in radon_dd.c:158-188).  Angles in [0, pi) are always accepted.
 * A distance-driven plan whose views include both narrow ones (at most 8 candidate pixels
 * per bin and line) and wider ones also creates a HIP stream (default priority) and two
 * events: its forward projections run the wider views' launch on that stream, forked
 * from the caller's stream and joined back into it, so callers see one ordered stream
 * (hipGraph capture records the fork and join).
 * Replaces class_projectors_radon.projectors.__init__ (class_projectors_radon.py:47-55).
 * ------------------------------------------------------------------------------- */
int itomo_plan_create(itomo_plan **plan, int projector, int npix, const float *angles,
                      int nang, int arith);
/* Pixel-driven plan with an explicit method (radon_pd.c:20 `method`); itomo_plan_create
 * with ITOMO_PROJ_PD uses method 1, the facade's (class_projectors_radon.py:62). */
int itomo_plan_create_pd(itomo_plan **plan, int npix, const float *angles, int nang,
                         int method, int arith);
/* Slant-stacking plan with an explicit method (radon_ss.c:93 `method`); itomo_plan_create
 * with ITOMO_PROJ_SS uses method 1, the facade's (class_projectors_radon.py:68).  The plan
 * holds a per-(view, bin) step-range table (8 bytes per bin). */
int itomo_plan_create_ss(itomo_plan **plan, int npix, const float *angles, int nang,
                         int method, int arith);
int itomo_plan_destroy(itomo_plan *plan);
/* Host copy of the per-view sin/cos the plan uses (for checking). */
int itomo_plan_trig(const itomo_plan *plan, float *s_host, float *c_host);

/* Launch shapes of the distance-driven kernels (no counterpart in the reference, whose
 * C has no launch shapes; used by the autotuner).  Forward variant f in
 * [0, ITOMO_FWD_VARIANTS): 256 / 128 / 64 bins per block.  Back variant b in
 * [0, ITOMO_BACK_VARIANTS): 64x64 tile with 4x4 pixels per thread / 32x32 with 2x2 /
 * 32x32 with 4x4.  Every shape produces bit-identical projections and iterates; FP64
 * monitor sums follow the block partition and may differ in the last bits.
 * Ray-, pixel-driven and slant-stacking plans accept only (0, 0).  Not to be called while work using the plan is
 * queued on another stream. */
#define ITOMO_FWD_VARIANTS 3
#define ITOMO_BACK_VARIANTS 3
int itomo_plan_set_shapes(itomo_plan *plan, int fwd_variant, int back_variant);
int itomo_plan_get_shapes(const itomo_plan *plan, int *fwd_variant, int *back_variant);

/* Ray-driven plans cache the projector's exact system matrix in HBM (both directions,
 * 8 bytes per walk step, at most 8 GiB together) on their first iteration call and
 * replay it in every later projection, bitwise equal to the walks.  No counterpart in
 * the reference (whose radon_rd.c walks every time).  mode: ITOMO_RDM_QUERY reports the
 * cached bytes (0 if none), ITOMO_RDM_BUILD builds it now (synchronous on stream; 0
 * bytes when the geometry does not fit or the device is out of memory), ITOMO_RDM_DROP
 * frees it and keeps the walks (EINVAL once a stream capture has recorded a projection
 * that replays it: the matrix then lives as long as the plan).  EINVAL for other
 * projectors. */
#define ITOMO_RDM_QUERY 0
#define ITOMO_RDM_BUILD 1
#define ITOMO_RDM_DROP 2
int itomo_plan_rd_matrix(itomo_plan *plan, int mode, size_t *bytes, void *stream);

/* Forward projection A (class_projectors_radon.py:60-68): sino = A image.
 * accumulate = 0 overwrites sino, 1 adds to it (the reference C always adds into the
 * zeroed array its wrapper allocates).  Deterministic. */
int itomo_forward(const itomo_plan *plan, const float *image, float *sino, int nslices,
                  int accumulate, void *stream);
/* Back projection A^T (class_projectors_radon.py:73-81): image = A^T sino. */
int itomo_back(const itomo_plan *plan, const float *sino, float *image, int nslices,
               int accumulate, void *stream);

/* ---------------------------------------------------------------------------------
 * SIRT (algorithms/art.py:211-237):
 *   R = 1 / A(1) and C = 1 / A^T(1), zero entries replaced by 1e-10 before inverting.
 *   x <- x + C * A^T( R * (b - A x) ), then, when pc_beta >= 0, x[x < 0] *= pc_beta
 *   (utils.phys_constr, art.py:314-321).
 * R: nang x npix, C: npix x npix (one geometry, shared by all slices).
 * work: nslices x nang x npix scratch (receives R*(b-Ax)).
 * stats (DEVICE, double[3*nslices] or NULL): per slice
 *   [0] sum (b - A x_k)^2  (x_k = the iterate entering this step)
 *   [1] sum (x_{k+1} - x_k)^2
 *   [2] sum x_k^2
 * all accumulated in FP64 in a fixed order (bitwise reproducible).
 * ------------------------------------------------------------------------------- */
int itomo_sirt_weights(const itomo_plan *plan, float *R, float *C, void *stream);
/* The two halves of one step (the iterate call is forward then back):
 *   forward: work = R * (b - A x), stats[0]
 *   back:    x = x + C * A^T work (+ non-negativity), stats[1], stats[2] */
int itomo_sirt_forward(const itomo_plan *plan, const float *x, const float *b,
                       const float *R, float *work, int nslices, double *stats,
                       void *stream);
int itomo_sirt_back(const itomo_plan *plan, const float *work, const float *C, float *x,
                    int nslices, float pc_beta, double *stats, void *stream);
int itomo_sirt_iterate(const itomo_plan *plan, float *x, const float *b, const float *R,
                       const float *C, float *work, int nslices, float pc_beta,
                       double *stats, void *stream);

/* The SIRT step of itomo_sirt_iterate with no device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * kernels; the plan's scratch is replaced by a fixed device array of
 * ITOMO_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Calls that need
 * more (itomo_sirt_iterate needs about 4*npix*npix bytes per slice plus FP64 partial
 * sums: up to 3 slices of 2048^2 x 1800) fail with ITOMO_EINVAL.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialized by the caller.  The first call on each device looks up the
 * array's device address. */
#define ITOMO_NOALLOC_SCRATCH_BYTES (64u << 20)
int itomo_sirt_iterate_noalloc(const itomo_plan *plan, float *x, const float *b,
                               const float *R, const float *C, float *work, int nslices,
                               float pc_beta, double *stats, void *stream);

/* SIRT constraints after a step (art.py:296-321), on the reconstruction square
 * [i1, i2) x [i1, i2) of each of nslices npix x npix images, in the reference's order:
 *   mask (m x m uint8, m = i2 - i1, or NULL): x = 0 where mask == 0 (utils.supp_constr,
 *     utils.py:264-266);
 *   mask_add (nmask x m x m uint8): for each mask in turn, x = mean of x over mask == 1
 *     there (utils.mask_constr, utils.py:279-281; the mean is summed in FP64 and rounded
 *     to float, where numpy sums the float32 values pairwise);
 *   pc == 1: x < 0 -> 0, then zero outside the resolution circle (centre m/2 - 1/2,
 *     radius m/2 - 1/2; utils.resol_circle_constr, utils.py:229-251); pc == 2: the circle
 *     only; other pc >= 0: x[x < 0] *= pc (utils.phys_constr); pc < 0: nothing.
 * x_old and stats (DEVICE double[3*nslices]) optional: stats[1] = sum (x - x_old)^2 and
 * stats[2] = sum x_old^2 over the whole image after the constraints (the monitor of
 * art.py:325).  work: DEVICE double scratch of ITOMO_CONSTR_WORK(nslices, npix) entries,
 * needed with masks or stats. */
#define ITOMO_CONSTR_WORK(nslices, npix) \
    ((int64_t)(nslices) * 3 * (((int64_t)(npix) * (npix) + 2047) / 2048))
int itomo_sirt_constraints(float *x, int npix, int nslices, int i1, int i2,
                           const unsigned char *mask, const unsigned char *mask_add, int nmask,
                           float pc, const float *x_old, double *stats, double *work,
                           void *stream);

/* The two halves of a SIRT step as elementwise passes, for projector pairs without
 * fused epilogues (the gridding pair): with ax = A x and atw = A^T work computed by the
 * caller, over nslices slices of n entries (R, C shared by all slices):
 *   itomo_sirt_residual: ax_io <- R * (b - ax_io), stats[0] = sum (b - ax)^2 (art.py:212);
 *   itomo_sirt_update:   x <- x + C * atw, then x[x < 0] *= pc_beta when pc_beta >= 0,
 *                        stats[1] = sum (x_new - x)^2, stats[2] = sum x^2;
 *   itomo_reciprocal:    out = 1 / v with zeros replaced by 1e-10 first (art.py:227-237).
 * stats (DEVICE double[3*nslices]) or NULL; work: DEVICE double scratch of
 * ITOMO_EPI_WORK(nslices, n) entries when stats is given. */
#define ITOMO_EPI_WORK(nslices, n) ((int64_t)(nslices) * 3 * (((int64_t)(n) + 2047) / 2048))
int itomo_sirt_residual(float *ax_io, const float *b, const float *R, int64_t n, int nslices,
                        double *stats, double *work, void *stream);
int itomo_sirt_update(const float *atw, const float *C, float *x, int64_t n, int nslices,
                      float pc_beta, double *stats, double *work, void *stream);
int itomo_reciprocal(const float *v, float *out, int64_t n, void *stream);

/* ---------------------------------------------------------------------------------
 * MLEM (algorithms/sir.py:204-283):
 *   inv_d = 1 / d, d = A^T(1) with zeros replaced by 1e-10 (sir.py:215-217, 283).
 *   Ax = A x; Ax[Ax == 0] += 2.220446e-16 (np.spacing(1)); q = b / Ax
 *   x <- x * (inv_d * A^T q)
 * work: nslices x nang x npix scratch (receives q).
 * stats (DEVICE, double[3*nslices] or NULL): [0] sum (Ax - b)^2 (after the zero fix),
 *   [1] sum (x_{k+1} - x_k)^2, [2] sum x_k^2.
 * ------------------------------------------------------------------------------- */
int itomo_mlem_weights(const itomo_plan *plan, float *inv_d, void *stream);
/* halves of one step: forward work = b / A x (stats[0]); back x *= inv_d * A^T work */
int itomo_mlem_forward(const itomo_plan *plan, const float *x, const float *b, float *work,
                       int nslices, double *stats, void *stream);
int itomo_mlem_back(const itomo_plan *plan, const float *work, const float *inv_d, float *x,
                    int nslices, double *stats, void *stream);
int itomo_mlem_iterate(const itomo_plan *plan, float *x, const float *b,
                       const float *inv_d, float *work, int nslices, double *stats,
                       void *stream);

/* ---------------------------------------------------------------------------------
 * Batched SIRT over several geometries in one persistent launch (no reference
 * counterpart: the reference fans slices out to a multiprocessing.Pool, art.py:289-295).
 * Member m is a distance-driven FAST plan with nslices[m] stacked slices and its own
 * x, b, R, C, work (as for itomo_sirt_iterate) and stats (DEVICE double[3*nslices[m]] or
 * NULL).  One iterate call = transpose + persistent forward + persistent back +
 * finalize; results equal itomo_sirt_iterate on each member bit for bit.  Creating a
 * batch allocates its work tables and scratch; iterate allocates nothing.
 * ------------------------------------------------------------------------------- */
typedef struct itomo_batch itomo_batch;
int itomo_batch_create(itomo_batch **batch, int nmembers, const itomo_plan *const *plans,
                       const int *nslices, float *const *x, const float *const *b,
                       const float *const *R, const float *const *C, float *const *work,
                       double *const *stats);
int itomo_batch_sirt_iterate(itomo_batch *batch, float pc_beta, void *stream);
int itomo_batch_destroy(itomo_batch *batch);

/* ---------------------------------------------------------------------------------
 * hipGraph capture of `iters` SIRT iterations (single plan or batch); the graph is
 * replayed with itomo_graph_launch on any stream.  Capture synchronizes once with the
 * work already queued on `stream`.  The graph reuses the arrays it was captured with.
 * itomo_graph_sirt allocates the graph's own scratch (the plan-scratch layout for
 * nslices, see above) before capture and itomo_graph_destroy frees it, so later calls
 * on the plan (more slices, other launch shapes) never touch memory a replay uses.  The
 * plan and the captured arrays must outlive the graph; destroy it only after its
 * replays have finished.  Batch graphs use the batch's own scratch.
 * ------------------------------------------------------------------------------- */
typedef struct itomo_graph itomo_graph;
int itomo_graph_sirt(itomo_graph **graph, const itomo_plan *plan, float *x, const float *b,
                     const float *R, const float *C, float *work, int nslices,
                     float pc_beta, double *stats, int iters, void *stream);
int itomo_graph_batch_sirt(itomo_graph **graph, itomo_batch *batch, float pc_beta,
                           int iters, void *stream);
int itomo_graph_launch(itomo_graph *graph, void *stream);
int itomo_graph_destroy(itomo_graph *graph);

/* ---------------------------------------------------------------------------------
 * SPS without penalty (algorithms/sir.py:348-540, beta = 0):
 *   setup: w = exp(-b / max(b)) over all slices, den = A^T(w * A(1)) per slice
 *   step:  work = (A x - b) * w;  x <- x - A^T(work) / (den + 2^-52), the quotient in FP64
 * w: nslices x nang x npix, den: nslices x npix x npix, work: nslices x nang x npix,
 * bmax_work: DEVICE float[1024].  stats as for SIRT with [0] = sum work^2.
 * ------------------------------------------------------------------------------- */
int itomo_sps_weights(const itomo_plan *plan, const float *b, int nslices, float *w,
                      float *den, float *work, float *bmax_work, void *stream);
int itomo_sps_iterate(const itomo_plan *plan, float *x, const float *b, const float *w,
                      const float *den, float *work, int nslices, double *stats,
                      void *stream);

/* ---------------------------------------------------------------------------------
 * SPS penalties (sir.py:455-560 over the reference's penalty module).
 * itomo_penalty replaces penalty.pyx:16-106 (huber_num/_den over huber.c:66-118,
 * tikhonov_num/_den over tikhonov.c:162-195, haar_num/_den over haar.c:242-275): for each
 * of nslices nr x nc images, r_num[i][j] = sum over the 3x3 neighbours inside the image of
 * w_k * phi'(I[i][j] - I[nb]) and r_den the curvature term (huber: 2 w_k / (1 + |d/delta|);
 * tikhonov: the constant 2; haar: 0), neighbour weights 1/sqrt(2), 1, 4
 * (penalty_weights.c:7-15), accumulated in the reference's order.  arith FAST = the
 * module's -ffast-math build (create_penalty_module.py), IEEE = the C as written.
 * itomo_sobel_magnitude: sqrt(gx^2 + gy^2) of scipy.ndimage.sobel(x, axis, mode='reflect')
 * for both axes (sir.py:339-343), the image the Huber penalty is applied to.
 * itomo_sps_step: one SPS iteration with the reference's options (sir.py:428-540):
 *   work = (A x - b) w; num = A^T work;
 *   x <- x - (num + beta R_num) / (den + beta R_den + 2^-52), quotient in FP64, with R from
 *   x (tikhonov, haar) or from sobel_magnitude(x) (huber, delta), penalty NONE: beta unused;
 *   clip != 0: x <- np.clip(x, 0, max x) over all slices (sir.py:539-540, the reference's
 *   default when there is no edge padding).
 * img_work: nslices x npix x npix floats (penalty NONE) or 4x that (with a penalty).
 * stats as for itomo_sps_iterate; [1] uses the clipped iterate.
 * ------------------------------------------------------------------------------- */
#define ITOMO_PENALTY_NONE 0
#define ITOMO_PENALTY_HUBER 1
#define ITOMO_PENALTY_TIKHONOV 2
#define ITOMO_PENALTY_HAAR 3
int itomo_penalty(const float *image, int nr, int nc, int nslices, int kind, float delta,
                  int arith, float *r_num, float *r_den, void *stream);
int itomo_sobel_magnitude(const float *x, int nr, int nc, int nslices, float *g, void *stream);
int itomo_sps_step(const itomo_plan *plan, float *x, const float *b, const float *w,
                   const float *den, float *work, float *img_work, int nslices, int penalty,
                   float beta, float delta, int clip, double *stats, void *stream);

/* ---------------------------------------------------------------------------------
 * Conjugate gradients on the normal equations (algorithms/cg.py:83-176) and the
 * ADMM-TV inner solve (cg.py:256-310, admm.py:470-646), float32 vectors with FP64
 * inner products and step sizes kept on the device.
 * scal: DEVICE double[4*nslices] per slice {rr, pHp, rr_new, tv}; work: DEVICE double
 * scratch of ITOMO_VEC_WORK(nslices) entries.  n = elements per slice.
 * Gradient images g hold 2 components per slice (nslices x 2 x npix x npix): d/drow
 * and d/dcol of np.gradient (edge_order 1, unit spacing).
 * ------------------------------------------------------------------------------- */
#define ITOMO_VEC_WORK(nslices) ((int64_t)(nslices) * 771)
/* g = G x (cg.py:62-63) and out = G^T g, its exact adjoint (cg.py:66-71). */
int itomo_grad(const float *x, int npix, int nslices, float *g, void *stream);
int itomo_grad_adjoint(const float *g, int npix, int nslices, float *out, void *stream);
/* Hp += lambd1*p + mu*G^T g (g may be NULL: no TV terms); scal[1] = <p, Hp>. */
int itomo_cg_hp(float *Hp, const float *p, const float *g, int npix, int nslices,
                float lambd1, float mu, double *scal, double *work, void *stream);
/* a = float(scal[0] / (scal[1] + 2^-52)); x += a p; r -= a Hp; scal[2] = <r, r>.
 * A slice whose scal[0] is negative has stopped (below): x and r stay, scal[2] = 0. */
int itomo_cg_update(float *x, float *r, const float *p, const float *Hp, int64_t n,
                    int nslices, double *scal, double *work, void *stream);
/* p = r + float(scal[2] / (scal[0] [+ 2^-52])) p; then scal[0] = scal[2].
 * The reference's loops stop once ||r|| <= 1e-12 (cg.py:101,128-142, 203,215-231,
 * 271,274-292): when sqrt(scal[2]) <= 1e-12, p stays and scal[0] becomes -1, which
 * itomo_cg_update and this call then treat as stopped until a residual call (cg_dot into
 * slot 0, tv_residual, l1_residual) stores a new scal[0]. */
int itomo_cg_direction(float *p, const float *r, int64_t n, int nslices, int with_eps,
                       double *scal, void *stream);
/* scal[slot] = <a, b> per slice. */
int itomo_cg_dot(const float *a, const float *b, int64_t n, int nslices, int slot,
                 double *scal, double *work, void *stream);
/* r = atb + G^T v - (atax + lambd1*x + G^T gx) (cg.py:258); scal[0] = <r, r>. */
int itomo_tv_residual(float *r, const float *atb, const float *atax, const float *x,
                      const float *gx, const float *v, int npix, int nslices, float lambd1,
                      double *scal, double *work, void *stream);
/* v = mu*u - alpha over n elements (admm.py:532-534). */
int itomo_tv_rhs(float *v, const float *u, const float *alpha, int64_t n, float mu,
                 void *stream);
/* u = soft(G x + alpha/mu, lambd2/mu); alpha += G x - u; scal[3] = sum (G x - u)^2
 * (admm.py:460-461, 565-593, utils.py:214-216). */
int itomo_admm_tv_update(const float *x, float *u, float *alpha, int npix, int nslices,
                         float mu, float lambd2, double *scal, double *work, void *stream);
/* Lasso-L1 ADMM (admm.py:291-447 with cg.py:189-245): the inner CG solves
 * (A^T A + (lambd1 + mu) I) x = A^T b + v with v = mu u - alpha (itomo_tv_rhs); its
 * residual r = (atb + v) - (atax + c x), c = lambd1 + mu, scal[0] = <r, r>; its H p is
 * itomo_cg_hp with g = NULL and lambd1 = c.  The multiplier step: u = soft(x + alpha/mu,
 * lambd2/mu), alpha += x - u, scal[3] = sum (x - u)^2. */
int itomo_l1_residual(float *r, const float *atb, const float *atax, const float *x,
                      const float *v, int npix, int nslices, float c, double *scal,
                      double *work, void *stream);
int itomo_admm_l1_update(const float *x, float *u, float *alpha, int npix, int nslices,
                         float mu, float lambd2, double *scal, double *work, void *stream);

/* ---------------------------------------------------------------------------------
 * Sinogram filtering for filtered back projection (filters.py:35-113, dpc = False).
 * itomo_filter_taps (HOST): taps[k], k < npix, of the real even impulse response of the
 *   reference's zero-padded FFT filter; ftype 0 ramp, 1 shepp-logan, 2 cosine,
 *   3 hamming, 4 hanning, 5 none.  npix <= 4096.
 * itomo_filter_rows: out[r][n] = sum_m in[r][m] * taps[|n-m|] (DEVICE arrays, FP64
 *   accumulation) -- identical to the reference's FFT filtering for n < npix.
 * ------------------------------------------------------------------------------- */
int itomo_filter_taps(int npix, int ftype, double *taps_host);
int itomo_filter_rows(const float *in, float *out, int nrows, int npix, const double *taps,
                      void *stream);

/* ---------------------------------------------------------------------------------
 * Gridding (gridrec) projector pair: the reference's Fourier-regridding back
 * projection / filtered back projection (gridrec_v4_backproj.c:64 `gridrec_v4_backproj`,
 * bound at gridrec_v4.pyx:7 and wrapped by gridrec_v4.pyx:32-62 `backproj`) and forward
 * projection (gridrec_v4_forwproj.c:59 `gridrec_v4_forwproj`, gridrec_v4.pyx:67-94
 * `forwproj`), used through class_projectors_grid.py:124-151 (A, At, fbp).
 * itomo_gridrec_create takes the set-up the reference computes in Python
 * (gridrec_lut.py:268 configure_regridding, class_projectors_grid.py:79-117) and in
 * filters.c:151 calc_filter, all HOST arrays:
 *   trig     : nang x {sin, cos} of the angles in degrees (backproj.c:308-309);
 *   lut      : the kernel look-up table (nlut floats), ltbl = nlut - 5 its resolution
 *              (class_projectors_grid.py:107);
 *   deapod   : the 1-D deapodiser with its centring signs (pdim floats); the 2-D matrix
 *              of the reference is deapod[j] * deapod[k];
 *   lconv_back, lconv_fwd : kernel window widths of the back projection (param[6]) and
 *              of the forward projection (fixed to 2*7/pi at forwproj.c:164);
 *   interp   : 0 nearest, 1 linear LUT interpolation; radon_degree: 1 for the
 *              first-derivative transform (zero edge padding, derivative filters);
 *   filt_at, filt_fbp : pdim/2 complex (re, im) filter values of the plain and the
 *              filtered back projection (calc_filter with type 0 and the chosen type);
 *   phase    : pdim/2 complex centre-of-rotation phase of the forward projection
 *              (forwproj.c:183-199, 363-364).
 * pdim must be a product of 2, 3 and 5, at most 4096 (npix <= 2048 at the reference's
 * oversampling 2 and 1.5).  Creating allocates the tables and two pdim x pdim complex
 * grids plus nang x pdim complex rows of scratch; the projections allocate nothing and
 * process `nslices` stacked slices one after another on `stream` (the object's scratch
 * is shared: calls on one object must not overlap).  Angles are DEGREES, images
 * npix x npix, sinograms nang x npix, row v = angle v.
 * itomo_fft_rows: in-place FFT of nrows rows of L complex floats (interleaved re, im),
 *   sign -1 = exp(-2 pi i jk/L) (FFTW_FORWARD), +1 = FFTW_BACKWARD, unnormalised;
 *   L a product of 2, 3 and 5 up to 4096.  Synchronous (allocates a twiddle table).
 * ------------------------------------------------------------------------------- */
typedef struct itomo_gridrec itomo_gridrec;
int itomo_gridrec_create(itomo_gridrec **g, int npix, int nang, int pdim,
                         const float *trig_host, const float *lut_host, int nlut, int ltbl,
                         const float *deapod_host, float lconv_back, float lconv_fwd,
                         int interp, int radon_degree, const float *filt_at_host,
                         const float *filt_fbp_host, const float *phase_host);
int itomo_gridrec_destroy(itomo_gridrec *g);
int itomo_gridrec_back(itomo_gridrec *g, const float *sino, float *image, int nslices,
                       int filtered, void *stream);
int itomo_gridrec_forward(itomo_gridrec *g, const float *image, float *sino, int nslices,
                          void *stream);
int itomo_fft_rows(void *data, int L, int nrows, int sign, void *stream);

/* ---------------------------------------------------------------------------------
 * FP64 reductions used by the monitoring / CG / ADMM / SPS scalars.
 * out (DEVICE double[1]) = sum_i a[i]*b[i] over n float32 entries, accumulated in FP64
 * in a fixed order.  b may equal a.  work: DEVICE scratch of ITOMO_DOT_WORK doubles.
 * ------------------------------------------------------------------------------- */
#define ITOMO_DOT_WORK 3072
int itomo_dot_f64(const float *a, const float *b, int64_t n, double *out, double *work,
                  void *stream);

/* ---------------------------------------------------------------------------------
 * Learned warm start: kernels of a small U-Net (no counterpart in the reference, which
 * has no learned component).  Tensors are DEVICE arrays in NHWC order; x, y, wt, mask,
 * and every other argument declared void* (except stream) hold BF16 (16-bit, round to nearest
 * even); float/double pointers are FP32/FP64.  Every product accumulates in FP32 on the
 * matrix cores (v_mfma_f32_16x16x32_bf16).
 * itomo_unet_conv3x3: y[p][co] = act(bias[co] + sum_{tap,ci} x[p+off(tap)][ci] *
 *   wt[co][tap][ci]) with zero padding, act = ReLU when relu != 0, then multiplied by
 *   (mask[p][co] > 0) when mask is non-null.  bias may be null.
 * itomo_unet_flip: wf[ci][tap][co] = wt[co][8-tap][ci], the weights that turn conv3x3
 *   into the gradient with respect to its input.
 * itomo_unet_wgrad: dw[co][tap][ci] += sum_p dy[p][co] x[p+off(tap)][ci] and, when db is
 *   non-null, db[co] += sum_p dy[p][co] (FP32 atomics: summation order not fixed).
 * itomo_unet_pool_fwd: 2x2 max pool (h, w even), idx (uint8) = argmax (first on ties).
 * itomo_unet_pool_bwd: dx = (skip + dy scattered to the argmax) * (xm > 0); skip may
 *   be null.
 * itomo_unet_upcat_fwd: y[n][y][x] = concat(a[n][y/2][x/2][0:ca], b[n][y][x][0:cb]).
 * itomo_unet_upcat_bwd: da = (2x2 sum of dy[.., 0:ca]) * (a > 0), db = dy[.., ca:].
 * itomo_unet_head_fwd: out[p] = img[p] + hb[0] + sum_c hw[c] f[p][c] (c <= 32).
 * itomo_unet_head_bwd: df[p][c] = dout[p] hw[c] (f > 0); dhw += sum_p dout f;
 *   dhb += sum_p dout.
 * itomo_unet_mse: grad = 2 (out - target) / n; loss[0] = mean (out - target)^2 (FP64);
 *   work: ITOMO_UNET_MSE_WORK doubles.
 * itomo_unet_adam: Adam with bias correction (step >= 1) on FP32 w; refreshes the BF16
 *   copy wbf when non-null.
 * itomo_unet_to_bf16: y = bf16(x).
 * ------------------------------------------------------------------------------- */
#define ITOMO_UNET_MSE_WORK 1024
int itomo_unet_conv3x3(const void *x, int n, int h, int w, int cin, const void *wt,
                       const float *bias, int cout, const void *mask, int relu, void *y,
                       void *stream);
int itomo_unet_flip(const void *wt, int cin, int cout, void *wf, void *stream);
int itomo_unet_wgrad(const void *x, const void *dy, int n, int h, int w, int cin, int cout,
                     float *dw, float *db, void *stream);
int itomo_unet_pool_fwd(const void *x, int n, int h, int w, int c, void *y, void *idx,
                        void *stream);
int itomo_unet_pool_bwd(const void *dy, const void *idx, const void *skip, const void *xm,
                        int n, int h, int w, int c, void *dx, void *stream);
int itomo_unet_upcat_fwd(const void *a, const void *b, int n, int h, int w, int ca, int cb,
                         void *y, void *stream);
int itomo_unet_upcat_bwd(const void *dy, const void *a, int n, int h, int w, int ca, int cb,
                         void *da, void *db, void *stream);
int itomo_unet_head_fwd(const void *f, const float *hw, const float *hb, const float *img,
                        int64_t npixels, int c, float *out, void *stream);
int itomo_unet_head_bwd(const void *f, const float *hw, const float *dout, int64_t npixels,
                        int c, void *df, float *dhw, float *dhb, void *stream);
int itomo_unet_mse(const float *out, const float *target, int64_t npixels, float *grad,
                   double *loss, double *work, void *stream);
int itomo_unet_adam(float *w, const float *g, float *m, float *v, void *wbf, int64_t n,
                    float lr, float beta1, float beta2, float eps, int step, void *stream);
int itomo_unet_to_bf16(const float *x, int64_t n, void *y, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ITOMO_H */
