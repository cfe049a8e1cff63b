/*
 * rajzel_amd.h -- C ABI of the MI355X (gfx950) implementation of RAJZEL's
 * adjoint-state first-arrival traveltime (FATT) gradient path.
 *
 * The path replaced is the reference's per-evaluation shot loop
 *   grad_obj (grad_obj.c:10-79)
 *     -> eikonal (eikonal.c:10-76, sweep.c:10-104)
 *     -> calc_FA / calc_FA_res (calc_FA.c:10-19, calc_FA_res.c:10-77)
 *     -> eikonal_adj (eikonal_adj.c:10-59, sweep_adj.c:10-77,
 *                     init_travel_adj.c:12-60, init_recflag.c:12-35)
 *     -> ass_grad (ass_grad.c:11-49, cp_grad_frame.c:10-53)
 *   summed over shots (grad_obj.c:33-68, sum_grad_MPI.c:19-47).
 *
 * Conventions
 *   - Fields are float32, row-major NY x NX with x fastest: the contiguous
 *     block &A[1][1] of the reference's matrix() (util.c:177-204).
 *   - Grid indices are 0-based: reference index (iround(pos/DH), fd.h:18)
 *     minus one.  rz_grid_index() reproduces that rounding.
 *   - Pointers flagged "device" are HIP device pointers; all others are host.
 *   - Every call returns an rz_status; on failure rz_last_error() names the
 *     cause.  Nothing here aborts the process (the reference's err() does;
 *     the reference-surface shim maps failures onto it).
 *   - Results are bitwise those of the reference's single-rank float path
 *     (one MPI rank), outer-iteration counts included: the stopping test
 *     (eikonal.c:44, eikonal_adj.c:37) is decided from a double sum of the
 *     per-sweep changes where a rigorous bound places norm1's float sum on
 *     one side of the threshold, and otherwise from norm1's float sum taken
 *     in the reference's order (norm1.c:21-29, x outer, y inner).
 *     Between the two, a count of the cells that changed tightens the
 *     bound.  RAJZEL_AMD_ORDERED_NORM=1 in the environment at rz_create
 *     takes the ordered sum for every nonzero decision, =2 the count (tests
 *     of those paths).
 *
 * A context is single-threaded: calls on one context must not overlap.
 */
#ifndef RAJZEL_AMD_H
#define RAJZEL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rz_status {
    RZ_OK = 0,
    RZ_ERR_INVALID = 1,        /* bad argument, geometry or slowness          */
    RZ_ERR_HIP = 2,            /* HIP runtime error                           */
    RZ_ERR_NOMEM = 3,          /* device allocation failed                    */
    RZ_ERR_NOT_CONVERGED = 4,  /* a shot hit max_iter (results still written) */
    RZ_ERR_STATE = 5           /* call out of order (e.g. no slowness yet)    */
} rz_status;

typedef struct rz_ctx rz_ctx;

/* One model (grid + acquisition).  Globals of the reference it replaces:
 * NX, NY, DH, TIME, TTNORM, EPS_ADJ (globvar.h:12-15,96); srcpos/recpos
 * (sources.c:68-71, receiver.c:55-58).  max_iter bounds the outer
 * fast-sweeping iterations (the reference has no bound, eikonal.c:44,
 * eikonal_adj.c:37); 0 selects 1000. */
typedef struct rz_model_desc {
    int nx, ny;                /* >= 5 each                                  */
    float dh;                  /* DH (m), > 0                                */
    float time;                /* TIME: T starts at 2*TIME, lambda at 10*TIME */
    float ttnorm;              /* TTNORM: stop when mean |dT| < TTNORM        */
    float eps_adj;             /* EPS_ADJ water level of sweep_adj.c:60       */
    int max_iter;
    int nshots;
    const int *src_ix, *src_iy;  /* nshots source cells (0-based)           */
    int ntr;
    const int *rec_ix, *rec_iy;  /* ntr receiver cells (0-based), shared by all shots */
} rz_model_desc;

/* 0-based grid index of a position in metres: (int)floor(pos/DH + 0.5) - 1
 * with pos/DH a float division (fd.h:18, eikonal.c:31-32, receiver.c:55). */
int rz_grid_index(float pos_m, float dh);

/* Create a context on HIP device `device` holding `nmodels` models.
 * On failure *out is NULL and rz_last_error(NULL) gives the reason.
 *   rows_per_lane  1, 2, 4 or 8 rows of a band per lane; 0 = chosen by
 *                  the band count; RZ_ROWS_AUTOTUNE = measured by
 *                  rz_autotune_rows().  A band is 64*R rows; a grid may
 *                  have at most 8 bands at R <= 2, 4 at R = 4 and 8 at
 *                  R = 8 (NY <= 512, 1024, 1024, 4096).  Above 4 bands
 *                  at R = 8 the four sweeps of an iteration run one after
 *                  the other instead of two at a time.  A grid with
 *                  NY > 4096 and NX <= 4096 is stored and solved
 *                  transposed (R = 8 only; fields, gradients and receiver
 *                  times keep the grid's orientation and values).
 *   hbm_budget     bytes of device memory for per-shot fields; 0 = enough
 *                  for every shot at once when that fits the device's free
 *                  memory less 1/16 of the device, else what does fit.
 *                  Below all shots, shots run in batches that fit and
 *                  forward fields are recomputed per batch (results are
 *                  unchanged). */
#define RZ_ROWS_AUTOTUNE (-1)
int rz_create(rz_ctx **out, int device, int nmodels, const rz_model_desc *models, int rows_per_lane,
              size_t hbm_budget);

/* Autotuner of the rows per lane (the tile height of a wave: 64*R rows per
 * band): times one outer forward + adjoint iteration of these models at
 * every feasible R on the device and returns the fastest (and its time in
 * ms, if ms is not NULL).  Creates and destroys temporary contexts.  No
 * counterpart in the reference. */
int rz_autotune_rows(int device, int nmodels, const rz_model_desc *models, size_t hbm_budget, int *rows_per_lane,
                     double *ms);
void rz_destroy(rz_ctx *ctx);
const char *rz_last_error(const rz_ctx *ctx);

/* Run all further work on the HIP stream `hip_stream` (a hipStream_t);
 * NULL selects HIP's null stream.  Until the first call the context uses a
 * non-blocking stream of its own, which rz_get_stream returns. */
int rz_set_stream(rz_ctx *ctx, void *hip_stream);
void *rz_get_stream(const rz_ctx *ctx);

/* Slowness S = 1/Vp (calc_S.c:18-24), NY*NX.  Replaces the reference's S
 * matrix passed to grad_obj (fd.h:72).  Rejected unless finite and > 0. */
int rz_set_slowness(rz_ctx *ctx, int model, const float *S, int on_device);

/* Observed picks, nshots*ntr, shot-major; negative = muted
 * (calc_FA_res.c:39).  Replaces the <DATA_DIR>_<shot>.dat files read by
 * calc_FA_res.c:28-37. */
int rz_set_picks(rz_ctx *ctx, int model, const float *tobs, int on_device);

/* Initial contents (ntr) of the residual array Tres that grad_obj's caller
 * passes in (fd.h:72); muted picks inherit it (calc_FA_res.c:39-70).
 * Default zero. */
int rz_set_tres_carry(rz_ctx *ctx, int model, const float *tres0);

/* Forward only, every shot of every model: eikonal() per shot (eikonal.c:
 * 10-76) with receiver times (calc_FA.c:15-17); the driver loop of
 * forward.c:19-24. */
int rz_forward(rz_ctx *ctx);

/* One objective + gradient evaluation of every model: grad_obj()
 * (grad_obj.c:10-79) with precond() left to the caller.  All shots in one
 * launch per phase. */
int rz_gradient(rz_ctx *ctx);

/* The same evaluation with one shot per launch (the plain GPU form, the
 * reference's shot loop with each routine a launch).  Same results. */
int rz_gradient_per_shot(rz_ctx *ctx);

/* Adjoint source given directly: dT (nshots*ntr, shot-major) replaces the
 * residual Tmod - Tobs of calc_FA_res.c:42 (no pick is muted); the slope
 * scaling of calc_FA_res.c:52-69, the adjoint and the gradient assembly are
 * unchanged.  With dT = dL/dTmod the gradient is the reference's adjoint-
 * state gradient of L.  No counterpart in the reference (its residual always
 * comes from picks); it serves autograd callers. */
int rz_set_residuals(rz_ctx *ctx, int model, const float *dT, int on_device);

/* Residual scaling + adjoint + gradient for the rz_set_residuals sources.
 * Reuses the forward fields of the last rz_forward when every shot's fields
 * are still resident and the slowness is unchanged; otherwise recomputes the
 * forward per batch (HBM budget).  Receiver times of the recomputation are
 * identical to the earlier ones. */
int rz_adjoint(rz_ctx *ctx);

/* Results of the last rz_forward / rz_gradient / rz_adjoint. */
int rz_get_tmod(rz_ctx *ctx, int model, float *out, int on_device);      /* nshots*ntr */
int rz_get_tres(rz_ctx *ctx, int model, float *out, int on_device);      /* nshots*ntr */
/* T at each receiver's slope neighbour (the row below it, or above it on the
 * last row; calc_FA_res.c:52-69), nshots*ntr, after a forward.  With
 * rz_get_tmod it is all of T that calc_FA_res() reads. */
int rz_get_tnb(rz_ctx *ctx, int model, float *out, int on_device);
int rz_get_grad(rz_ctx *ctx, int model, float *out, int on_device);      /* NY*NX      */
/* L2 = 0.5 * sum of per-shot sums (grad_obj.c:61); l2_shot may be NULL. */
int rz_get_misfit(rz_ctx *ctx, int model, double *L2, double *l2_shot);
/* Outer iterations per shot; -1 = hit max_iter.  Either pointer may be NULL. */
int rz_get_iterations(rz_ctx *ctx, int model, int *it_f, int *it_a);
/* Per shot, how often the stopping tests of the last run read the fields
 * because the double sum could not decide: counts of changed cells
 * (count_*) and norm1's float sum in the reference's order (ordered_*),
 * forward and adjoint; sum_* is that float sum (norm1.c:21-25, before the
 * division by NX*NY) of the last iteration's test where it was taken, else
 * -1.  Any pointer may be NULL.  No counterpart in the reference (a cost
 * diagnostic). */
int rz_get_stop_passes(rz_ctx *ctx, int model, int *count_f, int *count_a, int *ordered_f, int *ordered_a,
                       float *sum_f, float *sum_a);
/* Tres array after the evaluation (the last shot's residuals). */
int rz_get_tres_carry(rz_ctx *ctx, int model, float *out);
/* Field of one shot, NY*NX: which = 0 traveltime T, 1 adjoint lambda.
 * Only for shots of the last batch (all shots without an HBM budget). */
int rz_get_field(rz_ctx *ctx, int model, int shot, int which, float *out, int on_device);

/* Per-shot entry points with host arrays, used by the reference-surface
 * shim (eikonal(), eikonal_adj(), ass_grad()).  `model` fixes the grid and
 * receivers; its slowness must be set.
 *   rz_eikonal_shot      eikonal.c:10-76 for a source cell: T, receiver
 *                        times, outer iterations.
 *   rz_eikonal_adj_shot  eikonal_adj.c:10-59 for a given T and Tres.
 *   rz_ass_grad_shot     ass_grad.c:11-49 without the per-shot taper
 *                        (rz_ass_grad_shot_at below applies it):
 *                        grad_shot = frame-copied lambda*S^3, grad += it. */
int rz_eikonal_shot(rz_ctx *ctx, int model, int src_ix, int src_iy, float *T, float *tmod, int *iters);
int rz_eikonal_adj_shot(rz_ctx *ctx, int model, const float *T, const float *tres, float *lam, int *iters);
int rz_ass_grad_shot(rz_ctx *ctx, int model, const float *lam, float *grad_shot, float *grad);

/* Per-shot gradient taper of ass_grad.c:31-37 (SWS_TAPER_CIRCULAR_PER_SHOT)
 * for `model`: each shot's frame-copied lambda*S^3 is tapered before it joins
 * the shot sum, and rz_gradient / rz_adjoint then copy the frame of the sum
 * again (grad_obj.c:70), bitwise as the reference.
 *   mode 0   none (the default)
 *   mode 1   taper_grad_shot.c:43-84, radius_m = SRTRADIUS (> 0): the
 *            gradient is multiplied by f0 + (1-f0)(0.5 - 0.5 cos(PI r/R))^a
 *            (f0 = 1e-3, a from frh = 0.1) within r <= R of the shot's source
 *            and then of each receiver in turn.  Shot s of the model uses
 *            source first_shot + s, which must be < nsrc.
 *   mode 2   taper_grad_shot1.c:32-62, filtsize = FILTSIZE: cells of the
 *            (2F+1) x (2F+1) windows around all nsrc sources (iround(pos/DH))
 *            and all receivers take 1e-5, in every shot.  Columns outside the
 *            grid wrap onto the neighbouring row, as in the reference's
 *            contiguous matrix; rows outside the grid, and cells before the
 *            first or past the last, are left alone (the reference writes
 *            outside its matrix there).
 * src_x_m, src_z_m are the survey's nsrc source positions in metres
 * (srcpos[1][1..nsrc], srcpos[2][1..nsrc], sources.c:68-71).  The factors are
 * computed here, once per geometry, on the host with the C library's pow
 * and cos (as the reference does), and kept on the device: at most 2^28 of
 * them (RZ_ERR_INVALID beyond, e.g. for a radius that covers most of a large
 * grid); a call with the same settings keeps them. */
int rz_set_shot_taper(rz_ctx *ctx, int model, int mode, float radius_m, int filtsize, int nsrc, const float *src_x_m,
                      const float *src_z_m, int first_shot);
/* rz_ass_grad_shot with the model's per-shot taper for source `src` (the
 * reference's ishot - 1, an index into rz_set_shot_taper's sources):
 * ass_grad.c:11-49 in full. */
int rz_ass_grad_shot_at(rz_ctx *ctx, int model, int src, const float *lam, float *grad_shot, float *grad);

/* Phase timing: with profiling on, rz_forward / rz_gradient record HIP
 * events on the work stream around each phase; rz_get_profile returns the
 * summed milliseconds and the number of launches per phase since the last
 * rz_set_profiling call, in the order forward, residual, adjoint, gradient.
 * No counterpart in the reference (it times only the whole run,
 * rajzel.c:185-187). */
int rz_set_profiling(rz_ctx *ctx, int on);
int rz_get_profile(const rz_ctx *ctx, double *ms, long long *launches);

/* Record rz_forward and rz_gradient, on their first call, as HIP graphs of
 * the whole evaluation and replay those on later calls (on = 1); on = 0
 * drops them.  Applies while every shot fits the HBM budget in one batch,
 * profiling is off and the work stream is not HIP's null stream; otherwise
 * the calls launch as usual.  Results are identical either way.  No
 * counterpart in the reference (its evaluation is a host loop,
 * grad_obj.c:33-52); it serves callers that repeat evaluations of small
 * problems, where launch overhead matters. */
int rz_set_graph(rz_ctx *ctx, int on);

/* Gaussian smoothing of an NY*NX row-major field in place: gauss_filt.c:10-139
 * (precond.c:35-37 calls it for SPATFILTER = 4 on the gradient, fatt.c:179
 * on the search direction), bitwise.  The field is extended by edge replication and
 * convolved with the normalised |fsx| x |fsy| Gaussian, each output summed
 * in the reference's order.  *fsx, *fsy are FILT_SIZE_GRAD and
 * FILT_SIZE_GRAD1 and are updated as the reference updates those globals
 * (an even size is made odd); a size of 0 leaves the field unchanged.
 * field is a host (on_device = 0) or device pointer on HIP device `device`;
 * hip_stream (a hipStream_t, NULL = the null stream) orders the work, and
 * the call returns when it is done.  No context is needed; errors are
 * reported by rz_last_error(NULL). */
int rz_gauss_filt(int device, void *hip_stream, float *field, int nx, int ny, int *fsx, int *fsy, int on_device);

/* Gradient tapers of precond.c:43-53, taper_grad.c:11-299, on an NY*NX
 * row-major field in place, bitwise.  `sws` is taper_grad's argument:
 *   1  vertical taper (SWS_TAPER_GRAD_VERT): columns GRADT1 < i < GRADT4
 *      (1-based) take the Gaussian window of taper_grad.c:62-75, the others 0;
 *   2  horizontal taper (SWS_TAPER_GRAD_HOR): rows below GRADT1 0, rows
 *      GRADT1..GRADT2 the window of taper_grad.c:137-146, deeper rows
 *      pow(j*DH, EXP_TAPER_GRAD_HOR);
 *   3  radial taper at every source and every receiver
 *      (SWS_TAPER_GRAD_SOURCES), as taper_grad.c:199-258, with SRTRADIUS;
 *   4..6  the coefficients of taper.bin (SWS_TAPER_FILE), given here as its
 *      NX*NY floats in file order (x outer, y inner).
 * Window values, powers and disc factors are computed on the host with the C
 * library's exp, pow and cos, as the reference computes them; the device
 * multiplies each cell by its factors in the reference's order.  coeff (host,
 * NY*NX row-major, may be NULL) receives taper_coeff as the reference writes
 * it to taper_coeff_vert.bin, _hor.bin and _file.bin (sws 1, 2, 4..6; left
 * unchanged for sws 3, whose coefficients the reference does not write).
 * field is a host (on_device = 0) or device pointer on HIP device `device`;
 * hip_stream orders the work, and the call returns when it is done.  The
 * sws 3 factors (at most 2^28) are kept between calls with the same geometry.
 * Errors are reported by rz_last_error(NULL). */
typedef struct rz_grad_taper {
    int sws;                          /* 1, 2, 3, or 4..6 (file)                   */
    int gradt1, gradt2, gradt3, gradt4;   /* GRADT1..GRADT4 (sws 1, 2)             */
    float dh;                         /* DH (sws 2, 3)                             */
    float exp_taper_grad_hor;         /* EXP_TAPER_GRAD_HOR (sws 2)                */
    float srtradius;                  /* SRTRADIUS in metres (sws 3)               */
    int nsrc;                         /* sws 3: all sources, srcpos in metres      */
    const float *src_x_m, *src_z_m;
    int ntr;                          /* sws 3: all receivers, 0-based cells       */
    const int *rec_ix, *rec_iy;
    const float *file;                /* sws 4..6: NX*NY floats, x outer (host)    */
} rz_grad_taper;
int rz_taper_grad(int device, void *hip_stream, float *field, int nx, int ny, const rz_grad_taper *taper, float *coeff,
                  int on_device);

/* rz_gauss_filt on a device field without any device allocation.  No
 * counterpart in the reference: it serves callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The weights and the copy of the input live in a
 * static device array of RZ_GAUSS_NOALLOC_FLOATS floats (32 MiB), which
 * each device the call runs on holds until the process ends: weights at
 * its start (at most RZ_GAUSS_NOALLOC_WEIGHTS of them, i.e. windows up to
 * 127 x 127), the field after them, so NX*NY may be at most
 * RZ_GAUSS_NOALLOC_FLOATS - RZ_GAUSS_NOALLOC_WEIGHTS cells (the
 * 1001 x 3001 grid fits).  Larger inputs are rejected with RZ_ERR_INVALID
 * and nothing changed.  Results, size updates and errors are otherwise
 * those of rz_gauss_filt with on_device = 1.  All calls on one device share
 * that device's array: calls that overlap on one device race (the call
 * returns when its work on hip_stream is done). */
#define RZ_GAUSS_NOALLOC_FLOATS (8u << 20)
#define RZ_GAUSS_NOALLOC_WEIGHTS (16u << 10)
int rz_gauss_filt_noalloc(int device, void *hip_stream, float *field, int nx, int ny, int *fsx, int *fsy);

/* smooth_grad.c:10-99 (fd.h:131; precond.c:30-33, SPATFILTER = 3): each
 * cell becomes the median (median2d.c:12-40) of the *fsize x *fsize window
 * around it, the grid's edges replicated, bit for bit.  *fsize is
 * FILT_SIZE_GRAD: 0 leaves the field, an even size is moved one away from 0
 * as the reference does (the new value is returned); |*fsize| <= 129.
 * field: NY x NX row-major, on the host or (on_device = 1) on the device.
 * No context. */
int rz_median_filt(int device, void *hip_stream, float *field, int nx, int ny, int *fsize, int on_device);

/* smooth2.c:9-201 (fd.h:133; precond.c:25-28, SPATFILTER = 2): the
 * tridiagonal smoothing with WD_DAMP (along z) and WD_DAMP1 (along x), bit
 * for bit, on the gradient in the order smooth2.c writes and reads it
 * (x slow, z fast: in[(i-1)*n1 + (j-1)] = v[i][j], i <= n2 = NX,
 * j <= n1 = NY); out receives v[1..n2][1..n1] in the same order (the
 * contents of <JACOBIAN>_smooth2.old).  Host arrays; no context. */
int rz_smooth2(int device, void *hip_stream, const float *in, float *out, int n2, int n1, float wd_damp,
               float wd_damp1);

/* Split shots (on by default; RAJZEL_AMD_SPLIT=0 in the environment turns
 * the default off): a launch with at most one shot per two CUs runs each shot's
 * two paired sweeps in two workgroups on two CUs of one XCD, coupled
 * through L2, instead of one workgroup on one CU.  If a launch finds the two
 * workgroups of a shot on different XCDs, or a wait for the partner times
 * out, the context repeats the work unsplit; failure then reads 1 or 2.  It
 * stays unsplit for 16 runs (evaluations or forwards), then tries split
 * launches again; each further failure doubles that wait (32, 64, ...).
 * rz_set_split resets the wait.  Results are identical either way.  on = 2 also marks the
 * next split launch as failed before it starts (a test of that fallback).
 * No counterpart in the reference (its sweeps run one after another on one
 * core). */
int rz_set_split(rz_ctx *ctx, int on);
int rz_get_split(const rz_ctx *ctx, int *active, int *failure);

/* Shots per CU.  With more shots than half the CUs, a mixed launch runs the
 * n shots that took the most outer iterations last time split (two CUs
 * each) and the others with one workgroup: n = -1 (default) uses every CU
 * the others leave, 0 splits none; with at most one shot per two CUs, 0
 * runs no shot split and any other value all of them.  rz_tune_split times
 * the context's own evaluation (slowness and picks set) with each candidate
 * (n in {0, F/4, F/2, 3F/4, -1}, F = CUs - shots; or split or not) and
 * keeps the fastest, returning it and its time in ms.  Results are the same
 * bits for every setting.  No counterpart in the reference (its shots run
 * one after another per rank). */
int rz_set_split_shots(rz_ctx *ctx, int n);
int rz_get_split_shots(const rz_ctx *ctx);
int rz_tune_split(rz_ctx *ctx, int *split_shots, double *ms);

/* Pick file of one shot into Tobs[1..ntr] (1-based, as the reference's
 * vectors): calc_FA_res.c:28-37, one fscanf("%e") per receiver, where a
 * value that does not parse leaves Tobs[i] as it was.  The reference reads
 * every shot's file on every evaluation and for every grid-search model
 * (about 50 ms for 200 files of 1500 picks); this call keeps each file's
 * parse and parses it again only when the file's device, inode, size or
 * modification time, or ntr, changed (RAJZEL_AMD_PICK_CACHE=0 in the
 * environment: always).  The values written are those a new parse writes.
 * Host only, no context; returns 0, or -1 if the file cannot be opened
 * (Tobs unchanged).  rz_picks_parses counts the parses so far;
 * rz_picks_clear releases the kept values. */
int rz_read_picks(const char *name, float *Tobs, int ntr);
long rz_picks_parses(void);
void rz_picks_clear(void);

/* dotp.c:12-41 (fd.h:80) with sw = 0: the float sum tmp += a[i]*b[i] for
 * i = 0 .. n-1 in that order, computed on device `device` by one workgroup,
 * exactly (every partial sum is the serial one; a NaN result is a NaN, its
 * payload bits may differ).  a and b are host arrays, or device arrays with
 * on_device = 1.  No context. */
int rz_dotp_ordered(int device, const float *a, const float *b, long long n, int on_device, float *out);

/* The l-BFGS two-loop recursion of LBFGS.c:12-226 (fd.h:88) on the
 * device, for the unstrided case (IDX = IDY = 1, NLBFGS_vec = NX*NY): the
 * NLBFGS history slots y, s stay in device memory, every dot product is
 * dotp.c's serial float sum (as rz_dotp_ordered), every update the
 * reference's float operations, so the direction is the reference's bit
 * for bit.  Vectors in "h order" are the reference's vector and file order
 * (x slow, y fast: h = x*NY + y, LBFGS.c:45-58); fields are row-major
 * NY x NX (&A[1][1]).  One handle per NX x NY x NLBFGS; single-threaded. */
typedef struct rz_lbfgs rz_lbfgs;
int rz_lbfgs_create(rz_lbfgs **out, int device, int nx, int ny, int nlbfgs);
void rz_lbfgs_destroy(rz_lbfgs *lb);
const char *rz_lbfgs_error(const rz_lbfgs *lb);
/* zero_LBFGS.c / zero_LBFGS1.c (fd.h:172,175): every slot's y and s to 0 */
int rz_lbfgs_zero(rz_lbfgs *lb);
/* The previous model and gradient (the files <JACOBIAN>_p_vp.old and
 * _p.old, h order) are kept from the last rz_lbfgs_step; after a restart
 * they can be set from the files.  has_previous: 1 if there are some. */
int rz_lbfgs_has_previous(const rz_lbfgs *lb);
int rz_lbfgs_set_previous(rz_lbfgs *lb, const float *grad_h, const float *vp_h);
/* One LBFGS() call.  iter > 1: slot `pointer` (LBFGS_pointer, 1-based)
 * takes y = grad - previous gradient and s = vp - previous model, and hgrad
 * = H^-1 gradm; gamma, rho[0..NLBFGS-1] and alpha[0..NLBFGS-1] get the
 * reference's gamma_LBFGS, rho_LBFGS[1..] and alpha_LBFGS[1..].  iter = 1:
 * hgrad = gradm.  grad_h, vp_h, hgrad_h (any may be NULL) get grad, vp and
 * hgrad in h order, the contents of _p.old, _p_vp.old and _c.old
 * (LBFGS.c:178-216).  grad, gradm, vp, hgrad: row-major fields (host). */
int rz_lbfgs_step(rz_lbfgs *lb, int iter, int pointer, const float *grad, const float *gradm, const float *vp,
                  float *hgrad, float *grad_h, float *vp_h, float *hgrad_h, float *gamma, float *rho,
                  float *alpha);

/* Block until the context's stream is idle. */
int rz_sync(rz_ctx *ctx);

/* Rows per lane in use and bytes of device memory held. */
int rz_rows_per_lane(const rz_ctx *ctx);
/* Compute units of the context's device (one shot per CU fills it). */
int rz_compute_units(const rz_ctx *ctx);
size_t rz_device_bytes(const rz_ctx *ctx);
/* Shots per launch of rz_gradient (all shots, or fewer under a budget). */
int rz_batch_shots(const rz_ctx *ctx);
/* Bytes of device memory one shot's fields take (T in two layouts, lambda
 * in three buffers): the unit an hbm_budget is counted in. */
size_t rz_slot_bytes(const rz_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* RAJZEL_AMD_H */
