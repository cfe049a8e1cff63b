/*
 * rz_fd_surface.c -- the reference's C entry points of the FATT gradient
 * path, on top of the MI355X C ABI (rajzel_amd.h).
 *
 * Linked into the reference program in place of eikonal.c, sweep.c,
 * eikonal_adj.c, sweep_adj.c, ass_grad.c, grad_obj.c, forward.c,
 * gauss_filt.c, smooth2.c, smooth_grad.c, taper_grad.c, grid_search.c,
 * LBFGS.c, zero_LBFGS.c, zero_LBFGS1.c and calc_mat_change_wolfe.c, it keeps
 * their fd.h signatures, their globals and their side effects, so that
 * fatt.c, PCG.c, wolfels.c, parabolicls.c, step_length_est.c and rajzel.c
 * call it unchanged:
 *
 *   grad_obj    (fd.h:72, grad_obj.c:10-79)   every shot of this rank in one
 *               batched evaluation on the GPU, with the line search's
 *               predicted next trial in the same launch where CUs are idle
 *               (calc_mat_change_wolfe, fd.h:44, reports the trials); picks
 *               read from the
 *               <DATA_DIR>_<shot>.dat files as calc_FA_res.c:28-37 does
 *               (each file parsed again only when it changed, rz_read_picks);
 *               MPI sum, frame copy and precond() as grad_obj.c:57-73
 *   forward     (fd.h:68, forward.c:10-26)    all shots of this rank in one
 *               launch; SNAP and SEISMO outputs per shot as eikonal.c:57-72
 *   eikonal     (fd.h:60, eikonal.c:10-76)    one shot
 *   eikonal_adj (fd.h:62, eikonal_adj.c:10-59) one shot
 *   ass_grad    (fd.h:36, ass_grad.c:11-49)   one shot
 *   grid_search (fd.h:74, grid_search.c:10-96)  every model of the 1-D
 *               gradient grid and every shot of this rank in batched
 *               forward launches; residuals, misfits and the table through
 *               the reference's calc_FA_res / write_gridsearch
 *   gauss_filt  (fd.h:70, gauss_filt.c:10-139)   the Gaussian smoothing of
 *               gradient (precond.c:35-37, SPATFILTER = 4) and of the
 *               search direction (fatt.c:179), on the GPU; linked in
 *               place of gauss_filt.c
 *   smooth2, smooth_grad (fd.h:133,131)  precond.c's SPATFILTER 2 and 3
 *               smoothings on the GPU, same files and messages
 *   taper_grad  (fd.h:149, taper_grad.c:11-299)  the gradient tapers of
 *               precond.c:43-53, multiplied on the GPU; log lines, taper.bin
 *               and taper_coeff_*.bin files as the reference's
 *   LBFGS       (fd.h:88, LBFGS.c:12-226)  the two-loop recursion on the
 *               GPU with the reference's arithmetic (unstrided; strided
 *               on the host), each history file read or written in one
 *               call; zero_LBFGS and zero_LBFGS1 (fd.h:172,175) with it
 *
 * sweep() and sweep_adj() (fd.h:145,147) have no callers outside
 * eikonal.c and eikonal_adj.c; their work happens inside the kernels.
 *
 * Arrays are the reference's NR matrices: &A[1][1] is a dense row-major
 * NY x NX block (util.c:177-204), vectors are 1-based.  A GPU failure ends
 * the run through the reference's err() (util.c:7-17), as every other
 * error of the program does.  The device is RAJZEL_AMD_DEVICE (default 0).
 */
#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rajzel_amd.h"

/* globals of the reference (globvar.h:12-29,38,69,96) */
extern float DH, TIME, TTNORM, EPS_ADJ;
extern int NX, NY, SEISMO, SNAP, MYID, INFO, NSHOT1, NSHOT2, SWS_TAPER_CIRCULAR_PER_SHOT;
extern int FILT_SIZE_GRAD, FILT_SIZE_GRAD1, IDX, IDY, FILTSIZE, GRADT1, GRADT2, GRADT3, GRADT4;
extern float SRTRADIUS, EXP_TAPER_GRAD_HOR;
extern char JACOBIAN[], MFILE[];
extern FILE *FP;
extern float VP0_1, VP0_2, DVP0, GRAD0_1, GRAD0_2, DGRAD0;
extern char SNAP_FILE[], DATA_DIR[];
extern int LINESEARCH;
extern float SCALEFAC, VPUPPERLIM, VPLOWERLIM;
extern char INV_MODELFILE[];

/* routines of the reference that stay in use (fd.h:52,119,143,166,170,180;
 * util.c's matrix helpers) */
float **matrix(int nrl, int nrh, int ncl, int nch);
void free_matrix(float **m, int nrl, int nrh, int ncl, int nch);
void cp_grad_frame(float **A);
void precond(float **grad, int nsrc, float **srcpos, int **recpos, int ntr, int iter);
void sum_grad_MPI(float **grad);
void write_picks(float *Tmod, int ntr, int ishot);
void writemod(char modfile[], float **array, int format);
void err(char err_text[]);
double calc_FA_res(float **TT, float *Tmod, float *Tobs, float *Tres, int ntr, int ishot, int **recpos);
void calc_S(float **Vp, float **S);
void model_gridsearch(float **Vp, float vp0, float grad0);
void write_gridsearch(float L2, float vp0, float grad0, int count);
float dotp(float *vec1, float *vec2, int n1, int n2, int sw);
void writedsk(FILE *fp_out, float amp, int format);
void exchange_grad_MPI(float **grad);

#define RZS_PATH 74                 /* STRING_SIZE of fd.h:25 */

/* One context per distinct geometry (grid, constants, shot range,
 * receivers); a new geometry replaces it. */
static struct {
    rz_ctx *ctx;
    rz_model_desc d;
    int *src_ix, *src_iy, *rec_ix, *rec_iy;
    float *S_host;                  /* slowness last uploaded (NX*NY) */
    int s_valid;
} g_rzs;

static char g_msg[512];

/* The next line-search trial, evaluated ahead (see grad_obj below): a
 * two-model context of the same geometry (model 0 = the trial asked for,
 * model 1 = the predicted next trial), the predicted model's slowness and
 * the picks both used. */
static struct {
    rz_ctx *ctx;
    int nmodels;                    /* 2 or 3: the trial and up to two predicted next ones */
    float *S[2], *picks;            /* slowness of models 1 and 2 (host) */
    size_t npicks;
    int iter, nvalid;               /* models 1 .. nvalid hold results */
    long calls, served, ahead;
} g_spec;

static void rzs_fail(const char *what)
{
    snprintf(g_msg, sizeof g_msg, " rajzel_amd %s: %s", what, rz_last_error(g_rzs.ctx));
    err(g_msg);
}

static void rzs_check(int rc, const char *what)
{
    if (rc != RZ_OK) rzs_fail(what);
}

static void *rzs_alloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) err(" rajzel_amd: host allocation failed");
    return p;
}

/* an environment switch: unset -> dflt, else its integer value != 0 */
static int getenv_flag(const char *name, int dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) != 0 : dflt;
}

static void rzs_spec_release(void)
{
    rz_destroy(g_spec.ctx);
    free(g_spec.S[0]);
    free(g_spec.S[1]);
    free(g_spec.picks);
    g_spec.ctx = NULL;
    g_spec.S[0] = g_spec.S[1] = g_spec.picks = NULL;
    g_spec.nvalid = 0;
}

static void rzs_release(void)
{
    rzs_spec_release();
    rz_destroy(g_rzs.ctx);
    free(g_rzs.src_ix); free(g_rzs.src_iy); free(g_rzs.rec_ix); free(g_rzs.rec_iy); free(g_rzs.S_host);
    memset(&g_rzs, 0, sizeof g_rzs);
}

/* Context for shots [s1, s2) (1-based) and the receivers recpos[1..2][1..ntr]. */
static rz_ctx *rzs_context(float **srcpos, int s1, int s2, int **recpos, int ntr)
{
    const int ns = s2 > s1 ? s2 - s1 : 0;
    int same = g_rzs.ctx && g_rzs.d.nx == NX && g_rzs.d.ny == NY && g_rzs.d.dh == DH && g_rzs.d.time == TIME &&
               g_rzs.d.ttnorm == TTNORM && g_rzs.d.eps_adj == EPS_ADJ && g_rzs.d.nshots == ns &&
               g_rzs.d.ntr == ntr;
    int *sx = rzs_alloc(ns * sizeof(int)), *sy = rzs_alloc(ns * sizeof(int));
    int *rx = rzs_alloc(ntr * sizeof(int)), *ry = rzs_alloc(ntr * sizeof(int));
    for (int s = 0; s < ns; ++s) {
        sx[s] = rz_grid_index(srcpos[1][s1 + s], DH);      /* eikonal.c:31-32 */
        sy[s] = rz_grid_index(srcpos[2][s1 + s], DH);
    }
    for (int i = 0; i < ntr; ++i) {
        rx[i] = recpos[1][i + 1] - 1;
        ry[i] = recpos[2][i + 1] - 1;
    }
    if (same)
        same = !memcmp(sx, g_rzs.src_ix, ns * sizeof(int)) && !memcmp(sy, g_rzs.src_iy, ns * sizeof(int)) &&
               !memcmp(rx, g_rzs.rec_ix, ntr * sizeof(int)) && !memcmp(ry, g_rzs.rec_iy, ntr * sizeof(int));
    if (same) {
        free(sx); free(sy); free(rx); free(ry);
        return g_rzs.ctx;
    }
    rzs_release();
    rz_model_desc d;
    memset(&d, 0, sizeof d);
    d.nx = NX; d.ny = NY; d.dh = DH; d.time = TIME; d.ttnorm = TTNORM; d.eps_adj = EPS_ADJ;
    d.max_iter = 0;
    d.nshots = ns; d.src_ix = sx; d.src_iy = sy;
    d.ntr = ntr; d.rec_ix = rx; d.rec_iy = ry;
    const char *dev = getenv("RAJZEL_AMD_DEVICE");
    /* RAJZEL_AMD_HBM_MIB: device memory for the shots' fields in MiB; without
     * it every shot is resident when the device has room, else rz_create
     * takes what fits.  Under a budget the shots run in batches. */
    const char *mib = getenv("RAJZEL_AMD_HBM_MIB");
    const size_t budget = (mib && *mib) ? (size_t)strtoull(mib, NULL, 10) << 20 : 0;
    rz_ctx *c = NULL;
    if (rz_create(&c, dev ? atoi(dev) : 0, 1, &d, 0, budget) != RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd context: %s", rz_last_error(NULL));
        err(g_msg);
    }
    g_rzs.ctx = c;
    g_rzs.d = d;
    g_rzs.src_ix = sx; g_rzs.src_iy = sy; g_rzs.rec_ix = rx; g_rzs.rec_iy = ry;
    g_rzs.S_host = rzs_alloc((size_t)NX * NY * sizeof(float));
    g_rzs.s_valid = 0;
    return c;
}

/* Upload S unless it equals the slowness already on the device. */
static void rzs_slowness(float **S)
{
    const size_t n = (size_t)NX * NY;
    if (g_rzs.s_valid && !memcmp(g_rzs.S_host, &S[1][1], n * sizeof(float))) return;
    rzs_check(rz_set_slowness(g_rzs.ctx, 0, &S[1][1], 0), "slowness");
    memcpy(g_rzs.S_host, &S[1][1], n * sizeof(float));
    g_rzs.s_valid = 1;
}

/* Single-shot calls use a context of their own grid and receivers; its
 * shot list is that of this rank. */
static rz_ctx *rzs_single(float **srcpos, int **recpos, int ntr)
{
    if (g_rzs.ctx && g_rzs.d.nx == NX && g_rzs.d.ny == NY && g_rzs.d.ntr == ntr) {
        int same = 1;
        for (int i = 0; i < ntr && same; ++i)
            same = g_rzs.rec_ix[i] == recpos[1][i + 1] - 1 && g_rzs.rec_iy[i] == recpos[2][i + 1] - 1;
        if (same && g_rzs.d.dh == DH && g_rzs.d.time == TIME && g_rzs.d.ttnorm == TTNORM &&
            g_rzs.d.eps_adj == EPS_ADJ)
            return g_rzs.ctx;
    }
    return rzs_context(srcpos, NSHOT1, NSHOT1, recpos, ntr);
}

void eikonal(float **S, float **TT, float *Tmod, float **srcpos, int ishot, int nshots, int **recpos, int ntr)
{
    char filename[RZS_PATH];
    int iters = 0;
    if (MYID == 0 && INFO == 1)
        printf("\n *****  Solve Eikonal equation for shot %d of %d (MPI process %d, MI355X) ***** \n", ishot,
               nshots, MYID);
    rz_ctx *c = rzs_single(srcpos, recpos, ntr);
    rzs_slowness(S);
    rzs_check(rz_eikonal_shot(c, 0, rz_grid_index(srcpos[1][ishot], DH), rz_grid_index(srcpos[2][ishot], DH),
                              &TT[1][1], ntr > 0 ? &Tmod[1] : NULL, &iters),
              "eikonal");
    if (SNAP == 1) {                                        /* eikonal.c:57-60 */
        snprintf(filename, sizeof filename, "%s_shot_%d.tt", SNAP_FILE, ishot);
        writemod(filename, TT, 3);
    }
    if (SEISMO == 1) write_picks(Tmod, ntr, ishot);         /* eikonal.c:70-72 */
}

void eikonal_adj(float **TT, float *Tres, float **lam, int **recpos, int ntr, int ishot, int **recflag)
{
    int iters = 0;
    (void)ishot;
    (void)recflag;   /* init_recflag(recpos) of grad_obj.c:31; the context holds the same map */
    if (MYID == 0 && INFO == 1) printf("\n *****  Solve adjoint Eikonal equation (MI355X) ***** \n");
    rz_ctx *c = rzs_single(NULL, recpos, ntr);
    rzs_check(rz_eikonal_adj_shot(c, 0, &TT[1][1], ntr > 0 ? &Tres[1] : NULL, &lam[1][1], &iters), "eikonal_adj");
}

/* The per-shot gradient taper of ass_grad.c:31-37 on context c, whose shot
 * s is the reference's shot NSHOT1 + s: SWS_TAPER_CIRCULAR_PER_SHOT 1
 * (taper_grad_shot.c, SRTRADIUS) or 2 (taper_grad_shot1.c, FILTSIZE) over the
 * nshots sources of srcpos; other values apply none, as in ass_grad.c.  The
 * library keeps its tables while the settings stay the same. */
static int rzs_taper_at(rz_ctx *c, int model, float **srcpos, int nshots)
{
    const int mode = SWS_TAPER_CIRCULAR_PER_SHOT == 1 || SWS_TAPER_CIRCULAR_PER_SHOT == 2
                         ? SWS_TAPER_CIRCULAR_PER_SHOT : 0;
    if (mode == 0 || nshots <= 0) {
        rzs_check(rz_set_shot_taper(c, model, 0, 0.f, 0, 0, NULL, NULL, 0), "per-shot taper");
        return mode;
    }
    float *x = rzs_alloc((size_t)nshots * sizeof(float)), *z = rzs_alloc((size_t)nshots * sizeof(float));
    for (int n = 0; n < nshots; ++n) {
        x[n] = srcpos[1][n + 1];
        z[n] = srcpos[2][n + 1];
    }
    rzs_check(rz_set_shot_taper(c, model, mode, SRTRADIUS, FILTSIZE, nshots, x, z, NSHOT1 - 1), "per-shot taper");
    free(x);
    free(z);
    return mode;
}

static int rzs_taper(rz_ctx *c, float **srcpos, int nshots)
{
    return rzs_taper_at(c, 0, srcpos, nshots);
}

void ass_grad(float **grad, float **grad_shot, float **S, float **lam, float **srcpos, int nshots, int **recpos,
              int ntr, int ishot)
{
    rzs_single(srcpos, recpos, ntr);
    rzs_slowness(S);
    if (rzs_taper(g_rzs.ctx, srcpos, nshots) != 0)          /* ass_grad.c:31-37 */
        rzs_check(rz_ass_grad_shot_at(g_rzs.ctx, 0, ishot - 1, &lam[1][1], &grad_shot[1][1], &grad[1][1]),
                  "ass_grad");
    else
        rzs_check(rz_ass_grad_shot(g_rzs.ctx, 0, &lam[1][1], &grad_shot[1][1], &grad[1][1]), "ass_grad");
}

/* The last shot's fields and receiver data, which the reference leaves in
 * the caller's arrays. */
static void rzs_last_shot_of(rz_ctx *c, int model, int ns, int ntr, float **TT, float **lam, float *Tmod,
                             float *tm_all)
{
    const int last = ns - 1;
    if (TT) rzs_check(rz_get_field(c, model, last, 0, &TT[1][1], 0), "traveltime field");
    if (lam) rzs_check(rz_get_field(c, model, last, 1, &lam[1][1], 0), "adjoint field");
    if (Tmod && ntr > 0) memcpy(&Tmod[1], tm_all + (size_t)last * ntr, ntr * sizeof(float));
}

static void rzs_last_shot(int ns, int ntr, float **TT, float **lam, float *Tmod, float *tm_all)
{
    rzs_last_shot_of(g_rzs.ctx, 0, ns, ntr, TT, lam, Tmod, tm_all);
}

void forward(float **S, float **TT, float *Tmod, float **srcpos, int nshots, int **recpos, int ntr)
{
    char filename[RZS_PATH];
    const int ns = NSHOT2 > NSHOT1 ? NSHOT2 - NSHOT1 : 0;
    (void)nshots;
    if (ns == 0) return;
    rz_ctx *c = rzs_context(srcpos, NSHOT1, NSHOT2, recpos, ntr);
    rzs_slowness(S);
    rzs_check(rz_forward(c), "forward");
    float *tm = rzs_alloc((size_t)ns * (ntr > 0 ? ntr : 1) * sizeof(float));
    if (ntr > 0) rzs_check(rz_get_tmod(c, 0, tm, 0), "receiver times");
    /* under an HBM budget only the last batch's fields stay resident: the
     * snapshots of the others are solved again one shot at a time (the
     * same bits), which leaves the last shot's field in TT */
    const int resident = rz_batch_shots(c) >= ns;
    for (int s = 0; s < ns; ++s) {
        const int ishot = NSHOT1 + s;
        if (ntr > 0) memcpy(&Tmod[1], tm + (size_t)s * ntr, ntr * sizeof(float));
        if (SNAP == 1) {
            if (resident)
                rzs_check(rz_get_field(c, 0, s, 0, &TT[1][1], 0), "traveltime field");
            else
                rzs_check(rz_eikonal_shot(c, 0, g_rzs.src_ix[s], g_rzs.src_iy[s], &TT[1][1], NULL, NULL),
                          "traveltime field");
            snprintf(filename, sizeof filename, "%s_shot_%d.tt", SNAP_FILE, ishot);
            writemod(filename, TT, 3);
        }
        if (SEISMO == 1) write_picks(Tmod, ntr, ishot);
    }
    rzs_last_shot(ns, ntr, (SNAP == 1 && !resident) ? NULL : TT, NULL, Tmod, tm);
    free(tm);
}

/* ---------------------------------------------------------------------------
 * Line-search trials evaluated ahead: the batched line search.
 *
 * wolfels.c:55-101 and parabolicls.c:56-61 evaluate trial models one after
 * another, each grad_obj waiting for the decision on the previous one.  The
 * next trial is one of two, both known before the decision: the parabolic
 * search (LINESEARCH = 2, the stock one) continues with alpha +
 * alpha/SCALEFAC when the trial lowered the misfit or has not yet bracketed
 * the minimum, and with alpha/SCALEFAC when it did not lower it
 * (parabolicls.c:64-96); the Wolfe search (LINESEARCH = 1) sets mu or nu to
 * the trial's alpha and continues with SCALEFAC*alpha or (nu + mu)/SCALEFAC
 * (wolfels.c:55-58,90-97), whose mu and nu the shim follows from the
 * trials it sees.  When a trial's shots leave CUs idle even with each shot
 * split over two CUs, grad_obj evaluates the trial and the predicted next
 * trials that fit so (the likelier first) in one launch; when the search then asks for exactly one of those models
 * (the same slowness bits, picks and iteration), grad_obj returns its
 * results without evaluating again.  A wrong prediction only costs the idle
 * CUs (with fewer free CUs a second model costs about what it saves, and
 * the launch is left alone).  All models start from the caller's residual
 * carry, so trials are
 * evaluated ahead only when no pick is muted (the carry then does not enter
 * any result).
 *
 * The trials come from calc_mat_change_wolfe (calc_mat_change_wolfe.c:10-58,
 * fd.h:44), which the shim provides with the reference's operations, so
 * that it knows the direction, the model and the step of each trial.
 * RAJZEL_AMD_SPECULATE=0 turns this off, =2 evaluates ahead whatever the
 * CU count; RAJZEL_AMD_SPECULATE_STATS=1 prints the count at exit.
 * --------------------------------------------------------------------------- */
static struct {
    float **Hgrad, **Vp, **Vpnp1;     /* of the last trial (itest = 1) */
    float alpha;
    int fresh;                        /* a trial not yet evaluated */
    int open, known;                  /* a search is open; mu and nu follow it */
    float mu, nu, cand[2];            /* wolfels.c's mu, nu; the last trial's two successors */
} g_trial;

/* calc_mat_change_wolfe.c:29-47 on contiguous blocks: vp + eps*hg, kept
 * where it leaves [VPLOWERLIM, VPUPPERLIM] or is negative */
static void rzs_step_model(const float *vp, const float *hg, float eps, float *out, size_t n)
{
    for (size_t k = 0; k < n; ++k) {
        float v = vp[k] + eps * hg[k];
        if (v < VPLOWERLIM) v = vp[k];
        if (v > VPUPPERLIM) v = vp[k];
        if (v < 0.0) v = vp[k];
        out[k] = v;
    }
}

void calc_mat_change_wolfe(float **Hgrad, float **Vp, float **Vpnp1, float eps_scale, int itest)
{
    const size_t n = (size_t)NX * NY;
    rzs_step_model(&Vp[1][1], &Hgrad[1][1], eps_scale, &Vpnp1[1][1], n);
    if (itest == 0) {                                       /* the accepted step: fatt.c:210 */
        char modfile[RZS_PATH];
        memcpy(&Vp[1][1], &Vpnp1[1][1], n * sizeof(float));
        snprintf(modfile, sizeof modfile, "%s_vp.bin", INV_MODELFILE);
        writemod(modfile, Vpnp1, 3);
        g_trial.fresh = 0;
        g_trial.open = 0;
        return;
    }
    if (!g_trial.open) {                                    /* a search's first trial: wolfels.c:44-45 */
        g_trial.open = 1;
        g_trial.known = 1;
        g_trial.mu = 0.0f;
        g_trial.nu = 1e30f;
    } else if (g_trial.known) {                             /* which successor the search took */
        if (!memcmp(&eps_scale, &g_trial.cand[0], sizeof(float))) g_trial.mu = g_trial.alpha;
        else if (!memcmp(&eps_scale, &g_trial.cand[1], sizeof(float))) g_trial.nu = g_trial.alpha;
        else g_trial.known = 0;
    }
    /* wolfels.c:55-58 after mu = alpha (too short) and after nu = alpha (too long) */
    const float mu = g_trial.mu, nu = g_trial.nu;
    g_trial.cand[0] = nu < 1e30 ? (nu + eps_scale) / SCALEFAC : SCALEFAC * eps_scale;
    g_trial.cand[1] = (eps_scale + mu) / SCALEFAC;
    g_trial.Hgrad = Hgrad;
    g_trial.Vp = Vp;
    g_trial.Vpnp1 = Vpnp1;
    g_trial.alpha = eps_scale;
    g_trial.fresh = 1;
}

static void rzs_spec_stats(void)
{
    fprintf(stderr, "rajzel_amd line search: %ld grad_obj calls, %ld evaluated with the next trial ahead, "
                    "%ld served from it\n", g_spec.calls, g_spec.ahead, g_spec.served);
}

/* The current search's possible next trials, the likelier first; returns
 * how many (0 if the search is not one the shim can follow). */
static int rzs_next_alphas(float alpha[2])
{
    if (LINESEARCH == 2) {                                  /* parabolicls.c:76-95 */
        alpha[0] = g_trial.alpha + (g_trial.alpha / SCALEFAC);
        alpha[1] = g_trial.alpha / SCALEFAC;
        return 2;
    }
    if (LINESEARCH == 1 && g_trial.known) {                 /* wolfels.c:55-58 */
        alpha[0] = g_trial.cand[0];
        alpha[1] = g_trial.cand[1];
        return 2;
    }
    return 0;
}

/* S equals calc_S.c's 1.0/Vpnp1 of the last trial, bit for bit */
static int rzs_is_trial(float **S)
{
    const size_t n = (size_t)NX * NY;
    const float *v = &g_trial.Vpnp1[1][1], *s = &S[1][1];
    for (size_t k = 0; k < n; ++k) {
        const float q = (float)(1.0 / v[k]);
        if (memcmp(&q, s + k, sizeof q) != 0) return 0;
    }
    return 1;
}

static int rzs_spec_mode(void)
{
    const char *v = getenv("RAJZEL_AMD_SPECULATE");
    return (v && *v) ? atoi(v) : 1;
}

/* The results of (c, model) into the caller's arrays: gradient, this rank's
 * misfit sum, the residual carry and the last shot's fields. */
static double rzs_collect(rz_ctx *c, int model, int ns, int ntr, float **grad, float **TT, float **lam,
                          float *Tmod, float *Tres)
{
    double half = 0.0;
    rzs_check(rz_get_grad(c, model, &grad[1][1], 0), "gradient");
    rzs_check(rz_get_misfit(c, model, &half, NULL), "misfit");
    if (ntr > 0) {
        float *tm = rzs_alloc((size_t)ns * ntr * sizeof(float));
        rzs_check(rz_get_tmod(c, model, tm, 0), "receiver times");
        rzs_check(rz_get_tres_carry(c, model, &Tres[1]), "residual carry");
        rzs_last_shot_of(c, model, ns, ntr, TT, lam, Tmod, tm);
        free(tm);
    } else {
        rzs_last_shot_of(c, model, ns, 0, TT, lam, NULL, NULL);
    }
    return 2.0 * half;                                      /* this rank's sum, grad_obj.c:42 */
}

double grad_obj(float **grad, float **S, float **TT, float **lam, float *Tmod, float *Tobs, float *Tres,
                float **srcpos, int nshots, int **recpos, int ntr, int iter)
{
    const int ns = NSHOT2 > NSHOT1 ? NSHOT2 - NSHOT1 : 0;
    const size_t ncell = (size_t)NX * NY;
    double L2 = 0.0, L2sum = 0.0;
    SEISMO = 0;                                             /* grad_obj.c:24 */
    const int trial = g_trial.fresh;
    g_trial.fresh = 0;
    if (ns > 0) {
        rz_ctx *c = rzs_context(srcpos, NSHOT1, NSHOT2, recpos, ntr);
        const size_t np = (size_t)ns * (ntr > 0 ? ntr : 0);
        float *tobs = NULL;
        int muted = 0;
        if (ntr > 0) {
            /* calc_FA_res.c:28-37: every evaluation reads every shot's pick
             * file with fscanf("%e") into Tobs, in shot order (a value that
             * does not parse keeps what Tobs held) */
            tobs = rzs_alloc(np * sizeof(float));
            char name[RZS_PATH];
            for (int s = 0; s < ns; ++s) {
                snprintf(name, sizeof name, "%s_%i.dat", DATA_DIR, NSHOT1 + s);
                if (rz_read_picks(name, Tobs, ntr) != 0) err(" picks_?.dat could not be opened !");
                memcpy(tobs + (size_t)s * ntr, &Tobs[1], ntr * sizeof(float));
            }
            for (size_t k = 0; k < np && !muted; ++k) muted = !(tobs[k] >= 0.f);
        }
        const int mode = rzs_spec_mode();
        const int cus = rz_compute_units(c);
        /* predicted trials that fit beside this one with every shot still split over two CUs (a
         * lone shot already uses a second CU, so only CUs beyond that are free) */
        int fit = mode == 2 ? 2 : mode == 1 ? cus / (2 * ns) - 1 : 0;
        fit = fit < 0 ? 0 : fit > 2 ? 2 : fit;
        float alphas[2];
        int served = -1;
        if (g_spec.calls++ == 0 && MYID == 0 && getenv_flag("RAJZEL_AMD_SPECULATE_STATS", 0)) atexit(rzs_spec_stats);
        if (mode && np > 0 && g_spec.nvalid > 0 && g_spec.iter == iter && g_spec.npicks == np &&
            !memcmp(g_spec.picks, tobs, np * sizeof(float))) {
            for (int m = 0; m < g_spec.nvalid && served < 0; ++m)
                if (!memcmp(g_spec.S[m], &S[1][1], ncell * sizeof(float))) served = m + 1;
        }
        int na = 0;
        if (served > 0) {
            /* a trial evaluated ahead with the previous one */
            g_spec.nvalid = 0;
            ++g_spec.served;
            L2 = rzs_collect(g_spec.ctx, served, ns, ntr, grad, TT, lam, Tmod, Tres);
        } else if (mode && fit > 0 && trial && ntr > 0 && !muted && (na = rzs_next_alphas(alphas)) > 0 &&
                   rzs_is_trial(S)) {
            /* this trial and its predicted successors in one launch */
            if (na > fit) na = fit;
            if (g_spec.ctx == NULL) {
                rz_model_desc d3[3] = {g_rzs.d, g_rzs.d, g_rzs.d};
                const char *dev = getenv("RAJZEL_AMD_DEVICE");
                g_spec.nmodels = 1 + fit;
                if (rz_create(&g_spec.ctx, dev ? atoi(dev) : 0, g_spec.nmodels, d3, 0, 0) != RZ_OK) {
                    snprintf(g_msg, sizeof g_msg, " rajzel_amd context: %s", rz_last_error(NULL));
                    err(g_msg);
                }
                g_spec.S[0] = rzs_alloc(ncell * sizeof(float));
                g_spec.S[1] = rzs_alloc(ncell * sizeof(float));
            }
            if (na > g_spec.nmodels - 1) na = g_spec.nmodels - 1;
            rz_ctx *sc = g_spec.ctx;
            float *v1 = rzs_alloc(ncell * sizeof(float));
            for (int m = 0; m < g_spec.nmodels - 1; ++m) {
                /* the successors' models (calc_mat_change_wolfe.c, calc_S.c:20-22); a
                 * model without a successor repeats the trial's */
                if (m < na) {
                    rzs_step_model(&g_trial.Vp[1][1], &g_trial.Hgrad[1][1], alphas[m], v1, ncell);
                    for (size_t k = 0; k < ncell; ++k) g_spec.S[m][k] = (float)(1.0 / v1[k]);
                } else {
                    memcpy(g_spec.S[m], &S[1][1], ncell * sizeof(float));
                }
            }
            free(v1);
            rzs_check(rz_set_slowness(sc, 0, &S[1][1], 0), "slowness");
            for (int m = 1; m < g_spec.nmodels; ++m) rzs_check(rz_set_slowness(sc, m, g_spec.S[m - 1], 0), "slowness");
            for (int m = 0; m < g_spec.nmodels; ++m) {
                rzs_taper_at(sc, m, srcpos, nshots);        /* ass_grad.c:31-37 */
                rzs_check(rz_set_picks(sc, m, tobs, 0), "picks");
                rzs_check(rz_set_tres_carry(sc, m, &Tres[1]), "residual carry");
            }
            rzs_check(rz_gradient(sc), "gradient");
            free(g_spec.picks);
            g_spec.picks = rzs_alloc(np * sizeof(float));
            memcpy(g_spec.picks, tobs, np * sizeof(float));
            g_spec.npicks = np;
            g_spec.iter = iter;
            g_spec.nvalid = na;
            ++g_spec.ahead;
            L2 = rzs_collect(sc, 0, ns, ntr, grad, TT, lam, Tmod, Tres);
        } else {
            rzs_slowness(S);
            rzs_taper(c, srcpos, nshots);                   /* ass_grad.c:31-37 */
            if (ntr > 0) {
                rzs_check(rz_set_picks(c, 0, tobs, 0), "picks");
                rzs_check(rz_set_tres_carry(c, 0, &Tres[1]), "residual carry");
            }
            rzs_check(rz_gradient(c), "gradient");
            L2 = rzs_collect(c, 0, ns, ntr, grad, TT, lam, Tmod, Tres);
        }
        free(tobs);
    } else {
        memset(&grad[1][1], 0, ncell * sizeof(float));     /* init_grad.c */
    }
    MPI_Barrier(MPI_COMM_WORLD);                            /* grad_obj.c:58-61 */
    MPI_Allreduce(&L2, &L2sum, 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    L2 = 0.5 * L2sum;
    sum_grad_MPI(grad);                                     /* grad_obj.c:68-73 */
    cp_grad_frame(grad);
    precond(grad, nshots, srcpos, recpos, ntr, iter);
    return L2;
}

static size_t rzs_strided(void);
static void rzs_gather(float **A, float *buf);

/* smooth_grad.c:10-99 (fd.h:131; precond.c:30-33, SPATFILTER = 3): the
 * median filter over FILT_SIZE_GRAD x FILT_SIZE_GRAD windows on the GPU
 * (rz_median_filt), FILT_SIZE_GRAD made odd and its half printed as there. */
void smooth_grad(float **waveconv)
{
    if (FILT_SIZE_GRAD == 0) return;
    if (!(FILT_SIZE_GRAD % 2)) FILT_SIZE_GRAD += FILT_SIZE_GRAD > 0 ? 1 : -1;
    printf("\n hfs: %d \n", abs(FILT_SIZE_GRAD) / 2);
    int fs = FILT_SIZE_GRAD;
    const char *dev = getenv("RAJZEL_AMD_DEVICE");
    if (rz_median_filt(dev ? atoi(dev) : 0, NULL, &waveconv[1][1], NX, NY, &fs, 0) != RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd smooth_grad: %s", rz_last_error(NULL));
        err(g_msg);
    }
}

/* smooth2.c:9-201 (fd.h:133; precond.c:25-28, SPATFILTER = 2): the same
 * files and messages (<JACOBIAN>_wavenumber.old written by every rank with
 * strides IDX/IDY, read back and smoothed on rank 0, <JACOBIAN>_smooth2.old
 * written there and read back by every rank), each file in one call, the
 * tridiagonal smoothing on the GPU (rz_smooth2). */
void smooth2(float **grad)
{
    extern int NXG, NYG, POS[];
    extern float WD_DAMP, WD_DAMP1;
    char jac[RZS_PATH + 32];
    const size_t ns = rzs_strided(), n = (size_t)NXG * NYG;
    float *buf = rzs_alloc((ns > n ? ns : n) * sizeof(float));
    rzs_gather(grad, buf);                                  /* smooth2.c:55-65 */
    snprintf(jac, sizeof jac, "%s_wavenumber.old", JACOBIAN);
    FILE *fp = fopen(jac, "wb");
    if (fp) {
        fwrite(buf, sizeof(float), ns, fp);
        fclose(fp);
    }
    if (MYID == 0) {
        printf("\n Smooth2 is applied to gradient (written by PE %d)\n", MYID);
        fp = fopen(jac, "rb");
        if (fp == NULL) err(" Could not open gradient file ! ");
        float *in = rzs_alloc(n * sizeof(float)), *out = rzs_alloc(n * sizeof(float));
        const size_t got = fread(in, sizeof(float), n, fp);   /* a short file: fread leaves the last value */
        fclose(fp);
        for (size_t t = got; t < n; ++t) in[t] = got ? in[got - 1] : 0.f;
        const char *dev = getenv("RAJZEL_AMD_DEVICE");
        if (rz_smooth2(dev ? atoi(dev) : 0, NULL, in, out, NXG, NYG, WD_DAMP, WD_DAMP1) != RZ_OK) {
            snprintf(g_msg, sizeof g_msg, " rajzel_amd smooth2: %s", rz_last_error(NULL));
            err(g_msg);
        }
        snprintf(jac, sizeof jac, "%s_smooth2.old", JACOBIAN);
        fp = fopen(jac, "wb");
        if (fp) {
            fwrite(out, sizeof(float), n, fp);
            fclose(fp);
        }
        free(in);
        free(out);
    }
    MPI_Barrier(MPI_COMM_WORLD);
    snprintf(jac, sizeof jac, "%s_smooth2.old", JACOBIAN);  /* smooth2.c:176-199 */
    fp = fopen(jac, "rb");
    const size_t got = fp ? fread(buf, sizeof(float), n, fp) : 0;
    if (fp) fclose(fp);
    size_t t = 0;
    for (int i = 1; i <= NXG; i++)
        for (int j = 1; j <= NYG; j++, t++) {
            const float g = t < got ? buf[t] : (got ? buf[got - 1] : 0.f);
            if (POS[1] == (i - 1) / NX && POS[2] == (j - 1) / NY) grad[j - POS[2] * NY][i - POS[1] * NX] = g;
        }
    free(buf);
}

/* taper_grad.c:11-299 (precond.c:43-53) with the multiplication on the GPU:
 * the reference's log lines, its read of taper.bin (sws 4..6) and its
 * taper_coeff_vert/hor/file.bin are kept; sws 3 writes no file, as there. */
void taper_grad(float **waveconv, float **srcpos, int nshots, int **recpos, int ntr, int sws)
{
    const int msg = sws == 1 || sws == 2 || (sws >= 4 && sws <= 6);
    if (!(sws >= 1 && sws <= 6)) return;
    if (sws == 2 && MYID == 0) {                            /* taper_grad.c:124-127 */
        printf("%d \t %d \t %d \t %d \n", GRADT1, GRADT2, GRADT3, GRADT4);
        printf("%d \t %d \t %d \n", GRADT2 - GRADT1, GRADT4 - GRADT3, GRADT2 - GRADT1 + 1);
    }
    if (msg && MYID == 0) {
        fprintf(FP, "\n **Message from taper_grid (printed by PE %d):\n", MYID);
        fprintf(FP, " Coefficients for gradient taper are now calculated.\n");
    }
    const size_t n = (size_t)NX * NY;
    rz_grad_taper t;
    memset(&t, 0, sizeof t);
    t.sws = sws;
    t.gradt1 = GRADT1; t.gradt2 = GRADT2; t.gradt3 = GRADT3; t.gradt4 = GRADT4;
    t.dh = DH;
    t.exp_taper_grad_hor = EXP_TAPER_GRAD_HOR;
    t.srtradius = SRTRADIUS;
    float *x = NULL, *z = NULL, *file = NULL;
    int *rx = NULL, *ry = NULL;
    if (sws == 3) {
        t.nsrc = nshots > 0 ? nshots : 0;
        t.ntr = ntr > 0 ? ntr : 0;
        x = rzs_alloc((size_t)(t.nsrc + 1) * sizeof(float));
        z = rzs_alloc((size_t)(t.nsrc + 1) * sizeof(float));
        rx = rzs_alloc((size_t)(t.ntr + 1) * sizeof(int));
        ry = rzs_alloc((size_t)(t.ntr + 1) * sizeof(int));
        for (int i = 0; i < t.nsrc; ++i) { x[i] = srcpos[1][i + 1]; z[i] = srcpos[2][i + 1]; }
        for (int i = 0; i < t.ntr; ++i) { rx[i] = recpos[1][i + 1] - 1; ry[i] = recpos[2][i + 1] - 1; }
        t.src_x_m = x; t.src_z_m = z; t.rec_ix = rx; t.rec_iy = ry;
    }
    if (sws >= 4) {                                         /* taper_grad.c:272-282 */
        FILE *fp = fopen("taper.bin", "r");
        if (!fp) err(" rajzel_amd taper_grad: taper.bin could not be opened");
        file = rzs_alloc(n * sizeof(float));
        const size_t got = fread(file, sizeof(float), n, fp);
        for (size_t i = got; i < n; ++i) file[i] = got ? file[got - 1] : 0.f;   /* a failed fread keeps the value */
        fclose(fp);
        t.file = file;
    }
    float **coeff = sws == 3 ? NULL : matrix(1, NY, 1, NX);
    const char *dev = getenv("RAJZEL_AMD_DEVICE");
    if (rz_taper_grad(dev ? atoi(dev) : 0, NULL, &waveconv[1][1], NX, NY, &t, coeff ? &coeff[1][1] : NULL, 0) !=
        RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd taper_grad: %s", rz_last_error(NULL));
        err(g_msg);
    }
    if (coeff) {
        char name[64];
        snprintf(name, sizeof name, "%s", sws == 1 ? "taper_coeff_vert.bin"
                                          : (sws == 2 ? "taper_coeff_hor.bin" : "taper_coeff_file.bin"));
        writemod(name, coeff, 3);
        free_matrix(coeff, 1, NY, 1, NX);
    }
    free(x); free(z); free(rx); free(ry); free(file);
}

/* gauss_filt.c:10-139 on the GPU.  The sizes are adjusted in the globals as
 * the reference adjusts them (an even size is made odd, 0 returns), and the
 * reference's report of the half sizes is kept. */
void gauss_filt(float **waveconv)
{
    if (FILT_SIZE_GRAD == 0) return;
    if (!(FILT_SIZE_GRAD % 2)) FILT_SIZE_GRAD += FILT_SIZE_GRAD > 0 ? 1 : -1;
    if (FILT_SIZE_GRAD1 == 0) return;
    if (!(FILT_SIZE_GRAD1 % 2)) FILT_SIZE_GRAD1 += FILT_SIZE_GRAD1 > 0 ? 1 : -1;
    if (MYID == 0) {
        printf("\n hfsx: %d \n", abs(FILT_SIZE_GRAD) / 2);
        printf("\n hfsy: %d \n", abs(FILT_SIZE_GRAD1) / 2);
    }
    const char *dev = getenv("RAJZEL_AMD_DEVICE");
    if (rz_gauss_filt(dev ? atoi(dev) : 0, NULL, &waveconv[1][1], NX, NY, &FILT_SIZE_GRAD, &FILT_SIZE_GRAD1, 0) !=
        RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd gauss_filt: %s", rz_last_error(NULL));
        err(g_msg);
    }
}

/* grid_search.c:10-96.  The models of the (vp0, grad0) grid are built in the
 * reference's order with its model_gridsearch() and calc_S(), and their
 * forwards run as multi-model contexts, every shot of this rank resident
 * (one launch per chunk of models).  Each model is then finished exactly as
 * the reference finishes it: calc_FA_res per shot in order (pick files,
 * Tres carry and the slope scaling need T only at the receiver and at its
 * slope neighbour, which the library returns), the float L2, the
 * MPI_Allreduce and write_gridsearch().  With SNAP set the reference writes
 * a traveltime snapshot per shot; the per-shot path (the shim's eikonal())
 * is kept for that case. */
static rz_ctx *rzs_grid_context(int nmod, int ns, float **srcpos, int **recpos, int ntr)
{
    int *sx = rzs_alloc(ns * sizeof(int)), *sy = rzs_alloc(ns * sizeof(int));
    int *rx = rzs_alloc((ntr > 0 ? ntr : 1) * sizeof(int)), *ry = rzs_alloc((ntr > 0 ? ntr : 1) * sizeof(int));
    for (int s = 0; s < ns; ++s) {
        sx[s] = rz_grid_index(srcpos[1][NSHOT1 + s], DH);   /* eikonal.c:31-32 */
        sy[s] = rz_grid_index(srcpos[2][NSHOT1 + s], DH);
    }
    for (int i = 0; i < ntr; ++i) {
        rx[i] = recpos[1][i + 1] - 1;
        ry[i] = recpos[2][i + 1] - 1;
    }
    rz_model_desc *d = rzs_alloc((size_t)nmod * sizeof(rz_model_desc));
    for (int k = 0; k < nmod; ++k) {
        memset(&d[k], 0, sizeof d[k]);
        d[k].nx = NX; d[k].ny = NY; d[k].dh = DH; d[k].time = TIME; d[k].ttnorm = TTNORM;
        d[k].eps_adj = EPS_ADJ;
        d[k].nshots = ns; d[k].src_ix = sx; d[k].src_iy = sy;
        d[k].ntr = ntr; d[k].rec_ix = rx; d[k].rec_iy = ry;
    }
    const char *dev = getenv("RAJZEL_AMD_DEVICE");
    rz_ctx *c = NULL;
    if (rz_create(&c, dev ? atoi(dev) : 0, nmod, d, 0, 0) != RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd grid search context: %s", rz_last_error(NULL));
        err(g_msg);
    }
    free(d); free(sx); free(sy); free(rx); free(ry);
    return c;
}

/* calc_FA_res.c:10-77 for one shot from the receiver times and the times
 * at the receivers' slope neighbours: reads the shot's pick file into
 * Tobs, keeps Tres of muted picks (the carry), and returns the shot's
 * l2 = sum of Tres^2 (float products, double sum).  The reference leaves
 * its l2 uninitialised; like the GPU residual kernel this sum starts at
 * 0, which is what the reference program computes in grad_obj (the bitwise
 * program tests compare its misfit logs). */
static double rzs_calc_fa_res(const float *tm, const float *tn, float *Tobs, float *Tres, int ntr, int ishot,
                              int **recpos)
{
    char name[RZS_PATH];
    double l2 = 0.0;
    snprintf(name, sizeof name, "%s_%i.dat", DATA_DIR, ishot);
    if (rz_read_picks(name, Tobs, ntr) != 0) err(" picks_?.dat could not be opened !");
    for (int i = 1; i <= ntr; i++) {
        if (Tobs[i] >= 0.0) {
            Tres[i] = tm[i - 1] - Tobs[i];
            l2 += Tres[i] * Tres[i];
            float gradT;
            if (recpos[2][i] <= NY - 1) {
                gradT = (tn[i - 1] - tm[i - 1]) / DH;
                if (fabs(gradT) > 0.0) Tres[i] = Tres[i] / gradT;
                else Tres[i] = 0.0;
            } else if (NY == recpos[2][i]) {
                gradT = (tm[i - 1] - tn[i - 1]) / DH;
                if (fabs(gradT) > 0.0) Tres[i] = Tres[i] / gradT;
                else Tres[i] = 0.0;
            }
        }
    }
    return l2;
}

static void rzs_grid_check(rz_ctx *c, int rc, const char *what)
{
    if (rc != RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd grid search %s: %s", what, rz_last_error(c));
        err(g_msg);
    }
}

void grid_search(float **Vp, float **S, float **TT, float *Tmod, float *Tobs, float *Tres, float **srcpos,
                 int nshots, int **recpos, int ntr)
{
    if (MYID == 0) {
        printf("\n\n\n ------------------------------------------------------------------\n");
        printf("\n\n\n                1D linear gradient grid search                     \n");
        printf("\n\n\n ------------------------------------------------------------------\n");
    }
    SEISMO = 0;
    const int nvp0 = ceil((VP0_2 - VP0_1) / DVP0);
    const int ngrad0 = ceil((GRAD0_2 - GRAD0_1) / DGRAD0);
    const int nobj = nvp0 * ngrad0;
    if (MYID == 0) printf("nvp0 =  %d, ngrad0 = %d\n", nvp0, ngrad0);
    const int ns = NSHOT2 > NSHOT1 ? NSHOT2 - NSHOT1 : 0;
    const size_t ncell = (size_t)NX * NY;
    const size_t nrec = (size_t)ns * (ntr > 0 ? ntr : 1);

    /* the grid's parameters in the reference's order (float accumulation) */
    float *vp0s = rzs_alloc((size_t)(nobj > 0 ? nobj : 1) * sizeof(float));
    float *grad0s = rzs_alloc((size_t)(nobj > 0 ? nobj : 1) * sizeof(float));
    {
        int m = 0;
        float vp0 = VP0_1;
        for (int i = 1; i <= nvp0; i++) {
            float grad0 = GRAD0_1;
            for (int j = 1; j <= ngrad0; j++) {
                vp0s[m] = vp0; grad0s[m] = grad0; ++m;
                grad0 += DGRAD0;
            }
            vp0 += DVP0;
        }
    }
    const int batched = (SNAP == 0 && ns > 0);
    /* models per launch: every shot's four fields resident, within ~64 GiB
     * (a field is at most (NX+64) x (NY+511) floats) */
    int chunk = 1;
    if (batched) {
        const double per_model = (double)ns * 16.0 * (NX + 64.0) * (NY + 511.0);
        const double fit = (double)(64ull << 30) / per_model;
        chunk = fit >= 32.0 ? 32 : (fit >= 1.0 ? (int)fit : 1);
        if (chunk > nobj) chunk = nobj > 0 ? nobj : 1;
    }
    float *Sall = batched ? rzs_alloc((size_t)chunk * ncell * sizeof(float)) : NULL;
    float *tm = batched ? rzs_alloc((size_t)chunk * nrec * sizeof(float)) : NULL;
    float *tn = batched ? rzs_alloc((size_t)chunk * nrec * sizeof(float)) : NULL;
    rz_ctx *c = NULL;
    int c_models = 0;
    for (int m0 = 0; m0 < nobj; m0 += (batched ? chunk : nobj)) {
        const int m1 = batched ? (m0 + chunk < nobj ? m0 + chunk : nobj) : nobj;
        if (batched) {
            if (c_models != m1 - m0) {
                rz_destroy(c);
                c = rzs_grid_context(m1 - m0, ns, srcpos, recpos, ntr);
                c_models = m1 - m0;
            }
            for (int m = m0; m < m1; ++m) {
                model_gridsearch(Vp, vp0s[m], grad0s[m]);
                calc_S(Vp, S);
                memcpy(Sall + (size_t)(m - m0) * ncell, &S[1][1], ncell * sizeof(float));
            }
            for (int k = 0; k < m1 - m0; ++k)
                rzs_grid_check(c, rz_set_slowness(c, k, Sall + (size_t)k * ncell, 0), "slowness");
            rzs_grid_check(c, rz_forward(c), "forward");
            if (ntr > 0)
                for (int k = 0; k < m1 - m0; ++k) {
                    rzs_grid_check(c, rz_get_tmod(c, k, tm + (size_t)k * nrec, 0), "receiver times");
                    rzs_grid_check(c, rz_get_tnb(c, k, tn + (size_t)k * nrec, 0), "receiver times");
                }
        }
        for (int m = m0; m < m1; ++m) {
            if (MYID == 0) printf("Model %d of %d \n", m + 1, nobj);
            float L2 = 0.0;
            if (!batched) {
                model_gridsearch(Vp, vp0s[m], grad0s[m]);
                calc_S(Vp, S);
                for (int ishot = NSHOT1; ishot < NSHOT2; ishot++) {
                    eikonal(S, TT, Tmod, srcpos, ishot, nshots, recpos, ntr);
                    L2 += calc_FA_res(TT, Tmod, Tobs, Tres, ntr, ishot, recpos);
                }
            } else {
                for (int s = 0; s < ns; ++s) {
                    const float *a = tm + (size_t)(m - m0) * nrec + (size_t)s * ntr;
                    const float *b = tn + (size_t)(m - m0) * nrec + (size_t)s * ntr;
                    for (int i = 1; i <= ntr; ++i) Tmod[i] = a[i - 1];      /* calc_FA.c:15-17 */
                    L2 += rzs_calc_fa_res(a, b, Tobs, Tres, ntr, NSHOT1 + s, recpos);
                }
            }
            float L2sum = 0.0;
            MPI_Barrier(MPI_COMM_WORLD);
            MPI_Allreduce(&L2, &L2sum, 1, MPI_FLOAT, MPI_SUM, MPI_COMM_WORLD);
            L2 = 0.5 * L2sum;
            if (MYID == 0) write_gridsearch(L2, vp0s[m], grad0s[m], m + 1);
        }
    }
    /* the reference leaves the last model's last shot in TT */
    if (batched && nobj > 0) rzs_grid_check(c, rz_get_field(c, c_models - 1, ns - 1, 0, &TT[1][1], 0), "traveltimes");
    rz_destroy(c);
    free(Sall); free(tm); free(tn); free(vp0s); free(grad0s);
}

/* ---------------------------------------------------------------------------
 * LBFGS.c:12-226 (fd.h:88): the two-loop recursion of fatt.c:160, and the
 * previous model and gradient exchanged through the same files
 * <JACOBIAN>_p_vp.old, _p.old and _c.old (each read or written in one call;
 * the reference makes one fread/fwrite call per float, about 80 ms per 12 MB
 * file; a file shorter than the grid leaves the rest as the reference's
 * fread leaves it: the last value read).
 *
 * Unstrided (IDX = IDY = 1, the vectors NX*NY long), the recursion runs on
 * the GPU (rz_lbfgs_step): the history slots stay in device memory, and the
 * direction, gamma, rho and alpha are the reference's bit for bit, every
 * dot product being dotp()'s serial float sum.  The device keeps the
 * previous model and gradient itself (the files are still written); it
 * reads them from the files only when it has none.  Its history is reset by
 * zero_LBFGS and zero_LBFGS1 below, the only writers of the history besides
 * LBFGS (check_descent.c:39, fatt.c:232), so the caller's y_LBFGS, s_LBFGS,
 * q_LBFGS and r_LBFGS arrays are left as they are.  Strided, the reference's
 * arithmetic runs on the host on the caller's arrays (its dotp() for every
 * dot product, the same float operations in the same order).
 * --------------------------------------------------------------------------- */
static size_t rzs_strided(void)
{
    return (size_t)((NX + IDX - 1) / IDX) * (size_t)((NY + IDY - 1) / IDY);
}

/* matrix <-> buffer in the reference's file order */
static void rzs_gather(float **A, float *buf)
{
    size_t t = 0;
    for (int i = 1; i <= NX; i += IDX)
        for (int j = 1; j <= NY; j += IDY) buf[t++] = A[j][i];
}

static void rzs_scatter(const float *buf, float **A)
{
    size_t t = 0;
    for (int i = 1; i <= NX; i += IDX)
        for (int j = 1; j <= NY; j += IDY) A[j][i] = buf[t++];
}

static void rzs_read_old(const char *suffix, float *buf, size_t n)
{
    char name[RZS_PATH + 16];
    snprintf(name, sizeof name, "%s%s", JACOBIAN, suffix);
    FILE *fp = fopen(name, "rb");
    if (fp == NULL) {
        snprintf(g_msg, sizeof g_msg, " LBFGS: %s could not be opened", name);
        err(g_msg);
    }
    const size_t got = fread(buf, sizeof(float), n, fp);
    fclose(fp);
    for (size_t t = got; t < n; ++t) buf[t] = got ? buf[got - 1] : 0.f;
}

static void rzs_write_old(const char *suffix, const float *buf, size_t n)
{
    char name[RZS_PATH + 16];
    snprintf(name, sizeof name, "%s%s", JACOBIAN, suffix);
    FILE *fp = fopen(name, "wb");
    if (fp == NULL || fwrite(buf, sizeof(float), n, fp) != n) {
        snprintf(g_msg, sizeof g_msg, " LBFGS: %s could not be written", name);
        err(g_msg);
    }
    fclose(fp);
}

static rz_lbfgs *g_lb;              /* the device history (rank 0) */

static void rzs_lbfgs_release(void)
{
    rz_lbfgs_destroy(g_lb);
    g_lb = NULL;
}

static void rzs_lbfgs_check(int rc, const char *what)
{
    if (rc != RZ_OK) {
        snprintf(g_msg, sizeof g_msg, " rajzel_amd LBFGS %s: %s", what, rz_lbfgs_error(g_lb));
        err(g_msg);
    }
}

/* LBFGS on the GPU (unstrided); the log lines and files of LBFGS.c:75-216 */
static void rzs_lbfgs_device(float **Hgrad, float **grad, float **gradm, int iter, float *rho_LBFGS,
                             float *alpha_LBFGS, float **Vp, int LBFGS_pointer, int NLBFGS)
{
    const size_t n = (size_t)NX * NY;
    if (g_lb == NULL) {
        const char *dev = getenv("RAJZEL_AMD_DEVICE");
        if (rz_lbfgs_create(&g_lb, dev ? atoi(dev) : 0, NX, NY, NLBFGS) != RZ_OK) {
            snprintf(g_msg, sizeof g_msg, " rajzel_amd LBFGS: %s", rz_last_error(NULL));
            err(g_msg);
        }
        atexit(rzs_lbfgs_release);
    }
    float *buf = rzs_alloc(3 * n * sizeof(float));
    float *gh = buf, *vh = buf + n, *hh = buf + 2 * n;
    if (iter > 1 && !rz_lbfgs_has_previous(g_lb)) {
        rzs_read_old("_p.old", gh, n);
        rzs_read_old("_p_vp.old", vh, n);
        rzs_lbfgs_check(rz_lbfgs_set_previous(g_lb, gh, vh), "previous model");
    }
    float gamma = 0.f;
    rzs_lbfgs_check(rz_lbfgs_step(g_lb, iter, LBFGS_pointer, &grad[1][1], &gradm[1][1], &Vp[1][1], &Hgrad[1][1], gh,
                                  vh, hh, &gamma, rho_LBFGS + 1, alpha_LBFGS + 1),
                    "step");
    if (iter > 1) {
        printf("MYID = %d \n", MYID);
        printf("gamma_LBFGS = %e \n", gamma);
        for (int k = 1; k <= NLBFGS; ++k) printf("rho_LBFGS = %e of k = %d \n", rho_LBFGS[k], k);
    }
    rzs_write_old("_p_vp.old", vh, n);
    rzs_write_old("_p.old", gh, n);
    rzs_write_old("_c.old", hh, n);
    free(buf);
}

/* zero_LBFGS.c:12-31 and zero_LBFGS1.c:12-21 (fd.h:172,175), and the device
 * history with them */
void zero_LBFGS(int NLBFGS, int NLBFGS_vec, float *y_LBFGS, float *s_LBFGS, float *q_LBFGS, float *r_LBFGS,
                float *alpha_LBFGS, float *beta_LBFGS, float *rho_LBFGS)
{
    for (long i = 1; i <= (long)NLBFGS_vec * NLBFGS; i++) y_LBFGS[i] = s_LBFGS[i] = 0.0f;
    for (long i = 1; i <= NLBFGS_vec; i++) q_LBFGS[i] = r_LBFGS[i] = 0.0f;
    for (int i = 1; i <= NLBFGS; i++) alpha_LBFGS[i] = beta_LBFGS[i] = rho_LBFGS[i] = 0.0f;
    if (g_lb) rzs_lbfgs_check(rz_lbfgs_zero(g_lb), "zero");
}

void zero_LBFGS1(int NLBFGS, int NLBFGS_vec, float *y_LBFGS, float *s_LBFGS)
{
    for (long i = 1; i <= (long)NLBFGS_vec * NLBFGS; i++) y_LBFGS[i] = s_LBFGS[i] = 0.0f;
    if (g_lb) rzs_lbfgs_check(rz_lbfgs_zero(g_lb), "zero");
}

void LBFGS(float **Hgrad, float **grad, float **gradm, int iter, float *y_LBFGS, float *s_LBFGS, float *rho_LBFGS,
           float *alpha_LBFGS, float **Vp, float *q_LBFGS, float *r_LBFGS, float *beta_LBFGS, int LBFGS_pointer,
           int NLBFGS, int NLBFGS_vec)
{
    (void)beta_LBFGS;   /* the reference keeps beta in a local */
    if (MYID == 0 && IDX == 1 && IDY == 1 && NLBFGS_vec == NX * NY && iter >= 1 &&
        getenv_flag("RAJZEL_AMD_LBFGS", 1)) {
        rzs_lbfgs_device(Hgrad, grad, gradm, iter, rho_LBFGS, alpha_LBFGS, Vp, LBFGS_pointer, NLBFGS);
    } else if (MYID == 0) {
        const size_t n = rzs_strided();
        float *buf = rzs_alloc(2 * n * sizeof(float));
        float *old = buf + n;
        if (iter > 1) {
            /* y and s of this pair into the slot LBFGS_pointer (LBFGS.c:43-66) */
            const int h0 = NLBFGS_vec * (LBFGS_pointer - 1) + 1;
            rzs_read_old("_p.old", old, n);
            rzs_gather(grad, buf);
            for (size_t t = 0; t < n; ++t) y_LBFGS[h0 + t] = buf[t] - old[t];
            rzs_read_old("_p_vp.old", old, n);
            rzs_gather(Vp, buf);
            for (size_t t = 0; t < n; ++t) s_LBFGS[h0 + t] = buf[t] - old[t];
            /* initial Hessian scale and rho of every slot (LBFGS.c:68-96) */
            const float gamma = dotp(y_LBFGS, s_LBFGS, h0, h0 + NLBFGS_vec - 1, 0) /
                                dotp(y_LBFGS, y_LBFGS, h0, h0 + NLBFGS_vec - 1, 0);
            printf("MYID = %d \n", MYID);
            printf("gamma_LBFGS = %e \n", gamma);
            for (int k = 1; k <= NLBFGS; ++k) {
                const int a = NLBFGS_vec * (k - 1) + 1;
                const float ys = dotp(y_LBFGS, s_LBFGS, a, a + NLBFGS_vec - 1, 0);
                rho_LBFGS[k] = fabs(ys) > 0.0 ? (float)(1.0 / ys) : 0.0f;
                printf("rho_LBFGS = %e of k = %d \n", rho_LBFGS[k], k);
            }
            /* first loop, newest slot last in the reference's order k = NLBFGS..1 (LBFGS.c:98-125) */
            rzs_gather(gradm, q_LBFGS + 1);
            for (int k = NLBFGS; k >= 1; --k) {
                const int a = NLBFGS_vec * (k - 1) + 1;
                alpha_LBFGS[k] = rho_LBFGS[k] * dotp(s_LBFGS, q_LBFGS, a, a + NLBFGS_vec - 1, 1);
                const float *yk = y_LBFGS + a - 1;
                for (int i = 1; i <= NLBFGS_vec; ++i) q_LBFGS[i] = q_LBFGS[i] - alpha_LBFGS[k] * yk[i];
            }
            for (int i = 1; i <= NLBFGS_vec; ++i) r_LBFGS[i] = gamma * q_LBFGS[i];   /* LBFGS.c:128-131 */
            /* second loop (LBFGS.c:133-149) */
            for (int k = 1; k <= NLBFGS; ++k) {
                const int a = NLBFGS_vec * (k - 1) + 1;
                const float beta = rho_LBFGS[k] * dotp(y_LBFGS, r_LBFGS, a, a + NLBFGS_vec - 1, 1);
                const float *sk = s_LBFGS + a - 1;
                for (int i = 1; i <= NLBFGS_vec; ++i) r_LBFGS[i] = r_LBFGS[i] + sk[i] * (alpha_LBFGS[k] - beta);
            }
            rzs_scatter(r_LBFGS + 1, Hgrad);                 /* LBFGS.c:151-161 */
        } else if (iter == 1) {
            rzs_gather(gradm, buf);                          /* LBFGS.c:165-174 */
            rzs_scatter(buf, Hgrad);
        }
        /* this iteration's model, gradient and direction (LBFGS.c:178-216) */
        rzs_gather(Vp, buf);
        rzs_write_old("_p_vp.old", buf, n);
        rzs_gather(grad, buf);
        rzs_write_old("_p.old", buf, n);
        rzs_gather(Hgrad, buf);
        rzs_write_old("_c.old", buf, n);
        free(buf);
    }
    MPI_Barrier(MPI_COMM_WORLD);                             /* LBFGS.c:221-223 */
    exchange_grad_MPI(Hgrad);
}
