/*
 * ref_harness.c -- drives the UNMODIFIED reference hot-path sources, compiled
 * from /root/reference/src by oracle/Makefile into oracle/_ref/librzref.so.
 *
 * TEST INFRASTRUCTURE ONLY (golden-vector generation, parity pinning of the
 * restatement in rz_oracle.c, and the reference CPU baseline of bench.py).
 *
 * The reference keeps all state in globals defined by globvar.h, which this
 * file includes once (the reference's own rajzel.c does the same).  Outer
 * iterations of eikonal()/eikonal_adj() are counted by wrapping norm1() at
 * link time (-Wl,--wrap=norm1): every outer pass calls it exactly once
 * (eikonal.c:56, eikonal_adj.c:49).
 *
 * grad_obj.c itself is not compiled: it calls precond(), whose chain pulls
 * wavenumber.c and FFTW3, which this image lacks.  rzref_grad_obj() below
 * runs grad_obj.c's shot loop (grad_obj.c:27-70) over the reference's own
 * init_grad / init_recflag / eikonal / calc_FA_res / eikonal_adj / ass_grad /
 * cp_grad_frame, i.e. one rank, precond = identity (SPATFILTER=0, no tapers).
 *
 * Flat arrays are row-major NY x NX; indices passed here are 1-based (the
 * reference's own convention) and positions are in metres.
 */
#include "fd.h"
#include "globvar.h"

static int g_norm1_calls;
float __real_norm1(float **A, float **B);
float __wrap_norm1(float **A, float **B)
{
    ++g_norm1_calls;
    return __real_norm1(A, B);
}

void rzref_setup(int nx, int ny, float dh, float time, float ttnorm, float eps_adj)
{
    NX = nx; NY = ny; DH = dh; TIME = time; TTNORM = ttnorm; EPS_ADJ = eps_adj;
    SNAP = 0; SEISMO = 0; INFO = 0; MYID = 0; NP = 1;
    SWS_TAPER_CIRCULAR_PER_SHOT = 0;
}

/* Per-shot gradient taper of ass_grad.c:31-37: the reference's globals
 * SWS_TAPER_CIRCULAR_PER_SHOT, SRTRADIUS and FILTSIZE (read_par.c:183-192). */
void rzref_set_taper(int sws, float srtradius, int filtsize)
{
    SWS_TAPER_CIRCULAR_PER_SHOT = sws;
    SRTRADIUS = srtradius;
    FILTSIZE = filtsize;
}

static float **mat_from(const float *src)
{
    float **m = matrix(1, NY, 1, NX);
    if (src) memcpy(&m[1][1], src, (size_t)NX * NY * sizeof(float));
    return m;
}

static void mat_to(float **m, float *dst)
{
    if (dst) memcpy(dst, &m[1][1], (size_t)NX * NY * sizeof(float));
}

static int **recpos_from(int ntr, const int *rx1, const int *ry1)
{
    int **r = imatrix(1, 3, 1, ntr > 0 ? ntr : 1);
    for (int i = 1; i <= ntr; ++i) { r[1][i] = rx1[i - 1]; r[2][i] = ry1[i - 1]; r[3][i] = 0; }
    return r;
}

/* One reference sweep(): dir 0..3 = (x+,y+),(x-,y+),(x-,y-),(x+,y-). */
void rzref_sweep(const float *S, float *TT, int sx, int sy)
{
    float **s = mat_from(S), **t = mat_from(TT);
    sweep(s, t, 0, 0, sx > 0 ? 1 : NX, sx > 0 ? NX : 1, sx, sy > 0 ? 1 : NY, sy > 0 ? NY : 1, sy);
    mat_to(t, TT);
    free_matrix(s, 1, NY, 1, NX); free_matrix(t, 1, NY, 1, NX);
}

/* One reference sweep_adj() in direction (sx, sy). */
void rzref_sweep_adj_dir(float *lam, const float *TT, const int *recflag, int sx, int sy)
{
    float **l = mat_from(lam), **t = mat_from(TT);
    int **f = imatrix(1, NY, 1, NX);
    memcpy(&f[1][1], recflag, (size_t)NX * NY * sizeof(int));
    sweep_adj(l, t, NULL, NULL, 0, sx > 0 ? 2 : NX - 1, sx > 0 ? NX - 1 : 2, sx,
              sy > 0 ? 2 : NY - 1, sy > 0 ? NY - 1 : 2, sy, f);
    mat_to(l, lam);
    free_matrix(l, 1, NY, 1, NX); free_matrix(t, 1, NY, 1, NX); free_imatrix(f, 1, NY, 1, NX);
}

/* eikonal() for one shot at (srcx, srcy) metres; returns outer iterations. */
int rzref_eikonal(const float *S, float srcx, float srcy, int ntr, const int *rx1, const int *ry1,
                  float *TT_out, float *tmod_out)
{
    float **s = mat_from(S), **t = mat_from(NULL);
    float **srcpos = matrix(1, 8, 1, 1);
    float *tmod = vector(1, ntr > 0 ? ntr : 1);
    int **rec = recpos_from(ntr, rx1, ry1);
    srcpos[1][1] = srcx; srcpos[2][1] = srcy;
    g_norm1_calls = 0;
    eikonal(s, t, tmod, srcpos, 1, 1, rec, ntr);
    int it = g_norm1_calls;
    mat_to(t, TT_out);
    if (tmod_out) memcpy(tmod_out, &tmod[1], (size_t)ntr * sizeof(float));
    free_matrix(s, 1, NY, 1, NX); free_matrix(t, 1, NY, 1, NX); free_matrix(srcpos, 1, 8, 1, 1);
    free_vector(tmod, 1, ntr > 0 ? ntr : 1); free_imatrix(rec, 1, 3, 1, ntr > 0 ? ntr : 1);
    return it;
}

/* eikonal_adj() for one shot; lam initialised by the reference itself. */
int rzref_eikonal_adj(const float *TT, int ntr, const int *rx1, const int *ry1, const float *tres,
                      float *lam_out)
{
    float **t = mat_from(TT), **l = mat_from(NULL);
    int **rec = recpos_from(ntr, rx1, ry1);
    int **f = imatrix(1, NY, 1, NX);
    float *tr = vector(1, ntr > 0 ? ntr : 1);
    memcpy(&tr[1], tres, (size_t)ntr * sizeof(float));
    init_recflag(f, rec, ntr);
    g_norm1_calls = 0;
    eikonal_adj(t, tr, l, rec, ntr, 1, f);
    int it = g_norm1_calls;
    mat_to(l, lam_out);
    free_matrix(t, 1, NY, 1, NX); free_matrix(l, 1, NY, 1, NX); free_imatrix(f, 1, NY, 1, NX);
    free_imatrix(rec, 1, 3, 1, ntr > 0 ? ntr : 1); free_vector(tr, 1, ntr > 0 ? ntr : 1);
    return it;
}

/* init_recflag() into a flat int array. */
void rzref_init_recflag(int ntr, const int *rx1, const int *ry1, int *flag_out)
{
    int **rec = recpos_from(ntr, rx1, ry1);
    int **f = imatrix(1, NY, 1, NX);
    init_recflag(f, rec, ntr);
    memcpy(flag_out, &f[1][1], (size_t)NX * NY * sizeof(int));
    free_imatrix(f, 1, NY, 1, NX); free_imatrix(rec, 1, 3, 1, ntr > 0 ? ntr : 1);
}

/* ass_grad() for one shot: grad (in/out) += frame-copied lam*S^3. */
void rzref_ass_grad(const float *S, const float *lam, float *grad_shot_out, float *grad_io)
{
    float **s = mat_from(S), **l = mat_from(lam), **g = mat_from(grad_io), **gs = mat_from(NULL);
    float **srcpos = matrix(1, 8, 1, 1);
    int **rec = imatrix(1, 3, 1, 1);
    ass_grad(g, gs, s, l, srcpos, 1, rec, 0, 1);
    mat_to(g, grad_io); mat_to(gs, grad_shot_out);
    free_matrix(s, 1, NY, 1, NX); free_matrix(l, 1, NY, 1, NX); free_matrix(g, 1, NY, 1, NX);
    free_matrix(gs, 1, NY, 1, NX); free_matrix(srcpos, 1, 8, 1, 1); free_imatrix(rec, 1, 3, 1, 1);
}

/* ass_grad() for shot ishot (1-based) of nshots sources (metres) and ntr
 * receivers (1-based cells): the per-shot taper set by rzref_set_taper
 * applies. */
void rzref_ass_grad_at(const float *S, const float *lam, int nshots, const float *srcx, const float *srcy, int ntr,
                       const int *rx1, const int *ry1, int ishot, float *grad_shot_out, float *grad_io)
{
    float **s = mat_from(S), **l = mat_from(lam), **g = mat_from(grad_io), **gs = mat_from(NULL);
    float **srcpos = matrix(1, 8, 1, nshots);
    for (int i = 1; i <= nshots; ++i) { srcpos[1][i] = srcx[i - 1]; srcpos[2][i] = srcy[i - 1]; }
    int **rec = recpos_from(ntr, rx1, ry1);
    ass_grad(g, gs, s, l, srcpos, nshots, rec, ntr, ishot);
    mat_to(g, grad_io); mat_to(gs, grad_shot_out);
    free_matrix(s, 1, NY, 1, NX); free_matrix(l, 1, NY, 1, NX); free_matrix(g, 1, NY, 1, NX);
    free_matrix(gs, 1, NY, 1, NX); free_matrix(srcpos, 1, 8, 1, nshots);
    free_imatrix(rec, 1, 3, 1, ntr > 0 ? ntr : 1);
}

/* taper_grad() (precond.c:43-53) on a field (in/out): sws 1..6 with the
 * globals GRADT1..4, EXP_TAPER_GRAD_HOR, SRTRADIUS; nsrc sources (metres)
 * and ntr receivers (1-based cells) for sws 3.  The reference reads
 * taper.bin (sws 4..6) and writes taper_coeff_*.bin in the working
 * directory; its log lines go to FP (/dev/null here unless set). */
void rzref_taper_grad(float *field_io, int nsrc, const float *srcx, const float *srcy, int ntr, const int *rx1,
                      const int *ry1, int sws, int g1, int g2, int g3, int g4, float exp_hor, float srtradius)
{
    if (!FP) FP = fopen("/dev/null", "w");
    IDX = 1; IDY = 1;
    GRADT1 = g1; GRADT2 = g2; GRADT3 = g3; GRADT4 = g4;
    EXP_TAPER_GRAD_HOR = exp_hor;
    SRTRADIUS = srtradius;
    float **g = mat_from(field_io);
    const int ns = nsrc > 0 ? nsrc : 1;
    float **srcpos = matrix(1, 8, 1, ns);
    for (int i = 1; i <= nsrc; ++i) { srcpos[1][i] = srcx[i - 1]; srcpos[2][i] = srcy[i - 1]; }
    int **rec = recpos_from(ntr, rx1, ry1);
    taper_grad(g, srcpos, nsrc, rec, ntr, sws);
    mat_to(g, field_io);
    free_matrix(g, 1, NY, 1, NX); free_matrix(srcpos, 1, 8, 1, ns);
    free_imatrix(rec, 1, 3, 1, ntr > 0 ? ntr : 1);
}

static void write_picks_file(const char *stem, int shot1, int ntr, const float *v)
{
    char fn[STRING_SIZE2];
    snprintf(fn, sizeof fn, "%s_%i.dat", stem, shot1);
    FILE *fp = fopen(fn, "w");
    if (!fp) { fprintf(stderr, "ref_harness: cannot write %s\n", fn); exit(3); }
    for (int i = 0; i < ntr; ++i) fprintf(fp, "%e\n", v[i]);
    fclose(fp);
}

/*
 * grad_obj.c's shot loop over the reference per-shot routines.
 *   datadir_stem  DATA_DIR value (< 74 chars); pick files <stem>_<shot>.dat
 *                 are written here from tobs with "%e" and read back by the
 *                 reference's calc_FA_res; tobs_read receives what it read.
 *   tres_io       ntr: initial contents of the reference's Tres array (the
 *                 carry for muted picks), returns the last shot's values.
 *   per-shot outputs may be NULL.
 * Returns L2 (0.5 * sum of the reference's per-shot values).
 */
double rzref_grad_obj(const float *S, int nshots, const float *srcx, const float *srcy, int ntr,
                      const int *rx1, const int *ry1, const float *tobs, const char *datadir_stem,
                      float *tres_io, float *grad_out, float *TT_last, float *lam_last,
                      float *tobs_read, float *tmod_all, float *tres_all, double *l2_shot,
                      int *it_f, int *it_a, int write_files)
{
    float **s = mat_from(S), **t = mat_from(NULL), **l = mat_from(NULL);
    float **g = mat_from(NULL), **gs = mat_from(NULL);
    float **srcpos = matrix(1, 8, 1, nshots);
    int **rec = recpos_from(ntr, rx1, ry1);
    int **f = imatrix(1, NY, 1, NX);
    const int nv = ntr > 0 ? ntr : 1;
    float *tmod = vector(1, nv), *tob = vector(1, nv), *tr = vector(1, nv);
    memcpy(&tr[1], tres_io, (size_t)ntr * sizeof(float));
    for (int i = 1; i <= nshots; ++i) { srcpos[1][i] = srcx[i - 1]; srcpos[2][i] = srcy[i - 1]; }
    strncpy(DATA_DIR, datadir_stem, STRING_SIZE - 1);
    DATA_DIR[STRING_SIZE - 1] = 0;
    if (write_files)
        for (int sh = 1; sh <= nshots; ++sh) write_picks_file(DATA_DIR, sh, ntr, tobs + (size_t)(sh - 1) * ntr);

    /* grad_obj.c:24-31 */
    SEISMO = 0;
    init_grad(g);
    double L2 = 0.0;
    init_recflag(f, rec, ntr);
    for (int ishot = 1; ishot <= nshots; ++ishot) {           /* grad_obj.c:33 */
        init_grad(gs);
        g_norm1_calls = 0;
        eikonal(s, t, tmod, srcpos, ishot, nshots, rec, ntr);
        if (it_f) it_f[ishot - 1] = g_norm1_calls;
        const double l2 = calc_FA_res(t, tmod, tob, tr, ntr, ishot, rec);
        L2 += l2;
        g_norm1_calls = 0;
        eikonal_adj(t, tr, l, rec, ntr, ishot, f);
        if (it_a) it_a[ishot - 1] = g_norm1_calls;
        ass_grad(g, gs, s, l, srcpos, nshots, rec, ntr, ishot);
        if (l2_shot) l2_shot[ishot - 1] = l2;
        if (tmod_all) memcpy(tmod_all + (size_t)(ishot - 1) * ntr, &tmod[1], (size_t)ntr * sizeof(float));
        if (tres_all) memcpy(tres_all + (size_t)(ishot - 1) * ntr, &tr[1], (size_t)ntr * sizeof(float));
        if (tobs_read) memcpy(tobs_read + (size_t)(ishot - 1) * ntr, &tob[1], (size_t)ntr * sizeof(float));
    }
    L2 = 0.5 * L2;                                             /* grad_obj.c:61, one rank */
    cp_grad_frame(g);                                          /* grad_obj.c:70 */

    mat_to(g, grad_out); mat_to(t, TT_last); mat_to(l, lam_last);
    memcpy(tres_io, &tr[1], (size_t)ntr * sizeof(float));
    free_matrix(s, 1, NY, 1, NX); free_matrix(t, 1, NY, 1, NX); free_matrix(l, 1, NY, 1, NX);
    free_matrix(g, 1, NY, 1, NX); free_matrix(gs, 1, NY, 1, NX); free_matrix(srcpos, 1, 8, 1, nshots);
    free_imatrix(rec, 1, 3, 1, nv); free_imatrix(f, 1, NY, 1, NX);
    free_vector(tmod, 1, nv); free_vector(tob, 1, nv); free_vector(tr, 1, nv);
    return L2;
}

/* The reference's gauss_filt() on a flat NY x NX field (NX, NY from
 * rzref_setup).  The sizes go in through the globals it reads and come back
 * as it leaves them (made odd).  MYID is raised during the call so that its
 * report of hfsx/hfsy stays off stdout. */
void rzref_gauss_filt(float *field, int *fsx, int *fsy)
{
    float **m = mat_from(field);
    FILT_SIZE_GRAD = *fsx;
    FILT_SIZE_GRAD1 = *fsy;
    const int myid = MYID;
    MYID = 1;
    gauss_filt(m);
    MYID = myid;
    *fsx = FILT_SIZE_GRAD;
    *fsy = FILT_SIZE_GRAD1;
    mat_to(m, field);
    free_matrix(m, 1, NY, 1, NX);
}

/* The reference's LBFGS() (LBFGS.c:12-226) on flat row-major fields, with
 * the caller's 1-based history and work vectors (y, s: nlbfgs*NX*NY + 1
 * floats; q, r: NX*NY + 1; rho, alpha, beta: nlbfgs + 1) kept between calls
 * by the caller, IDX = idx, IDY = idy, and its files <jac>_p.old,
 * _p_vp.old and _c.old.  One rank: LBFGS.c's MPI_Barrier and
 * exchange_grad_MPI (a broadcast from rank 0) leave everything as it is,
 * so both are wrapped at link time (-Wl,--wrap) to no-ops. */
int __wrap_MPI_Barrier(MPI_Comm comm)
{
    (void)comm;
    return MPI_SUCCESS;
}

void __wrap_exchange_grad_MPI(float **grad)
{
    (void)grad;
}

void rzref_lbfgs(const char *jac, int idx, int idy, int iter, int pointer, int nlbfgs, float *y, float *s,
                 float *rho, float *alpha, float *beta, float *q, float *r, const float *grad, const float *gradm,
                 const float *vp, float *hgrad)
{
    float **G = mat_from(grad), **Gm = mat_from(gradm), **V = mat_from(vp), **H = mat_from(hgrad);
    IDX = idx;
    IDY = idy;
    MYID = 0;
    snprintf(JACOBIAN, STRING_SIZE, "%s", jac);
    LBFGS(H, G, Gm, iter, y, s, rho, alpha, V, q, r, beta, pointer, nlbfgs, NX * NY);
    mat_to(H, hgrad);
    free_matrix(G, 1, NY, 1, NX);
    free_matrix(Gm, 1, NY, 1, NX);
    free_matrix(V, 1, NY, 1, NX);
    free_matrix(H, 1, NY, 1, NX);
}

/* The reference's smooth_grad() (SPATFILTER = 3) on a flat NY x NX field;
 * the size goes in through FILT_SIZE_GRAD and comes back as it leaves it. */
void rzref_smooth_grad(float *field, int *fsize)
{
    float **m = mat_from(field);
    FILT_SIZE_GRAD = *fsize;
    smooth_grad(m);
    *fsize = FILT_SIZE_GRAD;
    mat_to(m, field);
    free_matrix(m, 1, NY, 1, NX);
}

/* The reference's smooth2() (SPATFILTER = 2) on a flat NY x NX field, one
 * rank (NXG = NX, NYG = NY, POS = 0, IDX = IDY = 1), its two files under
 * JACOBIAN = jac. */
void rzref_smooth2(float *field, float wd_damp, float wd_damp1, const char *jac)
{
    float **m = mat_from(field);
    NXG = NX;
    NYG = NY;
    POS[1] = POS[2] = 0;
    IDX = IDY = 1;
    MYID = 0;
    WD_DAMP = wd_damp;
    WD_DAMP1 = wd_damp1;
    snprintf(JACOBIAN, STRING_SIZE, "%s", jac);
    smooth2(m);
    mat_to(m, field);
    free_matrix(m, 1, NY, 1, NX);
}
