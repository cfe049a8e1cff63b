// HIP kernels of the FATT gradient path for gfx950 (CDNA4).
//
// One workgroup owns one shot for the whole forward (or adjoint) solve: its
// waves are the row bands of the grid, chained through an LDS ring, and the
// outer fast-sweeping iteration and its convergence test never leave the
// workgroup.  A launch carries every shot of the batch (all models), so the
// reference's MPI shot loop (grad_obj.c:33, init_MPIshot.c:19-30) becomes
// the grid of one launch.
//
// Inside a band, lane l owns R consecutive rows and walks along x; lane l+1
// trails lane l by one step (one column), so at every step the wave updates
// one anti-diagonal of the band.  With the band-skewed layouts of
// rz_internal.h that anti-diagonal is 64*R contiguous floats.  Cell values
// cross lanes with DPP wave shifts (row above/below), cross bands through
// an LDS ring (new values) and single-lane global loads (old values).
//
// Arithmetic is the reference's, operation by operation (compile with
// -ffp-contract=off): see godunov() and adj_cell().  Any topological order
// of a Gauss-Seidel sweep's dependency DAG gives the same bits, so each
// sweep here equals the reference's sweep bit for bit.
#include "rz_internal.h"

namespace rz {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_from_below(float v)  // lane l gets lane l+1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_from_above(float v)  // lane l gets lane l-1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, false));
}

template <int R>
__device__ __forceinline__ void ldv(const float *p, float (&o)[R])
{
    if constexpr (R == 1) {
        o[0] = p[0];
    } else if constexpr (R == 2) {
        const float2 t = *reinterpret_cast<const float2 *>(p);
        o[0] = t.x; o[1] = t.y;
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            const float4 t = *reinterpret_cast<const float4 *>(p + i);
            o[i] = t.x; o[i + 1] = t.y; o[i + 2] = t.z; o[i + 3] = t.w;
        }
    }
}

template <int R>
__device__ __forceinline__ void stv(float *p, const float (&v)[R])
{
    if constexpr (R == 1) {
        p[0] = v[0];
    } else if constexpr (R == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4)
            *reinterpret_cast<float4 *>(p + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
    }
}

template <int R>
__device__ __forceinline__ int addr_a(int k, int c, long long bstride)
{
    constexpr int BR = kWave * R;
    const int b = k / BR, rem = k - b * BR, l = rem / R, r = rem - l * R;
    return (int)(b * bstride) + (c + l) * BR + l * R + r;
}
template <int R>
__device__ __forceinline__ int addr_b(int k, int c, long long bstride)
{
    constexpr int BR = kWave * R;
    const int b = k / BR, rem = k - b * BR, l = rem / R, r = rem - l * R;
    return (int)(b * bstride) + (c - l + 63) * BR + l * R + r;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Godunov upwind update of sweep.c:89-95.  |a-b| and S*DH are float; the
// 2-D root is evaluated in double: 2*S^2*DH^2 is S^2 * (2 DH^2) (both
// products of exact values, hence the same rounding), (a-b)^2 is exact in
// double, (a+b) is a float sum, /2.0 is exact, one rounding back to float.
__device__ __forceinline__ float godunov(float a, float b, float s, float dh, double two_dh2)
{
    const float sdh = s * dh;
    const float d = a - b;
    if (fabsf(d) >= sdh) return fminf(a, b) + sdh;
    const double sd = (double)s;
    const double dd = (double)d;
    const double disc = (sd * sd) * two_dh2 - dd * dd;
    return (float)(((double)(a + b) + __builtin_sqrt(disc)) * 0.5);
}

// Block-wide double sum in a fixed order (identical in every wave).
__device__ __forceinline__ double block_sum(double v, double *red, int nwaves)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nwaves; ++i) t += red[i];
    __syncthreads();
    return t;
}

struct WaveCtx {
    int nx, ny, nb, beta, lane, active;
    long long bstride;
    float *ring;
};

__device__ __forceinline__ int sweep_steps(int nx, int nb)
{
    return ((nb - 1) * kBandOff + nx + 63 + kPF - 1) / kPF * kPF;
}

// ---------------------------------------------------------------------------
// Forward sweep (sweep.c:10-104) in direction (SX, SY) over one shot.
//   Tin    field in this sweep's family layout (A if SX==SY, else B)
//   Sf     slowness in the same layout
//   Tother written in the other family's layout   (if WOTHER)
//   Tsame  written in the same layout, may be Tin (if WSAME)
//   INIT   read the initial field (2*TIME, 0 at the source) instead of Tin
// Returns this lane's sum of (old - new) over its valid cells.
// ---------------------------------------------------------------------------
template <int R, int SX, int SY, bool INIT, bool WOTHER, bool WSAME>
__device__ double fwd_sweep(const WaveCtx &w, float dh, double two_dh2, const float *Tin,
                            const float *Sf, float *Tother, float *Tsame, int src_k, int src_c,
                            float tinit)
{
    constexpr bool FA = (SX == SY);
    constexpr int BR = kWave * R;
    const int nx = w.nx, ny = w.ny, nb = w.nb, beta = w.beta, l = w.lane;
    const int nkeys = nx + 63;
    const int gend = sweep_steps(nx, nb);
    if (!w.active) {
        for (int g0 = 0; g0 < gend; g0 += kPF) __syncthreads();
        return 0.0;
    }
    const int lag = SY > 0 ? l : 63 - l;
    const int off = (SY > 0 ? beta : nb - 1 - beta) * kBandOff;
    const long long boff = (long long)beta * w.bstride;
    const float *tband = Tin + boff;
    const float *sband = Sf + boff;
    const int k0 = beta * BR + l * R;

    bool rvalid[R], top[R], bot[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = k0 + r;
        rvalid[r] = k < ny;
        top[r] = (k == 0);
        bot[r] = (k == ny - 1);
    }
    // The edge lane needs the old value of the row across the band boundary
    // on the downstream side: below (SY>0, lane 63) or above (SY<0, lane 0).
    const int edge_lane = SY > 0 ? 63 : 0;
    const bool has_dn = SY > 0 ? (beta + 1 < nb) : (beta > 0);
    const bool has_up = SY > 0 ? (beta > 0) : (beta + 1 < nb);
    const bool do_edge = (l == edge_lane) && has_dn;
    const float *eband = tband + (SY > 0 ? w.bstride : -w.bstride);
    const int erow = SY > 0 ? (beta + 1) * BR : beta * BR - 1;
    const int ering = (SY > 0 ? beta - 1 : beta + 1);       // producer band read here
    const int ering_c = clampi(ering, 0, kMaxBands - 1);
    const bool ring_rd_lane = SY > 0 ? (l == 0) : (l == 63);
    const bool ring_wr = (SY > 0 ? (l == 63) : (l == 0)) && (SY > 0 ? (beta + 1 < nb) : (beta > 0));

    float qt[kPF][R], qs[kPF][R], qe[kPF];

#define RZ_FWD_LOAD(SLOT, TAU)                                                              \
    do {                                                                                    \
        const int tau_ = (TAU);                                                             \
        const int key_ = clampi(SX > 0 ? tau_ : nx + 62 - tau_, 0, nkeys - 1);              \
        ldv<R>(sband + key_ * BR + l * R, qs[SLOT]);                                        \
        if constexpr (INIT) {                                                               \
            const int c_ = FA ? key_ - l : key_ + l - 63;                                   \
            _Pragma("unroll") for (int r_ = 0; r_ < R; ++r_)                                \
                qt[SLOT][r_] = (k0 + r_ == src_k && c_ == src_c) ? 0.f : tinit;             \
        } else {                                                                            \
            ldv<R>(tband + key_ * BR + l * R, qt[SLOT]);                                    \
        }                                                                                   \
        if (do_edge) {                                                                      \
            const int ce_ = clampi(SX > 0 ? tau_ - lag : nx - 1 - tau_ + lag, 0, nx - 1);   \
            if constexpr (INIT) {                                                           \
                qe[SLOT] = (erow == src_k && ce_ == src_c) ? 0.f : tinit;                   \
            } else {                                                                        \
                const int ek_ = SY > 0 ? (FA ? ce_ : ce_ + 63) : (FA ? ce_ + 63 : ce_);     \
                const int eo_ = SY > 0 ? 0 : 63 * R + R - 1;                                \
                qe[SLOT] = eband[ek_ * BR + eo_];                                           \
            }                                                                               \
        }                                                                                   \
    } while (0)

#pragma unroll
    for (int u = 0; u < kPF; ++u) {
        qe[u] = 0.f;
        RZ_FWD_LOAD(u, u - off);
    }

    float pnew[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pnew[r] = tinit;
    double acc = 0.0;

    for (int g0 = 0; g0 < gend; g0 += kPF) {
#pragma unroll
        for (int u = 0; u < kPF; ++u) {
            const int un = (u + 1) % kPF;
            const int tau = g0 + u - off;
            if (tau >= 0 && tau < nkeys) {
                const int c = SX > 0 ? tau - lag : nx - 1 - tau + lag;
                const bool cvalid = (unsigned)c < (unsigned)nx;
                const bool first = SX > 0 ? (c == 0) : (c == nx - 1);
                const bool last = SX > 0 ? (c == nx - 1) : (c == 0);
                const int ckey = clampi(c, 0, nx - 1) & (kRingQ - 1);
                const float ringv = w.ring[ering_c * kRingQ + ckey];
                float up_new, dn_old;
                if (SY > 0) {
                    up_new = lane_from_above(pnew[R - 1]);
                    dn_old = lane_from_below(qt[un][0]);
                    if (l == 0) up_new = ringv;
                    if (l == 63) dn_old = qe[u];
                } else {
                    up_new = lane_from_below(pnew[0]);
                    dn_old = lane_from_above(qt[un][R - 1]);
                    if (l == 63) up_new = ringv;
                    if (l == 0) dn_old = qe[u];
                }
                (void)has_up;
                float nv[R];
                float fsum = 0.f;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = SY > 0 ? i : R - 1 - i;
                    const float self = qt[u][r];
                    const float xu = first ? self : pnew[r];
                    const float xd = last ? self : qt[un][r];
                    float yu, yd;
                    if (SY > 0) {
                        yu = (r == 0) ? up_new : nv[r > 0 ? r - 1 : 0];
                        yd = (r == R - 1) ? dn_old : qt[u][r < R - 1 ? r + 1 : r];
                        if (top[r]) yu = self;
                        if (bot[r]) yd = self;
                    } else {
                        yu = (r == R - 1) ? up_new : nv[r < R - 1 ? r + 1 : r];
                        yd = (r == 0) ? dn_old : qt[u][r > 0 ? r - 1 : 0];
                        if (bot[r]) yu = self;
                        if (top[r]) yd = self;
                    }
                    const float a = fminf(xu, xd);
                    const float b = fminf(yu, yd);
                    const float t = godunov(a, b, qs[u][r], dh, two_dh2);
                    nv[r] = fminf(self, t);
                    if (cvalid && rvalid[r]) fsum += self - nv[r];
                }
                acc += (double)fsum;
                if (cvalid) {
                    if constexpr (WSAME) {
                        const int key = SX > 0 ? tau : nx + 62 - tau;
                        stv<R>(Tsame + boff + key * BR + l * R, nv);
                    }
                    if constexpr (WOTHER) {
                        const int ko = FA ? (c - l + 63) : (c + l);
                        stv<R>(Tother + boff + ko * BR + l * R, nv);
                    }
                    if (ring_wr) w.ring[beta * kRingQ + (c & (kRingQ - 1))] = SY > 0 ? nv[R - 1] : nv[0];
                }
#pragma unroll
                for (int r = 0; r < R; ++r) pnew[r] = nv[r];
            }
            RZ_FWD_LOAD(u, tau + kPF);
        }
        __syncthreads();
    }
#undef RZ_FWD_LOAD
    (void)ring_rd_lane;
    return acc;
}

// ---------------------------------------------------------------------------
// Adjoint cell update (sweep_adj.c:33-68), float arithmetic in source order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float adj_cell(float tc, float txp, float txm, float typ, float tym,
                                          float lc, float lxm, float lxp, float lym, float lyp,
                                          float dh, float eps)
{
    const float ap = -(txp - tc) / dh;
    const float am = -(tc - txm) / dh;
    const float bp = -(typ - tc) / dh;
    const float bm = -(tc - tym) / dh;
    const float app = (ap + fabsf(ap)) * 0.5f, apm = (ap - fabsf(ap)) * 0.5f;
    const float amp = (am + fabsf(am)) * 0.5f, amm = (am - fabsf(am)) * 0.5f;
    const float bpp = (bp + fabsf(bp)) * 0.5f, bpm = (bp - fabsf(bp)) * 0.5f;
    const float bmp = (bm + fabsf(bm)) * 0.5f, bmm = (bm - fabsf(bm)) * 0.5f;
    const float lhs = (app - amm) / dh + (bpp - bmm) / dh;
    const float rhs = (amp * lxm - apm * lxp) / dh + (bmp * lym - bpm * lyp) / dh;
    const float lt = (lhs > eps) ? rhs / lhs : 0.f;
    float v = fminf(lc, lt);
    const float tmax = fmaxf(fmaxf(fmaxf(tc, typ), fmaxf(tym, txp)), txm);
    if (tmax >= kInf) v = 0.f;
    return v;
}

// ---------------------------------------------------------------------------
// Adjoint sweep (sweep_adj.c:10-77) in direction (SX, SY).
//   Tf    converged traveltime in this sweep's family layout; the sign bit
//         marks receiver cells (skipped, sweep_adj.c:30)
//   Lin   lambda in the same layout; outputs as in fwd_sweep.
// ---------------------------------------------------------------------------
template <int R, int SX, int SY, bool WOTHER, bool WSAME>
__device__ double adj_sweep(const WaveCtx &w, float dh, float eps, const float *Tf, const float *Lin,
                            float *Lother, float *Lsame)
{
    constexpr bool FA = (SX == SY);
    constexpr int BR = kWave * R;
    const int nx = w.nx, ny = w.ny, nb = w.nb, beta = w.beta, l = w.lane;
    const int nkeys = nx + 63;
    const int gend = sweep_steps(nx, nb);
    if (!w.active) {
        for (int g0 = 0; g0 < gend; g0 += kPF) __syncthreads();
        return 0.0;
    }
    const int lag = SY > 0 ? l : 63 - l;
    const int off = (SY > 0 ? beta : nb - 1 - beta) * kBandOff;
    const long long boff = (long long)beta * w.bstride;
    const float *tband = Tf + boff;
    const float *lband = Lin + boff;
    const int k0 = beta * BR + l * R;

    bool rvalid[R], frow[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = k0 + r;
        rvalid[r] = k < ny;
        frow[r] = (k == 0) || (k >= ny - 1);
    }
    // Lane 0 reads the row above from band beta-1, lane 63 the row below
    // from band beta+1 (T always; lambda on the downstream side only).
    const bool up_lane = (l == 0) && (beta > 0);
    const bool dn_lane = (l == 63) && (beta + 1 < nb);
    const bool do_edge = up_lane || dn_lane;
    const bool lam_edge = SY > 0 ? dn_lane : up_lane;
    const int ering_c = clampi(SY > 0 ? beta - 1 : beta + 1, 0, kMaxBands - 1);
    const bool ring_wr = (SY > 0 ? (l == 63) : (l == 0)) && (SY > 0 ? (beta + 1 < nb) : (beta > 0));

    float qt[kPF][R], ql[kPF][R], qet[kPF], qel[kPF];

#define RZ_ADJ_LOAD(SLOT, TAU)                                                              \
    do {                                                                                    \
        const int tau_ = (TAU);                                                             \
        const int key_ = clampi(SX > 0 ? tau_ : nx + 62 - tau_, 0, nkeys - 1);              \
        ldv<R>(tband + key_ * BR + l * R, qt[SLOT]);                                        \
        ldv<R>(lband + key_ * BR + l * R, ql[SLOT]);                                        \
        if (do_edge) {                                                                      \
            const int ce_ = clampi(SX > 0 ? tau_ - lag : nx - 1 - tau_ + lag, 0, nx - 1);   \
            long long eo_;                                                                  \
            if (l == 0) eo_ = -w.bstride + (long long)((FA ? ce_ + 63 : ce_) * BR + 63 * R + R - 1); \
            else        eo_ = w.bstride + (long long)((FA ? ce_ : ce_ + 63) * BR);         \
            qet[SLOT] = tband[eo_];                                                         \
            if (lam_edge) qel[SLOT] = lband[eo_];                                           \
        }                                                                                   \
    } while (0)

#pragma unroll
    for (int u = 0; u < kPF; ++u) {
        qet[u] = 0.f;
        qel[u] = 0.f;
        RZ_ADJ_LOAD(u, u - off);
    }

    float tprev[R], pnew[R];
#pragma unroll
    for (int r = 0; r < R; ++r) { tprev[r] = 0.f; pnew[r] = 0.f; }
    double acc = 0.0;

    for (int g0 = 0; g0 < gend; g0 += kPF) {
#pragma unroll
        for (int u = 0; u < kPF; ++u) {
            const int un = (u + 1) % kPF;
            const int tau = g0 + u - off;
            if (tau >= 0 && tau < nkeys) {
                const int c = SX > 0 ? tau - lag : nx - 1 - tau + lag;
                const bool cvalid = (unsigned)c < (unsigned)nx;
                const bool fcol = (c <= 0) || (c >= nx - 1);
                const int ckey = clampi(c, 0, nx - 1) & (kRingQ - 1);
                const float ringv = w.ring[ering_c * kRingQ + ckey];
                // T of the rows across the lane boundary, at column c
                float t_above, t_below;
                if (SY > 0) {   // lane l-1 is one step ahead, lane l+1 one step behind
                    t_above = lane_from_above(tprev[R - 1]);
                    t_below = lane_from_below(qt[un][0]);
                } else {        // lane l-1 behind, lane l+1 ahead
                    t_above = lane_from_above(qt[un][R - 1]);
                    t_below = lane_from_below(tprev[0]);
                }
                if (l == 0) t_above = qet[u];
                if (l == 63) t_below = qet[u];
                // lambda across the lane boundary: new upstream, old downstream
                float l_up, l_dn;   // up = row k-1, dn = row k+1 (for r = 0 / r = R-1)
                if (SY > 0) {
                    l_up = lane_from_above(pnew[R - 1]);
                    l_dn = lane_from_below(ql[un][0]);
                    if (l == 0) l_up = ringv;
                    if (l == 63) l_dn = qel[u];
                } else {
                    l_up = lane_from_above(ql[un][R - 1]);
                    l_dn = lane_from_below(pnew[0]);
                    if (l == 0) l_up = qel[u];
                    if (l == 63) l_dn = ringv;
                }
                float nv[R];
                float fsum = 0.f;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = SY > 0 ? i : R - 1 - i;
                    const float traw = qt[u][r];
                    const bool rec = __float_as_uint(traw) >> 31;
                    const float tc = fabsf(traw);
                    const float tup = fabsf(SX > 0 ? tprev[r] : qt[un][r]);   // T(k, c-1)
                    const float tdn = fabsf(SX > 0 ? qt[un][r] : tprev[r]);   // T(k, c+1)
                    const float tym = fabsf(r == 0 ? t_above : qt[u][r > 0 ? r - 1 : 0]);
                    const float typ = fabsf(r == R - 1 ? t_below : qt[u][r < R - 1 ? r + 1 : r]);
                    const float lc = ql[u][r];
                    const float lxm = SX > 0 ? pnew[r] : ql[un][r];            // lambda(k, c-1)
                    const float lxp = SX > 0 ? ql[un][r] : pnew[r];            // lambda(k, c+1)
                    float lym, lyp;
                    if (SY > 0) {
                        lym = (r == 0) ? l_up : nv[r > 0 ? r - 1 : 0];
                        lyp = (r == R - 1) ? l_dn : ql[u][r < R - 1 ? r + 1 : r];
                    } else {
                        lym = (r == 0) ? l_up : ql[u][r > 0 ? r - 1 : 0];
                        lyp = (r == R - 1) ? l_dn : nv[r < R - 1 ? r + 1 : r];
                    }
                    const float v = adj_cell(tc, tdn, tup, typ, tym, lc, lxm, lxp, lym, lyp, dh, eps);
                    const bool keep = !cvalid || fcol || frow[r] || rec;
                    nv[r] = keep ? lc : v;
                    if (cvalid && rvalid[r]) fsum += lc - nv[r];
                }
                acc += (double)fsum;
                if (cvalid) {
                    if constexpr (WSAME) {
                        const int key = SX > 0 ? tau : nx + 62 - tau;
                        stv<R>(Lsame + boff + key * BR + l * R, nv);
                    }
                    if constexpr (WOTHER) {
                        const int ko = FA ? (c - l + 63) : (c + l);
                        stv<R>(Lother + boff + ko * BR + l * R, nv);
                    }
                    if (ring_wr) w.ring[beta * kRingQ + (c & (kRingQ - 1))] = SY > 0 ? nv[R - 1] : nv[0];
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    pnew[r] = nv[r];
                    tprev[r] = qt[u][r];
                }
            }
            RZ_ADJ_LOAD(u, tau + kPF);
        }
        __syncthreads();
    }
#undef RZ_ADJ_LOAD
    return acc;
}

// ---------------------------------------------------------------------------
// Forward kernel: one workgroup per shot (eikonal.c:10-76 + calc_FA.c).
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kMaxThreads) void k_forward(const ModelDev *__restrict__ mds,
                                                         const ShotDev *__restrict__ shots, Slots slots,
                                                         const long long *__restrict__ rec_off,
                                                         const int *__restrict__ shot_base, ShotOut out)
{
    __shared__ float ring[kMaxBands * kRingQ];
    __shared__ double red[kMaxBands];
    const ShotDev sh = shots[blockIdx.x];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    WaveCtx w;
    w.nx = md.nx; w.ny = md.ny; w.nb = md.nb;
    w.beta = threadIdx.x >> 6; w.lane = threadIdx.x & 63;
    w.active = w.beta < md.nb;
    w.bstride = md.bstride;
    w.ring = ring;
    const long long so = (long long)sh.slot * slots.slot_stride;
    float *TA = slots.TA + so, *TB = slots.TB + so;
    const float tinit = (float)(2.0 * (double)md.time);          // zero_travel.c:22
    float lnorm = (float)(10.0 * (double)md.time);                // eikonal.c:39
    const float denom = (float)(md.nx * md.ny);
    int it = 0;
    while (lnorm >= md.ttnorm) {                                  // eikonal.c:44
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        double acc;
        if (it == 0)
            acc = fwd_sweep<R, +1, +1, true, true, false>(w, md.dh, md.two_dh2, TA, md.SA, TB, nullptr,
                                                          sh.src_k, sh.src_c, tinit);
        else
            acc = fwd_sweep<R, +1, +1, false, true, false>(w, md.dh, md.two_dh2, TA, md.SA, TB, nullptr,
                                                           sh.src_k, sh.src_c, tinit);
        acc += fwd_sweep<R, -1, +1, false, true, false>(w, md.dh, md.two_dh2, TB, md.SB, TA, nullptr,
                                                        sh.src_k, sh.src_c, tinit);
        acc += fwd_sweep<R, -1, -1, false, true, false>(w, md.dh, md.two_dh2, TA, md.SA, TB, nullptr,
                                                        sh.src_k, sh.src_c, tinit);
        acc += fwd_sweep<R, +1, -1, false, true, true>(w, md.dh, md.two_dh2, TB, md.SB, TA, TB,
                                                       sh.src_k, sh.src_c, tinit);
        const double tot = block_sum(acc, red, nwaves);
        lnorm = (float)tot / denom;                               // norm1.c:27
        ++it;
    }
    const int nfield = md.nb * (int)md.bstride;
    if (it == 0) {  // loop never entered: the field is the initial one in both layouts
        for (int p = threadIdx.x; p < nfield; p += blockDim.x) {
            const int b = p / (int)md.bstride, q = p - b * (int)md.bstride;
            const int key = q / (kWave * R), wv = q - key * (kWave * R), l = wv / R, r = wv - l * R;
            const int k = b * kWave * R + l * R + r;
            TA[p] = (k == sh.src_k && key - l == sh.src_c) ? 0.f : tinit;
            TB[p] = (k == sh.src_k && key + l - 63 == sh.src_c) ? 0.f : tinit;
        }
        __syncthreads();
    }
    // receiver times and the slope neighbour of calc_FA_res.c:52-69
    const long long ob = rec_off[sh.model] + (long long)sh.shot * md.ntr;
    for (int i = threadIdx.x; i < md.ntr; i += blockDim.x) {
        const int k = md.rec_k[i], c = md.rec_c[i];
        const int kn = (k <= md.ny - 2) ? k + 1 : (k == md.ny - 1 ? k - 1 : k);
        out.tmod[ob + i] = TA[addr_a<R>(k, c, md.bstride)];
        out.tnb[ob + i] = TA[addr_a<R>(kn, c, md.bstride)];
    }
    __syncthreads();
    // mark interior receiver cells for the adjoint (sign bit; T >= +0 everywhere)
    for (int u = threadIdx.x; u < md.nuq; u += blockDim.x) {
        const int k = md.uq_k[u], c = md.uq_c[u];
        const int pa = addr_a<R>(k, c, md.bstride), pb = addr_b<R>(k, c, md.bstride);
        TA[pa] = __uint_as_float(__float_as_uint(TA[pa]) | 0x80000000u);
        TB[pb] = __uint_as_float(__float_as_uint(TB[pb]) | 0x80000000u);
    }
    if (threadIdx.x == 0) out.it_f[shot_base[sh.model] + sh.shot] = it;
}

// ---------------------------------------------------------------------------
// Adjoint kernel: one workgroup per shot (eikonal_adj.c:10-59 with
// init_travel_adj.c as an O(ntr) scatter).
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kMaxThreads) void k_adjoint(const ModelDev *__restrict__ mds,
                                                         const ShotDev *__restrict__ shots, Slots slots,
                                                         const long long *__restrict__ rec_off,
                                                         const int *__restrict__ shot_base,
                                                         const float *__restrict__ tres, ShotOut out)
{
    constexpr int BR = kWave * R;
    __shared__ float ring[kMaxBands * kRingQ];
    __shared__ double red[kMaxBands];
    const ShotDev sh = shots[blockIdx.x];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    WaveCtx w;
    w.nx = md.nx; w.ny = md.ny; w.nb = md.nb;
    w.beta = threadIdx.x >> 6; w.lane = threadIdx.x & 63;
    w.active = w.beta < md.nb;
    w.bstride = md.bstride;
    w.ring = ring;
    const long long so = (long long)sh.slot * slots.slot_stride;
    const float *TA = slots.TA + so, *TB = slots.TB + so;
    float *LA = slots.LA + so, *LB = slots.LB + so;
    // init_travel_adj.c:22-58: 0 on the frame, 10*TIME inside, tres at receivers
    const float big = (float)(10.0 * (double)md.time);
    const int bs = (int)md.bstride;
    for (int b = 0; b < md.nb; ++b)
        for (int q = threadIdx.x; q < bs; q += blockDim.x) {
            const int key = q / BR, wv = q - key * BR, l = wv / R, r = wv - l * R;
            const int k = b * BR + l * R + r, c = key - l;
            const bool inner = k > 0 && k < md.ny - 1 && c > 0 && c < md.nx - 1;
            LA[b * bs + q] = inner ? big : 0.f;
        }
    __syncthreads();
    const long long ob = rec_off[sh.model] + (long long)sh.shot * md.ntr;
    for (int u = threadIdx.x; u < md.nuq; u += blockDim.x)
        LA[addr_a<R>(md.uq_k[u], md.uq_c[u], md.bstride)] = tres[ob + md.uq_r[u]];
    __syncthreads();

    float lnorm = big;                                            // eikonal_adj.c:34
    const float denom = (float)(md.nx * md.ny);
    int it = 0;
    while (lnorm >= md.ttnorm) {                                  // eikonal_adj.c:37
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        double acc = adj_sweep<R, +1, +1, true, false>(w, md.dh, md.eps_adj, TA, LA, LB, nullptr);
        acc += adj_sweep<R, +1, -1, false, true>(w, md.dh, md.eps_adj, TB, LB, nullptr, LB);
        acc += adj_sweep<R, -1, +1, true, false>(w, md.dh, md.eps_adj, TB, LB, LA, nullptr);
        acc += adj_sweep<R, -1, -1, false, true>(w, md.dh, md.eps_adj, TA, LA, nullptr, LA);
        const double tot = block_sum(acc, red, nwaves);
        lnorm = (float)tot / denom;
        ++it;
    }
    if (threadIdx.x == 0) out.it_a[shot_base[sh.model] + sh.shot] = it;
}

// ---------------------------------------------------------------------------
// Residuals, adjoint sources and misfit (calc_FA_res.c:10-77) for shots
// [s0, s1) of one model, one workgroup.  carry (ntr) is the reference's
// Tres array: a muted pick keeps the previous value.  l2_sum accumulates
// the per-shot sums in shot order (grad_obj.c:42).
// ---------------------------------------------------------------------------
__global__ void k_residual(const ModelDev *__restrict__ mds, int model, int ntr, int s0, int s1,
                           long long rec_off, const float *__restrict__ tmod, const float *__restrict__ tnb,
                           const float *__restrict__ tobs, float *carry, float *tres, double *l2_shot,
                           double *l2_sum)
{
    const ModelDev md = mds[model];
    for (int i = threadIdx.x; i < ntr; i += blockDim.x) {
        float cr = carry[i];
        const int k = md.rec_k[i];
        for (int s = s0; s < s1; ++s) {
            const long long x = rec_off + (long long)s * ntr + i;
            const float to = tobs[x];
            if (to >= 0.f) {
                const float tm = tmod[x];
                float r = tm - to;
                float g = 0.f;
                bool have = true;
                if (k <= md.ny - 2) g = (tnb[x] - tm) / md.dh;
                else if (k == md.ny - 1) g = (tm - tnb[x]) / md.dh;
                else have = false;
                if (have) r = (fabsf(g) > 0.f) ? r / g : 0.f;
                cr = r;
            }
            tres[x] = cr;
        }
        carry[i] = cr;
    }
    __syncthreads();
    for (int s = s0 + (int)threadIdx.x; s < s1; s += blockDim.x) {
        double l2 = 0.0;
        for (int i = 0; i < ntr; ++i) {
            const long long x = rec_off + (long long)s * ntr + i;
            const float to = tobs[x];
            if (to >= 0.f) {
                const float r = tmod[x] - to;
                l2 += (double)(r * r);
            }
        }
        l2_shot[s] = l2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = *l2_sum;
        for (int s = s0; s < s1; ++s) t += l2_shot[s];
        *l2_sum = t;
    }
}

// ---------------------------------------------------------------------------
// Gradient assembly and shot sum (ass_grad.c:20-46 + cp_grad_frame.c):
// grad(k,c) += (float)(lambda_s(k',c') * S(k',c')^3) over the given slots in
// order, (k',c') = (k,c) clamped to rows 2..NY-3 and columns 2..NX-3.
// ---------------------------------------------------------------------------
template <int R>
__global__ void k_gradient(const ModelDev *__restrict__ mds, int model, const int *__restrict__ slot_of,
                           int nsh, Slots slots, float *grad, int accumulate)
{
    constexpr int BR = kWave * R;
    const ModelDev md = mds[model];
    const long long total = (long long)md.nb * md.bstride;
    const int bs = (int)md.bstride;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < total;
         p += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(p / bs), q = (int)(p - (long long)b * bs);
        const int key = q / BR, wv = q - key * BR, l = wv / R, r = wv - l * R;
        const int k = b * BR + l * R + r, c = key - l;
        if (k >= md.ny || c < 0 || c >= md.nx) continue;
        const int kk = clampi(k, 2, md.ny - 3), cc = clampi(c, 2, md.nx - 3);
        const int pos = (kk == k && cc == c) ? (int)p : addr_a<R>(kk, cc, md.bstride);
        const double s = (double)md.SA[pos];
        const double s3 = s * s * s;
        float g = accumulate ? grad[(long long)k * md.nx + c] : 0.f;
        for (int j = 0; j < nsh; ++j) {
            const float lam = slots.LA[(long long)slot_of[j] * slots.slot_stride + pos];
            g += (float)((double)lam * s3);
        }
        grad[(long long)k * md.nx + c] = g;
    }
}

// Row-major slowness -> layouts A and B; flags non-finite or non-positive S.
template <int R>
__global__ void k_prepare_slowness(const ModelDev *__restrict__ mds, int model, float *SA, float *SB,
                                   int *err_flag)
{
    constexpr int BR = kWave * R;
    const ModelDev md = mds[model];
    const long long total = (long long)md.nb * md.bstride;
    const int bs = (int)md.bstride;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < total;
         p += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(p / bs), q = (int)(p - (long long)b * bs);
        const int key = q / BR, wv = q - key * BR, l = wv / R, r = wv - l * R;
        const int k = b * BR + l * R + r;
        const int ca = key - l, cb = key + l - 63;
        float va = 1.f, vb = 1.f;
        if (k < md.ny && ca >= 0 && ca < md.nx) {
            va = md.S[(long long)k * md.nx + ca];
            if (!(va > 0.f) || !(va < INFINITY)) atomicOr(err_flag, 1);
        }
        if (k < md.ny && cb >= 0 && cb < md.nx) vb = md.S[(long long)k * md.nx + cb];
        SA[p] = va;
        SB[p] = vb;
    }
}

// Layout-A field -> row-major (optionally clearing the receiver sign bits).
template <int R>
__global__ void k_field_to_rowmajor(int nx, int ny, long long bstride, const float *field, float *out,
                                    int strip)
{
    const long long n = (long long)nx * ny;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nx), c = (int)(i - (long long)k * nx);
        float v = field[addr_a<R>(k, c, bstride)];
        if (strip) v = fabsf(v);
        out[i] = v;
    }
}

// Row-major -> layout A (family 0) or B (family 1).
template <int R>
__global__ void k_rowmajor_to_field(int nx, int ny, long long bstride, const float *in, float *field,
                                    int family)
{
    const long long n = (long long)nx * ny;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nx), c = (int)(i - (long long)k * nx);
        const int p = family ? addr_b<R>(k, c, bstride) : addr_a<R>(k, c, bstride);
        field[p] = in[i];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define RZ_DISPATCH_R(R_, CALL)                                    \
    switch (R_) {                                                  \
    case 1: { constexpr int RR = 1; CALL; } break;                 \
    case 2: { constexpr int RR = 2; CALL; } break;                 \
    case 4: { constexpr int RR = 4; CALL; } break;                 \
    case 8: { constexpr int RR = 8; CALL; } break;                 \
    default: return hipErrorInvalidValue;                          \
    }

static int grid_for(long long n, int block)
{
    long long g = (n + block - 1) / block;
    if (g > 65535) g = 65535;
    if (g < 1) g = 1;
    return (int)g;
}

hipError_t launch_prepare_slowness(const ModelDev *mh, int model, const ModelDev *md, float *SA, float *SB,
                                   int *err_flag, hipStream_t st)
{
    const long long n = (long long)mh->nb * mh->bstride;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_prepare_slowness<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st,
                                            md, model, SA, SB, err_flag));
    return hipGetLastError();
}

hipError_t launch_forward(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots, Slots slots,
                          const long long *rec_off, const int *shot_base, ShotOut out, hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    RZ_DISPATCH_R(R, hipLaunchKernelGGL(k_forward<RR>, dim3(nshots), dim3(nb_max * kWave), 0, st, md, shots,
                                        slots, rec_off, shot_base, out));
    return hipGetLastError();
}

hipError_t launch_adjoint(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots, Slots slots,
                          const long long *rec_off, const int *shot_base, const float *tres, ShotOut out,
                          hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    RZ_DISPATCH_R(R, hipLaunchKernelGGL(k_adjoint<RR>, dim3(nshots), dim3(nb_max * kWave), 0, st, md, shots,
                                        slots, rec_off, shot_base, tres, out));
    return hipGetLastError();
}

hipError_t launch_residual(const ModelDev *md, int model, int ntr, int s0, int s1, long long rec_off,
                           const float *tmod, const float *tnb, const float *tobs, float *carry, float *tres,
                           double *l2_shot, double *l2_sum, hipStream_t st)
{
    hipLaunchKernelGGL(k_residual, dim3(1), dim3(256), 0, st, md, model, ntr, s0, s1, rec_off, tmod, tnb, tobs,
                       carry, tres, l2_shot, l2_sum);
    return hipGetLastError();
}

hipError_t launch_gradient(const ModelDev *mh, int model, const ModelDev *md, const int *slot_of, int nsh,
                           Slots slots, float *grad, int accumulate, hipStream_t st)
{
    const long long n = (long long)mh->nb * mh->bstride;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_gradient<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st, md, model,
                                            slot_of, nsh, slots, grad, accumulate));
    return hipGetLastError();
}

hipError_t launch_field_to_rowmajor(const ModelDev *mh, const float *field, float *out, int strip, hipStream_t st)
{
    const long long n = (long long)mh->nx * mh->ny;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_field_to_rowmajor<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st,
                                            mh->nx, mh->ny, mh->bstride, field, out, strip));
    return hipGetLastError();
}

hipError_t launch_rowmajor_to_field(const ModelDev *mh, const float *in, float *field, int family, hipStream_t st)
{
    const long long n = (long long)mh->nx * mh->ny;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_rowmajor_to_field<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st,
                                            mh->nx, mh->ny, mh->bstride, in, field, family));
    return hipGetLastError();
}

}  // namespace rz
