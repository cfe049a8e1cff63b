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
// -ffp-contract=off): see godunov_sel(), qdh() and adj_sweep().  Any topological order
// of a Gauss-Seidel sweep's dependency DAG gives the same bits, so each
// sweep here equals the reference's sweep bit for bit.
#include <climits>

#include "rz_internal.h"

namespace rz {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// Field pointers reach the sweeps through structs in memory, where the
// compiler cannot infer their address space and would emit FLAT accesses;
// a wait on a FLAT load needs vmcnt(0), which would also wait for every
// prefetch in flight.  All field accesses go through global pointers.
typedef __attribute__((address_space(1))) float gfloat;
template <typename T>
__device__ __forceinline__ const gfloat *gp(const T *p)
{
    return (const gfloat *)(p);
}
template <typename T>
__device__ __forceinline__ gfloat *gpw(T *p)
{
    return (gfloat *)(p);
}
__device__ __forceinline__ float lane_from_below(float v)  // lane l gets lane l+1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_from_above(float v)  // lane l gets lane l-1
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, false));
}

typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) vf2 gvf2;
typedef __attribute__((address_space(1))) vf4 gvf4;

// Buffer loads over a whole field: a raw buffer descriptor (gfx9 format,
// dword 3 = 0x00020000) with the field's size, the lane's byte offset in a
// VGPR and the wave-uniform part in an SGPR, which keeps per-step address
// arithmetic 32-bit and scalar.  Callers pass in-range offsets only.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float ldb1(__amdgpu_buffer_rsrc_t r, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, 0, soff, 0));
}

template <int R>
__device__ __forceinline__ void ldb(__amdgpu_buffer_rsrc_t r, int voff, int soff, float (&o)[R])
{
    if constexpr (R == 1) {
        o[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else if constexpr (R == 2) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
        o[0] = __uint_as_float(t[0]); o[1] = __uint_as_float(t[1]);
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            const auto t = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 4 * i, soff, 0);
            o[i] = __uint_as_float(t[0]); o[i + 1] = __uint_as_float(t[1]);
            o[i + 2] = __uint_as_float(t[2]); o[i + 3] = __uint_as_float(t[3]);
        }
    }
}

template <int R>
__device__ __forceinline__ void ldv(const gfloat *p, float (&o)[R])
{
    if constexpr (R == 1) {
        o[0] = p[0];
    } else if constexpr (R == 2) {
        const vf2 t = *reinterpret_cast<const gvf2 *>(p);
        o[0] = t.x; o[1] = t.y;
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            const vf4 t = *reinterpret_cast<const gvf4 *>(p + i);
            o[i] = t.x; o[i + 1] = t.y; o[i + 2] = t.z; o[i + 3] = t.w;
        }
    }
}

// Field stores.  Up to R = 4 they are nontemporal: measured 0.2-0.4 ms faster
// per 1001 x 3001, 200-shot evaluation (R = 4), while at R = 8 they cost
// 4 % on the 2000 x 6000 one, whose sweeps keep the default policy.
template <int R>
__device__ __forceinline__ void stv(gfloat *p, const float (&v)[R])
{
    if constexpr (R == 1) {
        __builtin_nontemporal_store(v[0], p);
    } else if constexpr (R == 2) {
        vf2 t;
        t.x = v[0]; t.y = v[1];
        __builtin_nontemporal_store(t, reinterpret_cast<gvf2 *>(p));
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            vf4 t;
            t.x = v[i]; t.y = v[i + 1]; t.z = v[i + 2]; t.w = v[i + 3];
            if constexpr (R == 4) __builtin_nontemporal_store(t, reinterpret_cast<gvf4 *>(p + i));
            else *reinterpret_cast<gvf4 *>(p + i) = t;
        }
    }
}

// R floats of one lane in LDS (a wave's staging line, see Stage).
typedef __attribute__((address_space(3))) float lfloat;
typedef __attribute__((address_space(3))) vf2 lvf2;
typedef __attribute__((address_space(3))) vf4 lvf4;

template <int R>
__device__ __forceinline__ void ldl(const lfloat *p, float (&o)[R])
{
    if constexpr (R == 1) {
        o[0] = p[0];
    } else if constexpr (R == 2) {
        const vf2 t = *reinterpret_cast<const lvf2 *>(p);
        o[0] = t.x; o[1] = t.y;
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            const vf4 t = *reinterpret_cast<const lvf4 *>(p + i);
            o[i] = t.x; o[i + 1] = t.y; o[i + 2] = t.z; o[i + 3] = t.w;
        }
    }
}

template <int R>
__device__ __forceinline__ void stl(lfloat *p, const float (&v)[R])
{
    if constexpr (R == 1) {
        p[0] = v[0];
    } else if constexpr (R == 2) {
        vf2 t;
        t.x = v[0]; t.y = v[1];
        *reinterpret_cast<lvf2 *>(p) = t;
    } else {
#pragma unroll
        for (int i = 0; i < R; i += 4) {
            vf4 t;
            t.x = v[i]; t.y = v[i + 1]; t.z = v[i + 2]; t.w = v[i + 3];
            *reinterpret_cast<lvf4 *>(p + i) = t;
        }
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

// min with the semantics of the reference's fminf (glibc x86-64: minss,
// i.e. x < y ? x : y for non-NaN arguments, the second one on ties, so also
// on +0/-0).  A compare and a select; no NaN canonicalisation.
__device__ __forceinline__ float mn(float x, float y) { return x < y ? x : y; }

// sqrt of a positive normal double by v_rsq_f64 and two Newton steps; within
// a few ulp of the correctly rounded root (correctly rounded on every one of
// 6.7e7 sampled inputs on gfx950).  Callers only use it behind a rounding
// guard, so exactness never depends on that observation.
__device__ __forceinline__ double sqrt_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// Godunov upwind update of sweep.c:89-95 (godunov_sel below).  |a-b| and
// S*DH are float; the 2-D root is evaluated in double: 2*S^2*DH^2 is
// S^2 * (2 DH^2) (both products of exact values, hence the same rounding),
// (a-b)^2 is exact in double, (a+b) is a float sum, /2.0 is exact, one
// rounding back to float.  The root comes from sqrt_nr(); the float rounding
// of the half-sum can only differ from the one with the correctly rounded
// root when the double lies within a few ulp of a float tie (low 29
// mantissa bits near 0x10000000).  Such cells (~1 in 1.6e7) take the
// correctly rounded sqrt.

// Minimum of two forward traveltimes.  They are never NaN and never -0
// (every value is a sum of non-negative terms starting from +0), so
// v_min_f32 returns exactly x < y ? x : y, the reference's fminf; the
// inline form keeps the compiler from canonicalising the operands first.
__device__ __forceinline__ float mn_t(float x, float y)
{
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
}

// Godunov update of sweep.c:89-95 without branches: both the 1-D and the
// 2-D value are formed and one is selected (lanes of a wave mostly
// disagree, so a branch would run both anyway).  The 2-D root comes from
// sqrt_nr(); `need` flags a 2-D cell whose half-sum lies near a float
// rounding tie, where only the correctly rounded root (EXACT) is safe.
template <bool EXACT>
__device__ __forceinline__ float godunov_sel(float a, float b, float s, float dh, double two_dh2, bool &need)
{
    const float sdh = s * dh;
    const float d = a - b;
    const bool one = fabsf(d) >= sdh;
    const float t1 = mn_t(a, b) + sdh;
    const double sd = (double)s;
    const double dd = (double)d;
    double disc = __builtin_fma(-dd, dd, (sd * sd) * two_dh2);
    const double ab = (double)(a + b);
    double root;
    if constexpr (EXACT) {
        asm volatile("" : "+v"(disc));   // keeps the exact root inside the rare branch
        root = __builtin_sqrt(disc);
    } else {
        root = sqrt_nr(disc);
    }
    const double half = (ab + root) * 0.5;
    if constexpr (!EXACT) {
        constexpr unsigned kGuard = 16;
        const unsigned lo = (unsigned)__double_as_longlong(half) & 0x1FFFFFFFu;
        need |= !one && (lo - (0x10000000u - kGuard) <= 2 * kGuard);
    }
    return one ? t1 : (float)half;
}

// One step of a forward sweep for the R cells of a lane (rows in sweep
// order).  cur/nxt: old values at this and the next key (nxt[r] is the
// downwind x-neighbour), prev: this lane's new values of the previous
// step (upwind x-neighbour), up_new/dn_old: rows across the lane boundary.
template <int R, int SY, bool EXACT>
__device__ __forceinline__ void fwd_cells(const float (&cur)[R], const float (&nxt)[R], const float (&sl)[R],
                                          const float (&prev)[R], float up_new, float dn_old, bool first,
                                          bool last, const bool (&top)[R], const bool (&bot)[R],
                                          const bool (&rvalid)[R], bool cvalid, float dh, double two_dh2,
                                          float (&nv)[R], float &fsum, bool &need)
{
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = SY > 0 ? i : R - 1 - i;
        const float self = cur[r];
        const float xu = first ? self : prev[r];
        const float xd = last ? self : nxt[r];
        float yu, yd;
        if (SY > 0) {
            yu = (r == 0) ? up_new : nv[r > 0 ? r - 1 : 0];
            yd = (r == R - 1) ? dn_old : cur[r < R - 1 ? r + 1 : r];
            if (top[r]) yu = self;
            if (bot[r]) yd = self;
        } else {
            yu = (r == R - 1) ? up_new : nv[r < R - 1 ? r + 1 : r];
            yd = (r == 0) ? dn_old : cur[r > 0 ? r - 1 : 0];
            if (bot[r]) yu = self;
            if (top[r]) yd = self;
        }
        const float a = mn_t(xu, xd);
        const float b = mn_t(yu, yd);
        const float t = godunov_sel<EXACT>(a, b, sl[r], dh, two_dh2, need);
        nv[r] = mn_t(self, t);
        if (cvalid && rvalid[r]) fsum += self - nv[r];
    }
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

// What a forward sweep returns for its lane: the sum of decreases and an
// upper bound on the cells it changed (R per step with any change).
struct Dec {
    double sum, cnt;
};

// block_sum of both parts.
__device__ __forceinline__ Dec block_sum2(Dec v, double *red, int nwaves)
{
    Dec r;
    r.sum = block_sum(v.sum, red, nwaves);
    r.cnt = block_sum(v.cnt, red, nwaves);
    return r;
}

// Barrier between step epochs of a sweep.  Within a sweep no wave reads
// global memory another wave writes: new values cross bands through the LDS
// ring, old values through loads that have already been consumed (waited
// for) before the barrier.  So only LDS needs ordering here, and the
// prefetched loads stay in flight.  Sweep boundaries use __syncthreads().
__device__ __forceinline__ void epoch_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Epoch barrier number i of a sweep pair.  Every eighth one is a full
// workgroup release/acquire: it waits for the wave's outstanding memory
// operations, so a value one sweep of a pair stores is visible to the other
// sweep's loads once such a barrier lies between them (pair_delay()).
__device__ __forceinline__ void epoch_sync(int i)
{
    if ((i & 7) == 7) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        epoch_barrier();
    }
}

// Epochs by which the trailing sweep of a pair follows the leading one.  Both
// run along x in the same direction and along y in opposite directions.  The
// trailing sweep loads cell (band p, lane l, column q) at step
// q + (63 - lag) + (nb-1-p)*kBandOff + kPF*D - kPF and needs the leading
// sweep's values up to column q+1, stored at q + 1 + lag + p*kBandOff; the
// difference is at least kPF*D - kBandOff*nb.  It must exceed one full-fence
// interval plus one epoch (8*kPF + kPF steps), plus `hold`, the steps by
// which a sweep may hold a value before storing it (staged stores).
__device__ __forceinline__ int pair_delay(int nb, int hold = 0)
{
    return (kBandOff * nb + 9 * kPF + hold + kPF - 1) / kPF;
}

struct WaveCtx {
    int nx, ny, nb, beta, lane, active, group;
    long long bstride;
    float *ring;
};

// A value every lane of the wave holds, as a scalar.  The sweeps are
// separate functions, whose arguments arrive in vector registers; without
// this the compiler computes every index derived from them per lane.
template <typename T>
__device__ __forceinline__ T uni(T v)
{
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    int u[sizeof(T) / 4];
    __builtin_memcpy(u, &v, sizeof(T));
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) u[k] = __builtin_amdgcn_readfirstlane(u[k]);
    __builtin_memcpy(&v, u, sizeof(T));
    return v;
}

__device__ __forceinline__ WaveCtx uni_ctx(const WaveCtx &w)
{
    WaveCtx u = uni(w);
    u.lane = w.lane;
    return u;
}

// norm1.c:17-31 in the reference's order: the float sum s, x outer and y
// inner, of the terms |Tend - Tstart| (a float difference) over the grid;
// Tstart == nullptr stands for the initial forward field (tinit, 0 at the
// source).  Both fields are in layout A.
//
// The sum is exact but not serial.  While s stays in one binade
// [2^e, 2^(e+1)) (e >= -126, where the floats are the multiples of
// u_e = 2^(e-23)), s + d is s plus d rounded to a multiple of u_e, whatever
// s is, except where d/u_e lies exactly half-way between two integers: such
// a tie rounds to the even neighbour, so it depends on the parity of s/u_e
// only.  A run of terms is therefore a function of the incoming parity, to
// (integer increment, outgoing parity), and those functions compose.  A
// thread takes kNormPer consecutive rows of one column (the order is padded
// with zero terms to whole blocks of rows per column: adding +0 changes no
// float sum), so that it loads whole R-row groups; per chunk of blockDim
// such blocks the workgroup scans the threads' functions, and each thread
// walks its terms from its exact incoming sum.  The first term that takes
// the sum to 2^(e+1) or beyond is added as a float by its thread, and the
// walk restarts after it in the new binade: a sum crosses about as many
// binades as it spans (a few dozen).  Increments stay below 2^24 within a
// binade, so 32-bit unsigned arithmetic is exact up to the first such term;
// what wraps after it is never used.
constexpr int kNormPer = 16;

// d in units of 2^(e-23), packed: bits 0-23 the floor, bit 29 a tie (the
// rest is exactly one half), bit 30 above one half, bit 31 the term alone
// takes a sum of this binade out of it
constexpr unsigned kIncTie = 1u << 29, kIncUp = 1u << 30, kIncBig = 1u << 31;

__device__ __forceinline__ unsigned term_inc(float d, int e, bool s_normal)
{
    const unsigned b = __float_as_uint(d);
    const int E = (b >> 23) & 0xFF;
    const unsigned M = (b & 0x7FFFFFu) | (E ? 0x800000u : 0u);
    if (M == 0) return 0;
    const int sh = (E ? E : 1) - 150 - e + 23;
    // a normal d of at least u_e: M * 2^sh >= 2^23 units, and s >= 2^23 units
    // when normal, so the sum leaves the binade
    if (sh >= 24 || (sh >= 0 && (s_normal || (M << sh) >= (1u << 24)))) return kIncBig;
    if (sh >= 0) return M << sh;
    const int k = -sh;
    if (k > 25) return 0;                 // below a quarter unit: rounds to 0
    const unsigned rem = M & ((1u << k) - 1u), half = 1u << (k - 1);
    return (M >> k) | (rem == half ? kIncTie : 0u) | (rem > half ? kIncUp : 0u);
}

// the increment of a packed term on a running sum of parity p (ties to even)
__device__ __forceinline__ unsigned inc_at(unsigned v, unsigned p)
{
    const unsigned f = v & 0xFFFFFFu;
    return f + (((v & kIncUp) || ((v & kIncTie) && ((p + f) & 1u))) ? 1u : 0u);
}

// A run of terms as a function of the incoming parity p: increment t[p];
// outgoing parity q[p] in bit p of q.
struct ParFn {
    unsigned t0, t1, q;
};

// first F, then G
__device__ __forceinline__ ParFn compose(const ParFn &F, const ParFn &G)
{
    const unsigned f0 = F.q & 1u, f1 = (F.q >> 1) & 1u;
    ParFn H;
    H.t0 = F.t0 + (f0 ? G.t1 : G.t0);
    H.t1 = F.t1 + (f1 ? G.t1 : G.t0);
    H.q = ((G.q >> f0) & 1u) | (((G.q >> f1) & 1u) << 1);
    return H;
}

// Workgroup-wide exclusive scan of the threads' functions in thread order;
// *all gets the composition of every thread's.
__device__ ParFn block_exscan_fn(const ParFn &mine, ParFn *all)
{
    __shared__ ParFn wtot[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const ParFn id{0u, 0u, 2u};
    ParFn inc = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        ParFn y;
        y.t0 = __shfl_up(inc.t0, o, 64);
        y.t1 = __shfl_up(inc.t1, o, 64);
        y.q = __shfl_up(inc.q, o, 64);
        if (lane >= o) inc = compose(y, inc);
    }
    ParFn ex;
    ex.t0 = __shfl_up(inc.t0, 1, 64);
    ex.t1 = __shfl_up(inc.t1, 1, 64);
    ex.q = __shfl_up(inc.q, 1, 64);
    if (lane == 0) ex = id;
    if (lane == 63) wtot[wave] = inc;
    __syncthreads();
    ParFn base = id, tot = id;
    for (int i = 0; i < nw; ++i) {
        if (i < wave) base = compose(base, wtot[i]);
        tot = compose(tot, wtot[i]);
    }
    __syncthreads();
    *all = tot;
    return compose(base, ex);
}

__device__ int block_min(int v)
{
    __shared__ int m_s;
    if (threadIdx.x == 0) m_s = INT_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) atomicMin(&m_s, v);
    __syncthreads();
    const int r = m_s;
    __syncthreads();
    return r;
}

template <int R>
__device__ __attribute__((noinline)) float ordered_norm1(int nx_, int ny_, long long bstride_, const float *Tend_,
                                                         const float *Tstart__, float tinit_, int src_k_, int src_c_,
                                                         int trans_)
{
    constexpr int BR = kWave * R;
    static_assert(kNormPer % R == 0, "whole row groups per thread");
    __shared__ float s_s;
    const int nx = uni(nx_), ny = uni(ny_), src_k = uni(src_k_), src_c = uni(src_c_), trans = uni(trans_);
    const long long bstride = uni(bstride_);
    const float tinit = uni(tinit_);
    const float *Tstart_ = uni(Tstart__);
    const gfloat *Tend = gp(uni(Tend_)), *Tstart = gp(Tstart_);
    // the grid's x outer: columns here, or rows of a transposed grid
    const int nseg = ((trans ? nx : ny) + kNormPer - 1) / kNormPer;   // blocks of terms per x
    const int nblk = (trans ? ny : nx) * nseg, t = threadIdx.x, nt = blockDim.x;
    float s = 0.f;                                      // the running sum, the same in every thread
    for (int g0 = 0; g0 < nblk; g0 += nt) {
        const int g = g0 + t;
        const bool live = g < nblk;
        float d[kNormPer];
        if (!trans) {
            // this thread's terms: rows k0 .. k0 + kNormPer - 1 of column c
            const int c = g / nseg, k0 = (g - c * nseg) * kNormPer;
#pragma unroll
            for (int j = 0; j < kNormPer; j += R) {
                const int k = k0 + j, bb = k / BR, l = (k % BR) / R;
                float te[R], ts[R];
                if (live && k < ny) {
                    const long long a = bb * bstride + (long long)(c + l) * BR + l * R;
                    ldv<R>(Tend + a, te);
                    if (Tstart_) ldv<R>(Tstart + a, ts);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float t0 = Tstart_ ? ts[r] : ((k + r == src_k && c == src_c) ? 0.f : tinit);
                    d[j + r] = (live && k + r < ny) ? fabsf(te[r] - t0) : 0.f;
                }
            }
        } else {
            // transposed: columns c0 .. c0 + kNormPer - 1 of row k
            const int k = g / nseg, c0 = (g - k * nseg) * kNormPer;
            const int bb = k / BR, l = (k % BR) / R, r = k % R;
#pragma unroll
            for (int j = 0; j < kNormPer; ++j) {
                const int c = c0 + j;
                float te = 0.f, t0 = 0.f;
                if (live && c < nx) {
                    const long long a = bb * bstride + (long long)(c + l) * BR + l * R + r;
                    te = Tend[a];
                    t0 = Tstart_ ? Tstart[a] : ((k == src_k && c == src_c) ? 0.f : tinit);
                }
                d[j] = (live && c < nx) ? fabsf(te - t0) : 0.f;
            }
        }
        const int mine = t * kNormPer;                 // chunk index of this thread's first term
        int start = 0;                                 // chunk index of the first term not yet added
        for (;;) {
            if (s == 0.f) {                            // 0 + d = d: the first nonzero term
                int f = INT_MAX;
#pragma unroll
                for (int i = kNormPer - 1; i >= 0; --i)
                    if (mine + i >= start && d[i] != 0.f) f = mine + i;
                const int fmin = block_min(f);
                if (fmin == INT_MAX) break;
#pragma unroll
                for (int i = 0; i < kNormPer; ++i)
                    if (mine + i == fmin) s_s = d[i];
                __syncthreads();
                s = s_s;
                __syncthreads();
                start = fmin + 1;
                continue;
            }
            // s = N * 2^(e-23), 2^23 <= N < 2^24 (N < 2^23 below the normal range)
            const unsigned sb = __float_as_uint(s);
            const int Es = (sb >> 23) & 0xFF;
            const int e = (Es ? Es : 1) - 127;
            const unsigned N = (sb & 0x7FFFFFu) | (Es ? 0x800000u : 0u);
            unsigned v[kNormPer];
            ParFn F{0u, 0u, 2u};
#pragma unroll
            for (int i = 0; i < kNormPer; ++i) {
                v[i] = mine + i >= start ? term_inc(d[i], e, Es != 0) : 0u;
                const unsigned a0 = inc_at(v[i], F.q & 1u), a1 = inc_at(v[i], (F.q >> 1) & 1u);
                F.t0 += a0;
                F.t1 += a1;
                F.q = (((F.q & 1u) + a0) & 1u) | (((((F.q >> 1) & 1u) + a1) & 1u) << 1);
            }
            ParFn all;
            const ParFn pre = block_exscan_fn(F, &all);
            const unsigned base = N + ((N & 1u) ? pre.t1 : pre.t0);
            const unsigned pin = (base & 1u);
            // this thread's run leaves the binade iff a term is big or its
            // sum from `base` reaches 2^24 units (base itself is exact while
            // every earlier run stays inside)
            bool big = false;
#pragma unroll
            for (int i = 0; i < kNormPer; ++i) big |= (v[i] & kIncBig) != 0;
            const unsigned mt = pin ? F.t1 : F.t0;
            const bool leaves = big || base >= (1u << 24) || mt >= (1u << 24) - base;
            if (!__syncthreads_or(leaves)) {           // the rest of the chunk stays in the binade
                s = __builtin_ldexpf((float)(N + ((N & 1u) ? all.t1 : all.t0)), e - 23);
                break;
            }
            // the first term that leaves the binade, in order
            unsigned run = base;
            int evc = INT_MAX;
            unsigned at = 0;
            if (leaves) {
#pragma unroll
                for (int i = 0; i < kNormPer; ++i) {
                    const unsigned a = inc_at(v[i], run & 1u);
                    if (evc == INT_MAX && mine + i >= start) {
                        if ((v[i] & kIncBig) || run + a >= (1u << 24)) { evc = mine + i; at = run; }
                        else run += a;
                    }
                }
            }
            const int j = block_min(evc);
            if (j == evc) {                            // its thread adds the term that leaves as a float
                const float sj = __builtin_ldexpf((float)at, e - 23);
#pragma unroll
                for (int i = 0; i < kNormPer; ++i)
                    if (mine + i == j) s_s = sj + d[i];
            }
            __syncthreads();
            s = s_s;
            __syncthreads();
            start = j + 1;
        }
    }
    return s;
}

// The stopping test after an outer iteration: the reference continues iff
// (float)sum / (float)(NX*NY) >= TTNORM with sum its float norm1 sum, i.e.
// iff sum >= stop_s.  G, the double sum of the iteration's per-sweep
// decreases, differs from the exact sum E of |T - Told| by at most 16 float
// units of E (a float difference per cell and a float sum of at most 8 rows
// per lane before the double sums), and the reference's float sum differs
// from E by at most (1+u)(1+gamma_{NX*NY}) - 1: band_lo/band_hi hold those
// bounds (host, set_stop).  G = 0 means no cell changed, where the float sum
// is exactly 0.  Returns 1 (continue), 0 (stop) or -1 (undecided: count the
// changed cells, stop_counted, and if that cannot decide either, take
// ordered_norm1; see decide()).  EXACT: the decreases do not bound the
// change (the INF guard can raise a value), so every decision takes the
// ordered sum.
template <bool EXACT = false>
__device__ __forceinline__ int stop_fast(const ModelDev &md, double G)
{
    if (EXACT) return -1;
    if (G == 0.0) return 0.0 >= md.stop_s ? 1 : 0;
    if (G * md.band_hi < md.stop_s) return 0;
    if (G * md.band_lo >= md.stop_s) return 1;
    return -1;
}

// Cells whose value differs between two fields of layout A (Tstart ==
// nullptr: the initial forward field): norm1's nonzero terms.  A coalesced
// pass of the whole workgroup over both fields, one lane's R rows at one key
// per load, eight loads of each field in flight per thread.
template <int R>
__device__ int count_changed(const ModelDev &md, const float *Tend_, const float *Tstart_, float tinit, int src_k,
                             int src_c)
{
    constexpr int BR = kWave * R, U = 8;
    __shared__ int cnt_s;
    if (threadIdx.x == 0) cnt_s = 0;
    __syncthreads();
    const gfloat *Tend = gp(Tend_), *Tstart = gp(Tstart_);
    const int gpb = (md.nx + 64) * kWave;   // groups of R rows per band
    const int nt = blockDim.x;
    int n = 0;
    for (int b = 0; b < md.nb; ++b) {
        const gfloat *e = Tend + (long long)b * md.bstride, *o = Tstart + (long long)b * md.bstride;
        for (int q0 = threadIdx.x; q0 < gpb; q0 += U * nt) {
            float te[U][R], ts[U][R];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = min(q0 + u * nt, gpb - 1);
                ldv<R>(e + q * R, te[u]);
                if (Tstart_) ldv<R>(o + q * R, ts[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int q = q0 + u * nt, key = q / kWave, l = q % kWave, c = key - l;
                const bool colok = q < gpb && c >= 0 && c < md.nx;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int k = b * BR + l * R + r;
                    const float t0 = Tstart_ ? ts[u][r] : ((k == src_k && c == src_c) ? 0.f : tinit);
                    n += (colok && k < md.ny && te[u][r] != t0) ? 1 : 0;
                }
            }
        }
    }
    atomicAdd(&cnt_s, n);
    __syncthreads();
    const int r = cnt_s;
    __syncthreads();
    return r;
}

// stop_fast's test with the bound of m nonzero terms instead of NX*NY.
__device__ __forceinline__ int stop_counted(const ModelDev &md, double G, int m)
{
    const double u = 0x1p-24, mu = (double)m * u;
    if (mu >= 0.5) return -1;
    const double gam = mu / (1.0 - mu), er = (1.0 + u) * (1.0 + gam) - 1.0, eg = 16.0 * u + 1e-9;
    if (G * ((1.0 + er) / (1.0 - eg) * (1.0 + 1e-9)) < md.stop_s) return 0;
    if (G * ((1.0 - er) / (1.0 + eg) * (1.0 - 1e-9)) >= md.stop_s) return 1;
    return -1;
}

__device__ __forceinline__ bool stop_exact(const ModelDev &md, float sum)
{
    const float denom = (float)(md.nx * md.ny);
    return sum / denom >= md.ttnorm;   // norm1.c:27, eikonal.c:44
}

// The stopping decision after an outer iteration whose fields at the start
// (Wi; nullptr: the initial forward field) and at the end (Wo) are intact:
// the fast test, then the bound with `mbound` changed cells when the
// sweeps counted them (mbound >= 0: the forward's count over its sweeps of
// R cells per lane step with a change, at least the number of cells that
// changed), then the bound with the exact
// count of changed cells, then the ordered sum.  Every thread of the
// workgroup calls it with the same G and gets the same answer.  `passes`
// is set when the fields were read (a split shot's partner must not
// overwrite Wi before both are done); nord counts the passes: 1 per count
// of changed cells, 1 << 16 per ordered sum; osum is the ordered sum when
// one was taken, else -1.  md.stop_mode 1 and 2 (tests) skip to the ordered
// sum, resp. to the count, for every nonzero G.
template <int R, bool EXACT>
__device__ bool decide(const ModelDev &md, double G, double mbound, const float *Wo, const float *Wi, float tinit,
                       int src_k, int src_c, bool &passes, int &nord, float &osum)
{
    osum = -1.f;
    int d = stop_fast<EXACT>(md, G);
    if (d < 0 && !EXACT && md.stop_mode == 0 && mbound >= 0.0) d = stop_counted(md, G, (int)fmin(mbound, 1e9));
    passes = d < 0;
    if (d >= 0) return d != 0;
    if (!EXACT && md.stop_mode != 1) {
        ++nord;
        d = stop_counted(md, G, count_changed<R>(md, Wo, Wi, tinit, src_k, src_c));
    }
    if (d >= 0) return d != 0;
    nord += 1 << 16;
    osum = ordered_norm1<R>(md.nx, md.ny, md.bstride, Wo, Wi, tinit, src_k, src_c, md.trans);
    return stop_exact(md, osum);
}

// Steps of a sweep: the last band's keys plus `hold` steps to drain staged
// stores, rounded up to whole epochs.
__device__ __forceinline__ int sweep_steps(int nx, int nb, int hold = 0)
{
    return ((nb - 1) * kBandOff + nx + 63 + hold + kPF - 1) / kPF * kPF;
}

// Staged stores of a forward sweep into the other family's layout.  In that
// layout the R cells lane l computes at one step land in a different key
// block for every lane (key + 63 - 2l from A to B, key + 2l - 63 from B to
// A); stored directly they would be 16-byte pieces of 64 different lines.
// Instead each lane keeps its pieces in an LDS delay line and stores the
// piece of step tau - d_l at step tau, where d_l = 2|j_last - j| + 1 for lane
// j = l mod P of its segment of P lanes: the P lanes of a segment then store
// the same segment of one block together, and every store instruction
// writes whole segments.  Segments are whole 128-byte cache lines (P = 32/R)
// where a workgroup has 8 waves (R >= 4: 16 KiB of LDS per wave) and
// 64-byte halves (P = 16/R) where it has 16 (8 KiB per wave).  Scattered
// half lines cost a sixth of the time per iteration with every CU busy
// (measured against contiguous stores), whole lines less.  With the extra
// step of delay every piece was written at least one step earlier, so a
// step reads its piece from LDS before its arithmetic and the LDS latency
// overlaps it.
template <int R>
struct Stage {
    static constexpr int BYTES = R >= 4 ? 128 : 64;  // bytes per segment
    static constexpr int P = BYTES / (4 * R);        // lanes per segment
    static constexpr int HOLD = 2 * (P - 1) + 1;     // longest delay, in steps
    static constexpr int Q = (HOLD < 4) ? 4 : (HOLD < 8) ? 8 : (HOLD < 16) ? 16 : 32;   // slots (power of 2)
    static constexpr int WAVE_FLOATS = Q * kWave * R;                                   // 16 or 8 KiB
    static_assert(HOLD < Q, "a slot is overwritten only after its piece is stored");
};

// The barriers of `nep` epochs without a sweep (a group with nothing to do).
__device__ __forceinline__ void idle_epochs(int nep)
{
    for (int i = 0; i < nep; ++i) epoch_sync(i);
    __syncthreads();
}

// ---------------------------------------------------------------------------
// How a sweep keeps pace with the other sweep of its pair.
//   kPaceLock   both sweeps in one workgroup, in lockstep: `pre` idle epochs
//               before the sweep and `post` after it, epoch_sync() between
//               epochs (the trailing sweep has pre = D, the leading post = D)
//   kPaceLead   split shot, leading sweep in a workgroup of its own: every
//               eighth epoch it drains its stores and publishes its epoch in
//               `lead` (L2, same XCD as the partner)
//   kPaceTrail  split shot, trailing sweep: before each block of eight
//               epochs it waits until the leading sweep has published the
//               epoch that the lockstep delay D would have guaranteed
//               (epoch + D - 9 for every load issued), then drops its L1.
// `base` makes the published values grow over a whole solve.
// ---------------------------------------------------------------------------
enum PaceMode { kPaceLock = 0, kPaceLead = 1, kPaceTrail = 2 };
constexpr int kPhaseStride = 1 << 20;      // epochs per pair phase are fewer
constexpr int kPhaseAll = kPhaseStride - 1;
constexpr unsigned kSpinLimit = 1u << 22;  // polls (about 1 s) before giving up

struct Pace {
    int pre, post;
    int *lead;      // leading sweep's published epoch (split modes)
    int *err;       // set when a wait gives up or the partner is elsewhere
    int base, D;
};

__device__ __forceinline__ Pace lockstep(int pre, int post)
{
    Pace p;
    p.pre = pre; p.post = post; p.lead = nullptr; p.err = nullptr; p.base = 0; p.D = 0;
    return p;
}

// all of this wave's vector memory operations complete (vmcnt(0))
__device__ __forceinline__ void drain_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Counters of split shots, addressed as global memory (the sweeps are
// separate functions, where a plain pointer would be a generic one: generic
// accesses make every later wait of the wave a full one).
typedef __attribute__((address_space(1))) int gint;

__device__ __forceinline__ void publish(int *p, int v)
{
    if (threadIdx.x == 0) __hip_atomic_store((gint *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0 polls *p >= target (bounded: on timeout, or when another wait of
// the launch gave up, it sets / sees *err); then the workgroup meets and
// drops its L1 so that later loads see what the partner stored.  Returns
// false once the launch is failing (every workgroup then runs to its end
// without waiting, and the host repeats the work unsplit).
__device__ bool pace_wait(int *p, int target, int *err)
{
    __shared__ int ok_s;
    if (threadIdx.x == 0) {
        int ok = 1;
        for (unsigned n = 0;; ++n) {
            if (__hip_atomic_load((gint *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load((gint *)err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
            if (n >= kSpinLimit) {
                __hip_atomic_store((gint *)err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        ok_s = ok;
    }
    __syncthreads();
    const bool ok = ok_s;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    return ok;
}

template <int MODE>
__device__ __forceinline__ void pace_begin(const Pace &pc)
{
    if constexpr (MODE == kPaceLock) {
        for (int i = 0; i < pc.pre; ++i) epoch_sync(i);
    } else if constexpr (MODE == kPaceTrail) {
        (void)pace_wait(pc.lead, pc.base + pc.D - 2, pc.err);
    }
}

template <int MODE>
__device__ __forceinline__ void pace_epoch(const Pace &pc, int e)
{
    if constexpr (MODE == kPaceLock) {
        epoch_sync(pc.pre + e);
    } else if constexpr (MODE == kPaceLead) {
        if ((e & 7) == 7) {
            drain_vmem();
            epoch_barrier();
            publish(pc.lead, pc.base + e);
        } else {
            epoch_barrier();
        }
    } else {
        epoch_barrier();
        if ((e & 7) == 7) (void)pace_wait(pc.lead, pc.base + e + pc.D - 1, pc.err);
    }
}

template <int MODE>
__device__ __forceinline__ void pace_end(const Pace &pc, int nep)
{
    if constexpr (MODE == kPaceLock) {
        for (int i = 0; i < pc.post; ++i) epoch_sync(pc.pre + nep + i);
    } else if constexpr (MODE == kPaceLead) {
        drain_vmem();
        __syncthreads();
        publish(pc.lead, pc.base + kPhaseAll);
    }
    __syncthreads();
}

// The pacing of a sweep for a wave without a band in it.
template <int MODE>
__device__ __forceinline__ void pace_idle(const Pace &pc, int nep)
{
    pace_begin<MODE>(pc);
    for (int e = 0; e < nep; ++e) pace_epoch<MODE>(pc, e);
    pace_end<MODE>(pc, nep);
}

// ---------------------------------------------------------------------------
// Forward sweep (sweep.c:10-104) in direction (SX, SY) over one shot.
//   Tin    field written by the previous sweep, in this sweep's family
//          layout (A if SX==SY, else B)
//   Sf     slowness in this sweep's family layout
//   Tout   written in the other family's layout, which the next sweep reads
//          (consecutive forward sweeps alternate between families A and B),
//          through the staged stores of Stage
//   stage  this wave's LDS delay line (Stage<R>::WAVE_FLOATS floats)
//   INIT   read the initial field (2*TIME, 0 at the source) instead of Tin
// Returns this lane's sum of (old - new) over its valid cells and the number
// of cells the wave changed.
//
// All loads are whole 64*R-float blocks of this sweep's layout and all
// stores whole segments (Stage) of the other layout.  Garbage (values of
// lanes outside the grid or of steps before the first one) only reaches
// positions of cells outside the grid, which no valid cell reads.
//
// Every step issues the same memory operations whatever the lane or step
// validity (invalid lanes write to the band's dummy key block nx+63 and to a
// dummy LDS word), so the compiler can wait on each prefetch with a partial
// vmcnt instead of draining the queue.
// ---------------------------------------------------------------------------
template <int R, int SX, int SY, bool INIT, int PACE = kPaceLock>
__device__ __attribute__((noinline)) Dec fwd_sweep(const WaveCtx w_, float dh_, double two_dh2_,
                                                      const float *Tin_, const float *Sf_, float *Tout_,
                                                      float *stage_, int src_k_, int src_c_, float tinit_,
                                                      const Pace pc_)
{
    const WaveCtx w = uni_ctx(w_);
    const float dh = uni(dh_), tinit = uni(tinit_);
    const double two_dh2 = uni(two_dh2_);
    const float *Tin = uni(Tin_), *Sf = uni(Sf_);
    float *Tout = uni(Tout_), *stage = uni(stage_);
    const int src_k = uni(src_k_), src_c = uni(src_c_);
    const Pace pc = uni(pc_);
    constexpr bool FA = (SX == SY);
    constexpr int BR = kWave * R;
    constexpr int SP = Stage<R>::P, SQ = Stage<R>::Q;
    constexpr int PS = pf_slots(R);
    const int nx = w.nx, ny = w.ny, nb = w.nb, beta = w.beta, l = w.lane;
    const int nkeys = nx + 63;          // data keys 0..nx+62; key nkeys is the dummy block
    const int gend = sweep_steps(nx, nb, Stage<R>::HOLD);
    const int nep = gend / kPF;
    if (!w.active) {
        pace_idle<PACE>(pc, nep);
        return Dec{0.0, 0.0};
    }
    pace_begin<PACE>(pc);
    const int lag = SY > 0 ? l : 63 - l;
    const int off = (SY > 0 ? beta : nb - 1 - beta) * kBandOff;
    const long long boff = (long long)beta * w.bstride;
    const gfloat *tband = gp(Tin) + boff;
    const gfloat *sband = gp(Sf) + boff;
    gfloat *gTout = gpw(Tout) + boff;
    lfloat *lstage = (lfloat *)stage + l * R;
    const int k0 = beta * BR + l * R;
    // key of this lane's cells in the other family: B key = A key - 2l + 63
    const int kshift = FA ? 63 - 2 * l : 2 * l - 63;
    // delay of this lane's staged stores (Stage)
    const int sj = l % SP, sjl = ((SX > 0) == FA) ? SP - 1 : 0;
    const int dl = 2 * (sjl > sj ? sjl - sj : sj - sjl) + 1;

    bool rvalid[R], top[R], bot[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = k0 + r;
        rvalid[r] = k < ny;
        top[r] = (k == 0);
        bot[r] = (k == ny - 1);
    }
    // Old value across the band boundary on the downstream side: the row
    // below (SY>0, needed by lane 63) or above (SY<0, lane 0).  Both edge
    // lanes have lag 63, so the address is the same for the whole wave.
    const bool has_dn = SY > 0 ? (beta + 1 < nb) : (beta > 0);
    const gfloat *eband = has_dn ? tband + (SY > 0 ? w.bstride : -w.bstride) : tband;
    const int erow = SY > 0 ? (beta + 1) * BR : beta * BR - 1;
    const int ering = clampi(SY > 0 ? beta - 1 : beta + 1, 0, kMaxBands - 1);
    const bool ring_writer = (SY > 0 ? (l == 63) : (l == 0)) && (SY > 0 ? (beta + 1 < nb) : (beta > 0));
    lfloat *ring = (lfloat *)w.ring;
    lfloat *ring_dummy = ring + kMaxBands * kRingQ + l;

    float qt[PS][R], qs[PS][R], qe[PS];

#define RZ_FWD_LOAD(SLOT, TAU)                                                              \
    do {                                                                                    \
        const int tau_ = (TAU);                                                             \
        const int key_ = clampi(SX > 0 ? tau_ : nx + 62 - tau_, 0, nkeys - 1);              \
        ldv<R>(sband + key_ * BR + l * R, qs[SLOT]);                                        \
        const int ce_ = clampi(SX > 0 ? tau_ - 63 : nx - 1 - tau_ + 63, 0, nx - 1);         \
        if constexpr (INIT) {                                                               \
            const int c_ = FA ? key_ - l : key_ + l - 63;                                   \
            _Pragma("unroll") for (int r_ = 0; r_ < R; ++r_)                                \
                qt[SLOT][r_] = (k0 + r_ == src_k && c_ == src_c) ? 0.f : tinit;             \
            qe[SLOT] = (erow == src_k && ce_ == src_c) ? 0.f : tinit;                       \
        } else {                                                                            \
            ldv<R>(tband + key_ * BR + l * R, qt[SLOT]);                                    \
            const int ek_ = SY > 0 ? (FA ? ce_ : ce_ + 63) : (FA ? ce_ + 63 : ce_);         \
            qe[SLOT] = eband[ek_ * BR + (SY > 0 ? 0 : 63 * R + R - 1)];                     \
        }                                                                                   \
    } while (0)

    // The first sweep from the initial field (INIT, always +x) changes no
    // cell left of column src_c - 1: such a cell's neighbours all hold the
    // initial value (the source's column is further right), and the update
    // of a cell among initial values returns at least that value.  Keys
    // below src_c - 1 hold only such cells (key = column + lane in layout
    // A), so the epochs of this band whose steps all lie there just store
    // the initial value: no loads, no arithmetic.  gbeg is the first epoch
    // computed in full (0 in every other sweep).
    int gbeg = 0;
    if constexpr (INIT && SX > 0) gbeg = min(max(src_c - 1 + off, 0) / kPF, nep) * kPF;
    float pnew[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pnew[r] = tinit;
    double acc = 0.0;
    unsigned chg = 0;        // steps at which some of this lane's cells changed
    float ring_next = 0.f;   // first consumer step is >= kBandOff steps away
    if constexpr (INIT && SX > 0) {
        for (int g0 = 0; g0 < gbeg; g0 += kPF) {
#pragma unroll
            for (int u = 0; u < kPF; ++u) {
                const int tau = g0 + u - off;
                const bool cvalid = tau >= 0 && tau < nkeys && (unsigned)(tau - lag) < (unsigned)nx;
                const int c = tau - lag;
                {   // read one step ahead, as below
                    const int cn = tau + 1 - lag;
                    ring_next = ring[ering * kRingQ + (clampi(cn, 0, nx - 1) & (kRingQ - 1))];
                }
                const int tp = tau - dl;
                float piece[R];
                ldl<R>(lstage + (tp & (SQ - 1)) * BR, piece);
                stl<R>(lstage + (tau & (SQ - 1)) * BR, pnew);
                int ko = tp + kshift;
                ko = (unsigned)ko < (unsigned)nkeys ? ko : nkeys;
                stv<R>(gTout + ko * BR + l * R, piece);
                lfloat *rw = (ring_writer && cvalid) ? &ring[beta * kRingQ + (c & (kRingQ - 1))] : ring_dummy;
                *rw = tinit;
            }
            pace_epoch<PACE>(pc, g0 / kPF);
        }
    }

#pragma unroll
    for (int u = 0; u < PS; ++u) RZ_FWD_LOAD(u, gbeg + u - off);
    // The set-up ends in a full wait: the waits the compiler derives at the
    // loop head then come from the loop's own pattern of prefetches (a slot
    // is waited for PS - 1 steps after its load), not from a set-up
    // load just before the loop, which would cost a full memory latency in
    // every epoch.
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)

    for (int g0 = gbeg; g0 < gend; g0 += kPF) {
#pragma unroll
        for (int u = 0; u < kPF; ++u) {
            const int sl = u % PS, un = (u + 1) % PS;   // g0 is a multiple of PS
            const int tau = g0 + u - off;
            const bool active = tau >= 0 && tau < nkeys;   // wave-uniform
            const int c = SX > 0 ? tau - lag : nx - 1 - tau + lag;
            const bool cvalid = active && (unsigned)c < (unsigned)nx;
            // ring value of this step, read one step ahead (it was written
            // eight or more steps before, in an earlier epoch)
            const float ringv = ring_next;
            {
                const int cn = SX > 0 ? tau + 1 - lag : nx - 2 - tau + lag;
                ring_next = ring[ering * kRingQ + (clampi(cn, 0, nx - 1) & (kRingQ - 1))];
            }
            float nv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) nv[r] = qt[sl][r];
            // staged store, part 1: the piece of step tau - dl (written at
            // least one step ago), read before this step's arithmetic
            const int tp = tau - dl;
            float piece[R];
            ldl<R>(lstage + (tp & (SQ - 1)) * BR, piece);
            if (active) {
                const bool first = SX > 0 ? (c == 0) : (c == nx - 1);
                const bool last = SX > 0 ? (c == nx - 1) : (c == 0);
                float up_new, dn_old;
                if (SY > 0) {
                    up_new = lane_from_above(pnew[R - 1]);
                    dn_old = lane_from_below(qt[un][0]);
                    if (l == 0) up_new = ringv;
                    if (l == 63) dn_old = qe[sl];
                } else {
                    up_new = lane_from_below(pnew[0]);
                    dn_old = lane_from_above(qt[un][R - 1]);
                    if (l == 63) up_new = ringv;
                    if (l == 0) dn_old = qe[sl];
                }
                // The R cells of a lane form a chain along y; all of them are
                // first updated with the fast root.  If any cell of the wave
                // lies near a float rounding tie (rare), the whole step is
                // redone with the correctly rounded root.
                float fsum = 0.f;
                bool need = false;
                fwd_cells<R, SY, false>(qt[sl], qt[un], qs[sl], pnew, up_new, dn_old, first, last, top, bot,
                                        rvalid, cvalid, dh, two_dh2, nv, fsum, need);
                if (__builtin_amdgcn_ballot_w64(need)) {
                    fsum = 0.f;
                    fwd_cells<R, SY, true>(qt[sl], qt[un], qs[sl], pnew, up_new, dn_old, first, last, top, bot,
                                           rvalid, cvalid, dh, two_dh2, nv, fsum, need);
                }
                acc += (double)fsum;
                chg += fsum > 0.f ? 1u : 0u;   // a change makes the (non-negative) sum positive
#pragma unroll
                for (int r = 0; r < R; ++r) pnew[r] = nv[r];
            }
            {   // staged store, part 2: keep this step's cells, store the piece
                stl<R>(lstage + (tau & (SQ - 1)) * BR, nv);
                int ko = (SX > 0 ? tp : nx + 62 - tp) + kshift;
                ko = (unsigned)ko < (unsigned)nkeys ? ko : nkeys;
                stv<R>(gTout + ko * BR + l * R, piece);
            }
            lfloat *rw = (ring_writer && cvalid) ? &ring[beta * kRingQ + (c & (kRingQ - 1))] : ring_dummy;
            *rw = SY > 0 ? nv[R - 1] : nv[0];
            RZ_FWD_LOAD(sl, tau + PS);
        }
        pace_epoch<PACE>(pc, g0 / kPF);
    }
    pace_end<PACE>(pc, nep);
#undef RZ_FWD_LOAD
    return Dec{acc, (double)chg * R};
}

// ---------------------------------------------------------------------------
// Division by DH (sweep_adj.c:33-54).  FD: (float)(x * RN64(1/DH)) in
// double, which is the correctly rounded x/DH for every float x: a quotient
// of two floats is never within 2^-49 (relative) of a float rounding tie,
// while the double product is within 2^-52 of it.  The host also checks it
// against the IEEE division for the model's DH before enabling it.
// ---------------------------------------------------------------------------
template <int FD>
__device__ __forceinline__ float qdh(float x, float dh, double rdh)
{
    if constexpr (FD >= 1) return (float)((double)x * rdh);
    else return x / dh;
}

// x/DH for the T-only quotients (slopes and the terms of lhs).  FD == 2:
// fma(x, rh, x * rl) with rh + rl ~ 1/DH, equal to the IEEE quotient for the
// model's DH and for every input these quotients can see (the host checks
// both before it selects this variant); otherwise as qdh.
template <int FD>
__device__ __forceinline__ float qdh_t(float x, float dh, double rdh, float rh, float rl)
{
    if constexpr (FD == 2) return __builtin_fmaf(x, rh, x * rl);
    else return qdh<FD>(x, dh, rdh);
}

// (x + |x|)/2 and (x - |x|)/2 of sweep_adj.c:39-49 for finite |x| < 2^127:
// max(x, +0) and min(x, +0).  For x != 0 both are exact; for x = -0 the
// formulas give +0 and -0, and v_max_f32 / v_min_f32 order -0 below +0, so
// max(-0, +0) = +0 and min(-0, +0) = -0.  Operands are never NaN (slopes
// of finite traveltimes); the inline form keeps the compiler from
// canonicalising them first.
__device__ __forceinline__ float pos_part(float x)
{
    float r;
    asm("v_max_f32 %0, 0, %1" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ float neg_part(float x)
{
    float r;
    asm("v_min_f32 %0, 0, %1" : "=v"(r) : "v"(x));
    return r;
}

// ---------------------------------------------------------------------------
// Adjoint sweep (sweep_adj.c:10-77) in direction (SX, SY).
//   Tf    converged traveltime in this sweep's family layout; the sign bit
//         marks the cells that keep their value: receivers
//         (sweep_adj.c:30) and the frame (mark_kept_cells)
//   Lin   lambda from the previous sweep, in this sweep's layout
//   Lout  lambda written: in this sweep's layout (LSAME, Lout == Lin, the
//         sweep updates it in place) or the other family's, through the
//         staged stores of fwd_sweep (stage: this wave's LDS delay line)
// The T-only terms (slopes, lhs, skip and INF conditions) are evaluated
// from T in every sweep: recomputing them costs less than the HBM traffic of
// storing them.  The memory operations per step are fixed, as in fwd_sweep.
// ---------------------------------------------------------------------------
template <int R, int SX, int SY, bool LSAME, int FD, bool CHKINF, int PACE = kPaceLock>
__device__ __attribute__((noinline)) double adj_sweep(const WaveCtx w_, float dh_, double rdh_, float rh_,
                                                      float rl_, float eps_, const float *Tf_, const float *Lin_,
                                                      float *Lout_, float *stage_, const Pace pc_)
{
    const WaveCtx w = uni_ctx(w_);
    const float dh = uni(dh_), eps = uni(eps_), rh = uni(rh_), rl = uni(rl_);
    const double rdh = uni(rdh_);
    const float *Tf = uni(Tf_), *Lin = uni(Lin_);
    float *Lout = uni(Lout_), *stage = uni(stage_);
    const Pace pc = uni(pc_);
    constexpr bool FA = (SX == SY);
    constexpr int BR = kWave * R;
    constexpr int SP = Stage<R>::P, SQ = Stage<R>::Q;
    constexpr int PS = pf_slots(R);
    const int nx = w.nx, ny = w.ny, nb = w.nb, beta = w.beta, l = w.lane;
    const int nkeys = nx + 63;
    const int gend = sweep_steps(nx, nb, Stage<R>::HOLD);   // the same for both kinds of output
    const int nep = gend / kPF;
    if (!w.active) {
        pace_idle<PACE>(pc, nep);
        return 0.0;
    }
    pace_begin<PACE>(pc);
    const int lag = SY > 0 ? l : 63 - l;
    const int off = (SY > 0 ? beta : nb - 1 - beta) * kBandOff;
    const long long boff = (long long)beta * w.bstride;
    gfloat *gLout = gpw(Lout) + boff;
    // T and lambda (old) through buffer loads over the whole field: the
    // lane's byte offset is fixed, the step's is scalar and 32-bit
    const unsigned fbytes = 4u * (unsigned)(nb * w.bstride);
    const __amdgpu_buffer_rsrc_t rT = field_rsrc(Tf, fbytes), rL = field_rsrc(Lin, fbytes);
    const int vlane = 4 * l * R, sband = 4 * (int)boff;
    lfloat *lstage = (lfloat *)stage + l * R;
    const int k0 = beta * BR + l * R;
    // key of this lane's cells in the other family, and the delay of its
    // staged stores (as in fwd_sweep)
    const int kshift = FA ? 63 - 2 * l : 2 * l - 63;
    const int sj = l % SP, sjl = ((SX > 0) == FA) ? SP - 1 : 0;
    const int dl = 2 * (sjl > sj ? sjl - sj : sj - sjl) + 1;

    bool rvalid[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rvalid[r] = k0 + r < ny;
    // Rows across the band boundaries: above lane 0 (band beta-1, lane 63,
    // r = R-1) and below lane 63 (band beta+1, lane 0, r = 0); each address
    // is the same for the whole wave.  T on both sides, lambda (old) on the
    // downstream side.
    const int lag0 = SY > 0 ? 0 : 63, lag63 = SY > 0 ? 63 : 0;
    const bool has_up = beta > 0, has_dn = beta + 1 < nb;
    const int up_base = has_up ? -(int)w.bstride + 63 * R + R - 1 : 0;
    const int dn_base = has_dn ? (int)w.bstride : 0;
    const int ering = clampi(SY > 0 ? beta - 1 : beta + 1, 0, kMaxBands - 1);
    const bool ring_writer = (SY > 0 ? (l == 63) : (l == 0)) && (SY > 0 ? has_dn : has_up);
    lfloat *ring = (lfloat *)w.ring;
    lfloat *ring_dummy = ring + kMaxBands * kRingQ + l;

    float qt[PS][R], ql[PS][R], qeu[PS], qed[PS], qel[PS];

#define RZ_ADJ_LOAD(SLOT, TAU)                                                              \
    do {                                                                                    \
        const int tau_ = (TAU);                                                             \
        const int key_ = clampi(SX > 0 ? tau_ : nx + 62 - tau_, 0, nkeys - 1);              \
        ldb<R>(rT, vlane, sband + 4 * BR * key_, qt[SLOT]);                                 \
        ldb<R>(rL, vlane, sband + 4 * BR * key_, ql[SLOT]);                                 \
        const int c0_ = clampi(SX > 0 ? tau_ - lag0 : nx - 1 - tau_ + lag0, 0, nx - 1);     \
        const int c63_ = clampi(SX > 0 ? tau_ - lag63 : nx - 1 - tau_ + lag63, 0, nx - 1);  \
        const int eu_ = sband + 4 * (up_base + (FA ? c0_ + 63 : c0_) * BR);                \
        const int ed_ = sband + 4 * (dn_base + (FA ? c63_ : c63_ + 63) * BR);               \
        qeu[SLOT] = ldb1(rT, eu_);                                                          \
        qed[SLOT] = ldb1(rT, ed_);                                                          \
        qel[SLOT] = ldb1(rL, SY > 0 ? ed_ : eu_);                                           \
    } while (0)

    // Slot u holds the blocks of step u (mod PS).  A step reads its own
    // slot, the next one (T and lambda at the next key) and the previous one
    // (T at the previous key), then reloads the previous slot with the
    // blocks of PS - 1 steps ahead: no value outlives its slot, so the
    // steady-state loop needs no register copies, and its waits are counted
    // ones (a slot is waited for a few steps after its load).  The set-up
    // ends in a full wait, so that the loop head's waits do not depend on
    // set-up loads.  Slot PS-1 is first loaded by step 0; before that it
    // stands for T at the key before the first, which only cells that keep
    // their value read.
#pragma unroll
    for (int u = 0; u < PS - 1; ++u) RZ_ADJ_LOAD(u, u - off);
#pragma unroll
    for (int r = 0; r < R; ++r) qt[PS - 1][r] = 0.f;
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)

    // xcp/xcn: positive and negative parts of the x quotient carried to the
    // next step (ap(c) is am(c+1)); the parts of +0 are +0
    float pnew[R], xcp[R], xcn[R];
#pragma unroll
    for (int r = 0; r < R; ++r) { pnew[r] = 0.f; xcp[r] = 0.f; xcn[r] = 0.f; }
    double acc = 0.0;
    float ring_next = 0.f;

    for (int g0 = 0; g0 < gend; g0 += kPF) {
#pragma unroll
        for (int u = 0; u < kPF; ++u) {
            // g0 is a multiple of PS: step g0 + u uses slot u % PS
            const int sl = u % PS, un = (u + 1) % PS, up = (u + PS - 1) % PS;
            const float (&tprev)[R] = qt[up];   // T at the previous key
            const int tau = g0 + u - off;
            const bool active = tau >= 0 && tau < nkeys;   // wave-uniform
            const int c = SX > 0 ? tau - lag : nx - 1 - tau + lag;
            const bool cvalid = active && (unsigned)c < (unsigned)nx;
            const float ringv = ring_next;   // read one step ahead, as in fwd_sweep
            {
                const int cn = SX > 0 ? tau + 1 - lag : nx - 2 - tau + lag;
                ring_next = ring[ering * kRingQ + (clampi(cn, 0, nx - 1) & (kRingQ - 1))];
            }
            float nv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) nv[r] = ql[sl][r];

            // Every step is computed: an inactive step only has lanes that
            // keep their value (cvalid is false).  Without a branch the eight
            // steps of an epoch form one block, in which the scheduler
            // overlaps a step's T-only terms with the previous step's lambda
            // chain, and the carried values need no copies at a join; both
            // outweigh the work of the inactive steps (measured).
            {
                // T of the rows across the lane boundary, at column c
                float t_above, t_below;
                if (SY > 0) {   // lane l-1 is one step ahead, lane l+1 one step behind
                    t_above = lane_from_above(tprev[R - 1]);
                    t_below = lane_from_below(qt[un][0]);
                } else {        // lane l-1 behind, lane l+1 ahead
                    t_above = lane_from_above(qt[un][R - 1]);
                    t_below = lane_from_below(tprev[0]);
                }
                if (l == 0) t_above = qeu[sl];
                if (l == 63) t_below = qed[sl];
                // lambda across the lane boundary: new upstream, old downstream
                float l_up, l_dn;   // row k-1 for r = 0, row k+1 for r = R-1
                if (SY > 0) {
                    l_up = lane_from_above(pnew[R - 1]);
                    l_dn = lane_from_below(ql[un][0]);
                    if (l == 0) l_up = ringv;
                    if (l == 63) l_dn = qel[sl];
                } else {
                    l_up = lane_from_above(ql[un][R - 1]);
                    l_dn = lane_from_below(pnew[0]);
                    if (l == 0) l_up = qel[sl];
                    if (l == 63) l_dn = ringv;
                }
                // T-only part of the cell equations (no dependence on lambda).
                // Each quotient's positive and negative parts are taken once:
                // am(c+1) = ap(c) comes from the previous step, bm(k) = bp(k-1)
                // from the row above in this lane.
                float lhs[R], amp[R], apm[R], bmp[R], bpm[R], xqp[R], xqn[R], bpp[R];
                bool keep[R], infv[R];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float traw = qt[sl][r];
                    const float tc = fabsf(traw);
                    const float tl = fabsf(SX > 0 ? tprev[r] : qt[un][r]);   // T(k, c-1)
                    const float tr = fabsf(SX > 0 ? qt[un][r] : tprev[r]);   // T(k, c+1)
                    const float tu = fabsf(r == 0 ? t_above : qt[sl][r > 0 ? r - 1 : 0]);
                    const float td = fabsf(r == R - 1 ? t_below : qt[sl][r < R - 1 ? r + 1 : r]);
                    float app, amm;   // pos_part(ap), neg_part(am)
                    if (SX > 0) {
                        const float ap = qdh_t<FD>(-(tr - tc), dh, rdh, rh, rl);
                        app = pos_part(ap);
                        apm[r] = neg_part(ap);
                        amp[r] = xcp[r];
                        amm = xcn[r];
                        xqp[r] = app;
                        xqn[r] = apm[r];
                    } else {
                        const float am = qdh_t<FD>(-(tc - tl), dh, rdh, rh, rl);
                        amp[r] = pos_part(am);
                        amm = neg_part(am);
                        app = xcp[r];
                        apm[r] = xcn[r];
                        xqp[r] = amp[r];
                        xqn[r] = amm;
                    }
                    const float bp = qdh_t<FD>(-(td - tc), dh, rdh, rh, rl);
                    bpp[r] = pos_part(bp);
                    bpm[r] = neg_part(bp);
                    float bmm;        // neg_part(bm)
                    if (r > 0) {
                        bmp[r] = bpp[r > 0 ? r - 1 : 0];
                        bmm = bpm[r > 0 ? r - 1 : 0];
                    } else {
                        const float bm = qdh_t<FD>(-(tc - tu), dh, rdh, rh, rl);
                        bmp[r] = pos_part(bm);
                        bmm = neg_part(bm);
                    }
                    lhs[r] = qdh_t<FD>(app - amm, dh, rdh, rh, rl) + qdh_t<FD>(bpp[r] - bmm, dh, rdh, rh, rl);
                    keep[r] = __float_as_uint(traw) >> 31;   // receiver or frame (mark_kept_cells)
                    if constexpr (CHKINF) infv[r] = fmaxf(fmaxf(fmaxf(tc, td), fmaxf(tu, tr)), tl) >= kInf;
                    else infv[r] = false;
                }
                float fsum = 0.f;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const int r = SY > 0 ? i : R - 1 - i;
                    const float lc = ql[sl][r];
                    const float lxm = SX > 0 ? pnew[r] : ql[un][r];            // lambda(k, c-1)
                    const float lxp = SX > 0 ? ql[un][r] : pnew[r];            // lambda(k, c+1)
                    float lym, lyp;
                    if (SY > 0) {
                        lym = (r == 0) ? l_up : nv[r > 0 ? r - 1 : 0];
                        lyp = (r == R - 1) ? l_dn : ql[sl][r < R - 1 ? r + 1 : r];
                    } else {
                        lym = (r == 0) ? l_up : ql[sl][r > 0 ? r - 1 : 0];
                        lyp = (r == R - 1) ? l_dn : nv[r < R - 1 ? r + 1 : r];
                    }
                    const float rhs = qdh<FD>(amp[r] * lxm - apm[r] * lxp, dh, rdh)
                                    + qdh<FD>(bmp[r] * lym - bpm[r] * lyp, dh, rdh);
                    // sweep_adj.c:60-64 without a branch: every row divides (by 1 where
                    // lhs <= EPS_ADJ) and the rows' chains can overlap
                    const bool pos = lhs[r] > eps;
                    const float lt = pos ? rhs / (pos ? lhs[r] : 1.0f) : 0.f;
                    float v = mn(lc, lt);                                          // sweep_adj.c:66
                    if (infv[r]) v = 0.f;                                          // sweep_adj.c:68
                    nv[r] = keep[r] ? lc : v;
                    if (cvalid && rvalid[r]) fsum += lc - nv[r];
                }
                acc += (double)fsum;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    xcp[r] = xqp[r];
                    xcn[r] = xqn[r];
                    pnew[r] = nv[r];
                }
            }
            if constexpr (LSAME) {
                const int ks = cvalid ? (SX > 0 ? tau : nx + 62 - tau) : nkeys;
                stv<R>(gLout + ks * BR + l * R, nv);
            } else {   // staged store, as in fwd_sweep (the piece is read here: reading
                       // it before the arithmetic measured no faster in this sweep)
                const int tp = tau - dl;
                float piece[R];
                ldl<R>(lstage + (tp & (SQ - 1)) * BR, piece);
                stl<R>(lstage + (tau & (SQ - 1)) * BR, nv);
                int ko = (SX > 0 ? tp : nx + 62 - tp) + kshift;
                ko = (unsigned)ko < (unsigned)nkeys ? ko : nkeys;
                stv<R>(gLout + ko * BR + l * R, piece);
            }
            lfloat *rw = (ring_writer && cvalid) ? &ring[beta * kRingQ + (c & (kRingQ - 1))] : ring_dummy;
            *rw = SY > 0 ? nv[R - 1] : nv[0];
            RZ_ADJ_LOAD(up, tau + PS - 1);
            // keeps the reloads in their step (the scheduler would otherwise sink
            // them toward their use, shortening the prefetch distance)
            __builtin_amdgcn_sched_barrier(0);
        }
        pace_epoch<PACE>(pc, g0 / kPF);
    }
    pace_end<PACE>(pc, nep);
#undef RZ_ADJ_LOAD
    return acc;
}

// ---------------------------------------------------------------------------
// Copy a field from layout A to layout B (one wave per band, 8 keys per
// round): each lane gathers its R cells of B key `key` from A key
// key + 2l - 63 (the A dummy block for cells outside the grid), then whole B
// blocks are stored.  With fa2 != nullptr the A blocks are also copied to
// fa2 (the converged field ended in the other A buffer).
// ---------------------------------------------------------------------------
template <int R>
__device__ void copy_a_to_b(const WaveCtx &w, const float *fa, float *fb, float *fa2 = nullptr)
{
    constexpr int BR = kWave * R;
    if (!w.active) return;
    const int nx = w.nx, l = w.lane;
    const int nkeys = nx + 63;
    const long long boff = (long long)w.beta * w.bstride;
    const gfloat *src = gp(fa) + boff;
    gfloat *dst = gpw(fb) + boff;
    for (int g = 0; g < nkeys; g += 8) {
        float v[8][R];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int key = g + u, c = key + l - 63;
            const int ka = (key < nkeys && (unsigned)c < (unsigned)nx) ? c + l : nkeys;
            ldv<R>(src + ka * BR + l * R, v[u]);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) stv<R>(dst + min(g + u, nkeys) * BR + l * R, v[u]);
    }
    if (fa2) {
        gfloat *dst2 = gpw(fa2) + boff;
        for (int g = 0; g < nkeys; g += 8) {
            float v[8][R];
#pragma unroll
            for (int u = 0; u < 8; ++u) ldv<R>(src + min(g + u, nkeys) * BR + l * R, v[u]);
#pragma unroll
            for (int u = 0; u < 8; ++u) stv<R>(dst2 + min(g + u, nkeys) * BR + l * R, v[u]);
        }
    }
}

// The cells whose adjoint value never changes carry the sign bit in the
// adjoint's copy of T (T >= +0 everywhere, so the bit is free): interior
// receiver cells (sweep_adj.c:30) and the frame, rows 0 and NY-1 and
// columns 0 and NX-1 (sweep_adj.c:26-28, the loops run over the interior).
// A sweep's keep test is then one sign test per cell.  Cells outside the
// grid (skew keys, rows past NY) are never read by a grid cell's update, so
// what the adjoint computes for them is never used.
template <int R>
__device__ void mark_kept_cells(const ModelDev &md, float *TA, float *TB)
{
    const int nx = md.nx, ny = md.ny, nframe = 2 * nx + 2 * ny;
    for (int u = threadIdx.x; u < md.nuq + nframe; u += blockDim.x) {
        int k, c;
        if (u < md.nuq) {
            k = md.uq_k[u]; c = md.uq_c[u];
        } else {
            const int v = u - md.nuq;
            if (v < 2 * nx) { k = v < nx ? 0 : ny - 1; c = v < nx ? v : v - nx; }
            else { const int w = v - 2 * nx; k = w < ny ? w : w - ny; c = w < ny ? 0 : nx - 1; }
        }
        const int pa = addr_a<R>(k, c, md.bstride), pb = addr_b<R>(k, c, md.bstride);
        TA[pa] = __uint_as_float(__float_as_uint(TA[pa]) | 0x80000000u);
        TB[pb] = __uint_as_float(__float_as_uint(TB[pb]) | 0x80000000u);
    }
}

// ---------------------------------------------------------------------------
// Forward kernel: one workgroup per shot (eikonal.c:10-76 + calc_FA.c).
// ---------------------------------------------------------------------------
// Wave context of a two-group workgroup: waves [0, nbg) form group 0 (the
// leading sweep of a pair), waves [nbg, 2 nbg) group 1 (the trailing one);
// within a group, wave b handles band b.  Each group has its own LDS ring.
// An unpaired workgroup (groups = 1) is group 0 alone.
__device__ __forceinline__ WaveCtx make_wave(const ModelDev &md, float *ring_all, int groups)
{
    WaveCtx w;
    const int nbg = (blockDim.x >> 6) / groups;                          // bands per group
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    w.nx = md.nx; w.ny = md.ny; w.nb = md.nb;
    w.group = wave / nbg;
    w.beta = wave - w.group * nbg;
    w.lane = threadIdx.x & 63;
    w.active = w.beta < md.nb;
    w.bstride = md.bstride;
    w.ring = ring_all + w.group * (kMaxBands * kRingQ + kWave);
    return w;
}

// A forward sweep that changes no cell leaves every later sweep nothing to
// change: a cell's update (sweep.c:37-97) is min(T, G) with G a function of
// its four neighbours, the same in every direction, so when none of them
// changed since the cell's last update, G is what it was then and T <= G
// already.  With no change in one sweep, induction over the next sweep's
// order gives none in it, and so on.
//
// Whether this workgroup's sweep that just ended changed no cell (cnt: a
// lane's count of steps with a change).
__device__ bool block_unchanged(double cnt)
{
    __shared__ int changed_s;
    if (threadIdx.x == 0) changed_s = 0;
    __syncthreads();
    if (__builtin_amdgcn_ballot_w64(cnt > 0.0) != 0 && (threadIdx.x & 63) == 0) changed_s = 1;
    __syncthreads();
    const bool r = changed_s == 0;
    __syncthreads();
    return r;
}

// Whether the sweep of one of the two wave groups in the phase that just
// ended changed no cell.
__device__ bool pair_unchanged(double cnt, int group)
{
    __shared__ int changed_s[kGroups];
    if (threadIdx.x < kGroups) changed_s[threadIdx.x] = 0;
    __syncthreads();
    if (__builtin_amdgcn_ballot_w64(cnt > 0.0) != 0 && (threadIdx.x & 63) == 0) changed_s[group] = 1;
    __syncthreads();
    const bool r = changed_s[0] == 0 || changed_s[1] == 0;
    __syncthreads();
    return r;
}

// After an iteration with a sweep that changed no cell and a decision to go
// on: the next iteration changes no cell, its norm1 is 0, and 0 < TTNORM
// ends the loop (eikonal.c:44).  Counts that iteration (or stops at the
// iteration limit, as the loop's head would) without sweeping.  Returns
// whether the loop goes on, which it does only when 0 >= TTNORM.
__device__ __forceinline__ bool fixed_point(const ModelDev &md, int &it, float &osum)
{
    if (0.f >= md.ttnorm) return true;
    if (md.max_iter > 0 && it >= md.max_iter) {
        it = -1;
    } else {
        ++it;
        osum = -1.f;   // decided by the zero sum, without an ordered sum
    }
    return false;
}

template <int R, bool PAIRED>
__global__ __launch_bounds__(kGroups * max_bands_for(R) * kWave) void k_forward(
    const ModelDev *__restrict__ mds, const ShotDev *__restrict__ shots, Slots slots,
    const long long *__restrict__ rec_off, const int *__restrict__ shot_base, ShotOut out)
{
    static_assert(PAIRED || max_bands_unpaired(R) <= kGroups * max_bands_for(R), "unpaired bands fit the block");
    __shared__ float ring[kGroups * (kMaxBands * kRingQ + kWave)];
    __shared__ double red[kGroups * kMaxBands];
    __shared__ float stage_all[kGroups * max_bands_for(R) * Stage<R>::WAVE_FLOATS];
    const ShotDev sh = shots[blockIdx.x];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    const WaveCtx w = make_wave(md, ring, PAIRED ? kGroups : 1);
    float *stage = stage_all + (threadIdx.x >> 6) * Stage<R>::WAVE_FLOATS;
    const int nep = sweep_steps(md.nx, md.nb, Stage<R>::HOLD) / kPF;
    const int D = pair_delay(md.nb, Stage<R>::HOLD);
    const long long so = (long long)sh.slot * slots.slot_stride;
    float *TA = slots.TA + so, *TB = slots.TB + so;
    // A-family output of iteration it: W[it & 1] (TA, then LA, then TA ...)
    float *const W[2] = {TA, slots.LA + so};
    const float tinit = (float)(2.0 * (double)md.time);          // zero_travel.c:22
    bool cont = (float)(10.0 * (double)md.time) >= md.ttnorm;    // eikonal.c:39,44
    // Sweep order (+,+) A, (-,+) B, (-,-) A, (+,-) B (eikonal.c:50-53); each
    // reads its own family's layout and writes the next sweep's.  Sweeps 2+3 and
    // sweep 4 + the next iteration's sweep 1 run as pairs (same x
    // direction); that next sweep 1 is run ahead of the stopping test and
    // simply dropped if the loop ends.
    int it = 0, last = -1, nord = 0;
    float osum = -1.f;   // ordered sum of the last decision, if it took one
    [[maybe_unused]] Dec acc_next{0.0, 0.0};  // group 1: the sweep 1 run ahead
    bool fixed = false;                       // a sweep of this iteration changed no cell
    int phases = 0;
    while (cont) {
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        float *Wo = W[it & 1];
        const float *Wi = it > 0 ? W[(it - 1) & 1] : nullptr;   // T at the start (null: the initial field)
        Dec acc{0.0, 0.0};
        if constexpr (!PAIRED) {
            // one group: the four sweeps of the iteration one after the other
            acc = it == 0 ? fwd_sweep<R, +1, +1, true>(w, md.dh, md.two_dh2, TA, md.SA, TB, stage, sh.src_k,
                                                       sh.src_c, tinit, lockstep(0, 0))
                          : fwd_sweep<R, +1, +1, false>(w, md.dh, md.two_dh2, Wi, md.SA, TB, stage, sh.src_k,
                                                        sh.src_c, tinit, lockstep(0, 0));
            // a transposed grid swaps the second and fourth directions: the
            // grid's (-,+) and (+,-) are (+,-) and (-,+) here
            const Dec a2 = md.trans
                ? fwd_sweep<R, +1, -1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k, sh.src_c, tinit,
                                              lockstep(0, 0))
                : fwd_sweep<R, -1, +1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k, sh.src_c, tinit,
                                              lockstep(0, 0));
            acc.sum += a2.sum;
            acc.cnt += a2.cnt;
            phases += 2;
            // After a sweep that changed no cell the rest are skipped (see
            // block_unchanged; an unchanged sweep 1 leaves sweep 2 unchanged
            // too): Wo holds the field of sweep 2, which sweep 3 kept.
            fixed = block_unchanged(a2.cnt);
            if (!fixed) {
                const Dec a3 = fwd_sweep<R, -1, -1, false>(w, md.dh, md.two_dh2, Wo, md.SA, TB, stage, sh.src_k,
                                                           sh.src_c, tinit, lockstep(0, 0));
                acc.sum += a3.sum;
                acc.cnt += a3.cnt;
                ++phases;
                fixed = block_unchanged(a3.cnt);
            }
            if (!fixed) {
                const Dec a4 = md.trans
                    ? fwd_sweep<R, -1, +1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k, sh.src_c,
                                                  tinit, lockstep(0, 0))
                    : fwd_sweep<R, +1, -1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k, sh.src_c,
                                                  tinit, lockstep(0, 0));
                acc.sum += a4.sum;
                acc.cnt += a4.cnt;
                ++phases;
                fixed = block_unchanged(a4.cnt);
            }
        } else {
        phases += (it == 0 ? 2 : 1) + 1;   // sweep 1 alone, pair 2+3, pair 4+1 (less one if skipped)
        if (it == 0) {
            if (w.group == 0)
                acc = fwd_sweep<R, +1, +1, true>(w, md.dh, md.two_dh2, TA, md.SA, TB, stage, sh.src_k, sh.src_c,
                                                 tinit, lockstep(0, 0));
            else
                idle_epochs(nep);
        } else {
            acc = acc_next;
        }
        const Dec a2 = w.group == 0
            ? fwd_sweep<R, -1, +1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k, sh.src_c, tinit,
                                          lockstep(0, D))
            : fwd_sweep<R, -1, -1, false>(w, md.dh, md.two_dh2, Wo, md.SA, TB, stage, sh.src_k, sh.src_c, tinit,
                                          lockstep(D, 0));
        acc.sum += a2.sum;
        acc.cnt += a2.cnt;
        // Sweeps 2 and 3 left the field in Wo (A) and TB (B).  If one of
        // them changed no cell, sweep 4 and the next sweep 1 change none
        // either (see block_unchanged) and are skipped.
        fixed = pair_unchanged(a2.cnt, w.group);
        if (fixed) {
            acc_next = Dec{0.0, 0.0};
            --phases;
        } else if (w.group == 0) {
            const Dec a4 = fwd_sweep<R, +1, -1, false>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k,
                                                       sh.src_c, tinit, lockstep(0, D));
            acc.sum += a4.sum;
            acc.cnt += a4.cnt;
            acc_next = Dec{0.0, 0.0};
            fixed = pair_unchanged(a4.cnt, w.group);
        } else {
            acc_next = fwd_sweep<R, +1, +1, false>(w, md.dh, md.two_dh2, Wo, md.SA, TB, stage, sh.src_k, sh.src_c,
                                                   tinit, lockstep(D, 0));
            fixed = pair_unchanged(acc_next.cnt, w.group);
        }
        }
        const Dec tot = block_sum2(acc, red, nwaves);
        last = it++;
        bool passes;                                              // norm1.c:27, eikonal.c:44
        cont = decide<R, false>(md, tot.sum, tot.cnt, Wo, Wi, tinit, sh.src_k, sh.src_c, passes, nord, osum);
        if (cont && fixed) cont = fixed_point(md, it, osum);
    }
    const int nfield = md.nb * (int)md.bstride;
    if (last >= 0) {  // the last sweep wrote W[last & 1]; the adjoint reads TA and TB
        float *F = W[last & 1];
        if (w.group == 0) copy_a_to_b<R>(w, F, TB, F == TA ? nullptr : TA);
        __syncthreads();
    } else {        // loop never entered: the field is the initial one in both layouts
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
        int kn = k, cn = c;   // the next row of the grid (a column here if transposed)
        if (md.trans) cn = (c <= md.nx - 2) ? c + 1 : (c == md.nx - 1 ? c - 1 : c);
        else kn = (k <= md.ny - 2) ? k + 1 : (k == md.ny - 1 ? k - 1 : k);
        out.tmod[ob + i] = TA[addr_a<R>(k, c, md.bstride)];
        out.tnb[ob + i] = TA[addr_a<R>(kn, cn, md.bstride)];
    }
    __syncthreads();
    mark_kept_cells<R>(md, TA, TB);
    if (threadIdx.x == 0) {
        out.it_f[shot_base[sh.model] + sh.shot] = it;
        out.ord_f[shot_base[sh.model] + sh.shot] = nord;
        out.osum_f[shot_base[sh.model] + sh.shot] = osum;
        out.phases_f[shot_base[sh.model] + sh.shot] = phases;
    }
}

// ---------------------------------------------------------------------------
// Adjoint kernel: one workgroup per shot (eikonal_adj.c:10-59 with
// init_travel_adj.c as an O(ntr) scatter).
// ---------------------------------------------------------------------------
template <int R, int FD, bool CHKINF, bool PAIRED>
__global__ __launch_bounds__(kGroups * max_bands_for(R) * kWave) void k_adjoint(
    const ModelDev *__restrict__ mds, const ShotDev *__restrict__ shots, Slots slots,
    const long long *__restrict__ rec_off, const int *__restrict__ shot_base, const float *__restrict__ tres,
    ShotOut out)
{
    static_assert(PAIRED || max_bands_unpaired(R) <= kGroups * max_bands_for(R), "unpaired bands fit the block");
    constexpr int BR = kWave * R;
    __shared__ float ring[kGroups * (kMaxBands * kRingQ + kWave)];
    __shared__ double red[kGroups * kMaxBands];
    __shared__ float stage_all[kGroups * max_bands_for(R) * Stage<R>::WAVE_FLOATS];
    const ShotDev sh = shots[blockIdx.x];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    const WaveCtx w = make_wave(md, ring, PAIRED ? kGroups : 1);
    float *stage = stage_all + (threadIdx.x >> 6) * Stage<R>::WAVE_FLOATS;
    const int D = pair_delay(md.nb, Stage<R>::HOLD);
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

    // A-family output of iteration it: W[it & 1] (LC, then LA, then LC ...);
    // the first iteration starts from the initial lambda in LA
    float *const W[2] = {slots.LC + so, LA};
    bool cont = big >= md.ttnorm;                                 // eikonal_adj.c:34,37
    int it = 0, last = -1, nord = 0;
    float osum = -1.f;   // ordered sum of the last decision, if it took one
    while (cont) {
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        float *Wo = W[it & 1];
        const float *Wi = it > 0 ? W[(it - 1) & 1] : LA;           // lambda at the start
        // (+,+) A, (+,-) B, (-,+) B, (-,-) A (eikonal_adj.c:43-46): sweep 1
        // reads lambda in A and writes B, sweep 2 updates B in place, sweep 3
        // reads B and writes A, sweep 4 updates A in place; sweeps 1+2 and
        // 3+4 run as pairs (same x direction).
        double acc;
        if constexpr (!PAIRED) {   // one group: the four sweeps one after the other
            acc = adj_sweep<R, +1, +1, false, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wi, LB, stage, lockstep(0, 0));
            // a transposed grid swaps the second and third directions
            if (md.trans) {
                acc += adj_sweep<R, -1, +1, true, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, LB, stage, lockstep(0, 0));
                acc += adj_sweep<R, +1, -1, false, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, Wo, stage, lockstep(0, 0));
            } else {
                acc += adj_sweep<R, +1, -1, true, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, LB, stage, lockstep(0, 0));
                acc += adj_sweep<R, -1, +1, false, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, Wo, stage, lockstep(0, 0));
            }
            acc += adj_sweep<R, -1, -1, true, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wo, Wo, stage, lockstep(0, 0));
        } else if (w.group == 0) {
            acc = adj_sweep<R, +1, +1, false, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wi, LB, stage, lockstep(0, D));
            acc += adj_sweep<R, -1, +1, false, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, Wo, stage, lockstep(0, D));
        } else {
            acc = adj_sweep<R, +1, -1, true, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, LB, stage, lockstep(D, 0));
            acc += adj_sweep<R, -1, -1, true, FD, CHKINF>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wo, Wo, stage, lockstep(D, 0));
        }
        const double tot = block_sum(acc, red, nwaves);
        last = it++;
        bool passes;                                              // norm1.c:27, eikonal_adj.c:37
        cont = decide<R, CHKINF>(md, tot, -1.0, Wo, Wi, 0.f, -1, -1, passes, nord, osum);
    }
    if (threadIdx.x == 0) {
        const int g = shot_base[sh.model] + sh.shot;
        out.it_a[g] = it;
        out.lsel[g] = (last >= 0 && (last & 1) == 0) ? 1 : 0;   // the converged lambda is in LC
        out.ord_a[g] = nord;
        out.osum_a[g] = osum;
    }
}

// ---------------------------------------------------------------------------
// Split shots: the two sweeps of a pair run in two workgroups, each on a CU
// of its own, instead of two wave groups of one workgroup.  A launch has
// 16 workgroups per 8 shots; workgroup b serves shot (b/16)*8 + b%8 as the
// leading (role 0, (b/8) even) or trailing (role 1) group.  Workgroups are
// dealt to the XCDs round-robin by index, so the two of a shot (b and b+8)
// share an XCD and its L2, through which they exchange fields, progress
// counters and partial sums; each checks this at its start (XCC_ID), and
// the launch reports failure if it does not hold (the host then repeats
// the work with the one-workgroup kernels).  Per shot: kSyncInts ints of
// counters (one 64-byte line each) and 16 doubles of partial sums.
// ---------------------------------------------------------------------------
constexpr int kSyncLead = 0, kSyncDone0 = 16, kSyncDone1 = 32, kSyncXcc = 48;

__device__ __forceinline__ WaveCtx make_wave_split(const ModelDev &md, float *ring, int role)
{
    WaveCtx w;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    w.nx = md.nx; w.ny = md.ny; w.nb = md.nb;
    w.group = role;
    w.beta = wave;
    w.lane = threadIdx.x & 63;
    w.active = w.beta < md.nb;
    w.bstride = md.bstride;
    w.ring = ring;
    return w;
}

// Both workgroups of a shot have started and run on the same XCD.
__device__ bool pair_handshake(int *cnt, int role, int *err)
{
    if (threadIdx.x == 0) {
        const int xcc = __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15;   // HW_REG_XCC_ID[3:0]
        __hip_atomic_store(cnt + kSyncXcc + role, xcc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!pace_wait(cnt + kSyncXcc + (1 - role), 1, err)) return false;
    __shared__ int same_s;
    if (threadIdx.x == 0) {
        const int a = __hip_atomic_load(cnt + kSyncXcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int b = __hip_atomic_load(cnt + kSyncXcc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        same_s = (a == b);
        if (a != b) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return same_s != 0;
}

// End of a pair phase: each workgroup drains its stores, says it is done
// and waits for the other (their next sweeps read what both wrote).
__device__ bool rendezvous(int *cnt, int role, int phase, int *err)
{
    drain_vmem();
    __syncthreads();
    publish(cnt + (role ? kSyncDone1 : kSyncDone0), phase + 1);
    return pace_wait(cnt + (role ? kSyncDone0 : kSyncDone1), phase + 1, err);
}

// block_sum() over the waves of both workgroups, in the order of the
// one-workgroup kernel (group 0's waves, then group 1's), with the
// rendezvous that ends the phase.
__device__ double pair_sum(double v, double *red, int nwaves, double *part, int *cnt, int role, int phase,
                           int *err, bool &ok)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < nwaves; ++i) part[role * kMaxBands + i] = red[i];
    ok = rendezvous(cnt, role, phase, err);
    double t = 0.0;
    for (int i = 0; i < nwaves; ++i) t += part[i];
    for (int i = 0; i < nwaves; ++i) t += part[kMaxBands + i];
    __syncthreads();
    return t;
}

// pair_sum of the decreases, with the waves' counts summed alongside (part
// holds the counts after the 2*kMaxBands partial sums).
__device__ Dec pair_sum2(Dec v, double *red, int nwaves, double *part, int *cnt, int role, int phase, int *err,
                         bool &ok)
{
    double c = v.cnt;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) part[2 * kMaxBands + role * kMaxBands + (threadIdx.x >> 6)] = c;
    Dec r;
    r.sum = pair_sum(v.sum, red, nwaves, part, cnt, role, phase, err, ok);
    double t = 0.0;
    for (int i = 0; i < nwaves; ++i) t += part[2 * kMaxBands + i] + part[3 * kMaxBands + i];
    __syncthreads();
    r.cnt = t;
    return r;
}

__device__ __forceinline__ bool split_role(int nshots, int &shot, int &role)
{
    const int b = blockIdx.x;
    role = (b >> 3) & 1;
    shot = (b >> 4) * 8 + (b & 7);
    return shot < nshots;
}

template <int R>
__global__ __launch_bounds__(max_bands_for(R) * kWave, max_bands_for(R) / 2) void k_forward_split(
    const ModelDev *__restrict__ mds, const ShotDev *__restrict__ shots, int nshots, Slots slots,
    const long long *__restrict__ rec_off, const int *__restrict__ shot_base, ShotOut out, SplitSync ss)
{
    __shared__ float ring[kMaxBands * kRingQ + kWave];
    __shared__ double red[kMaxBands];
    __shared__ float stage_all[max_bands_for(R) * Stage<R>::WAVE_FLOATS];
    int s, role;
    if (!split_role(nshots, s, role)) return;
    const ShotDev sh = shots[s];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    int *cnt = ss.cnt + (long long)s * kSyncInts;
    double *part = ss.part + (long long)s * kSyncPart;
    if (!pair_handshake(cnt, role, ss.err)) return;
    const WaveCtx w = make_wave_split(md, ring, role);
    float *stage = stage_all + (threadIdx.x >> 6) * Stage<R>::WAVE_FLOATS;
    Pace pc;
    pc.pre = pc.post = 0; pc.lead = cnt + kSyncLead; pc.err = ss.err; pc.base = 0;
    pc.D = pair_delay(md.nb, Stage<R>::HOLD);
    const long long so = (long long)sh.slot * slots.slot_stride;
    float *TA = slots.TA + so, *TB = slots.TB + so;
    float *const W[2] = {TA, slots.LA + so};                      // as in k_forward
    const float tinit = (float)(2.0 * (double)md.time);          // zero_travel.c:22
    bool cont = (float)(10.0 * (double)md.time) >= md.ttnorm;    // eikonal.c:39,44
    // The sweeps and pairs of k_forward: role 0 runs sweep 1 of the first
    // iteration alone, then sweeps 2 and 4; role 1 runs sweeps 3 and the
    // next iteration's sweep 1.
    int it = 0, last = -1, phase = 0, nord = 0;
    float osum = -1.f;
    bool ok = true;
    Dec acc_next{0.0, 0.0};
    int phases = 0;
    while (cont) {
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        float *Wo = W[it & 1];
        const float *Wi = it > 0 ? W[(it - 1) & 1] : nullptr;
        Dec acc{0.0, 0.0};
        phases += (it == 0 ? 2 : 1) + 1;
        if (it == 0) {
            if (role == 0) {
                pc.base = phase * kPhaseStride;
                acc = fwd_sweep<R, +1, +1, true, kPaceLead>(w, md.dh, md.two_dh2, TA, md.SA, TB, stage, sh.src_k,
                                                            sh.src_c, tinit, pc);
            }
            ok = rendezvous(cnt, role, phase++, ss.err);
            if (!ok) break;
        } else {
            acc = acc_next;
        }
        pc.base = phase * kPhaseStride;
        const Dec a2 = role == 0
            ? fwd_sweep<R, -1, +1, false, kPaceLead>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage, sh.src_k,
                                                     sh.src_c, tinit, pc)
            : fwd_sweep<R, -1, -1, false, kPaceTrail>(w, md.dh, md.two_dh2, Wo, md.SA, TB, stage, sh.src_k,
                                                      sh.src_c, tinit, pc);
        acc.sum += a2.sum;
        acc.cnt += a2.cnt;
        // as in k_forward: if sweep 2 or 3 changed no cell, sweep 4 and the
        // next sweep 1 are skipped (each workgroup tells the other, in part,
        // whether its sweep changed a cell)
        bool mine = block_unchanged(a2.cnt);
        if (threadIdx.x == 0) part[kPartFixed + role] = mine ? 1.0 : 0.0;
        ok = rendezvous(cnt, role, phase++, ss.err);
        if (!ok) break;
        bool fixed = mine || part[kPartFixed + 1 - role] != 0.0;
        pc.base = phase * kPhaseStride;
        if (fixed) {
            acc_next = Dec{0.0, 0.0};
            --phases;
        } else if (role == 0) {
            const Dec a4 = fwd_sweep<R, +1, -1, false, kPaceLead>(w, md.dh, md.two_dh2, TB, md.SB, Wo, stage,
                                                                  sh.src_k, sh.src_c, tinit, pc);
            acc.sum += a4.sum;
            acc.cnt += a4.cnt;
            acc_next = Dec{0.0, 0.0};
            mine = block_unchanged(a4.cnt);
        } else {
            acc_next = fwd_sweep<R, +1, +1, false, kPaceTrail>(w, md.dh, md.two_dh2, Wo, md.SA, TB, stage,
                                                               sh.src_k, sh.src_c, tinit, pc);
            mine = block_unchanged(acc_next.cnt);
        }
        if (!fixed && threadIdx.x == 0) part[kPartFixed + 2 + role] = mine ? 1.0 : 0.0;
        const Dec tot = pair_sum2(acc, red, nwaves, part, cnt, role, phase++, ss.err, ok);
        if (!ok) break;
        if (!fixed) fixed = mine || part[kPartFixed + 2 + 1 - role] != 0.0;
        last = it++;
        bool passes;                                              // norm1.c:27, eikonal.c:44
        cont = decide<R, false>(md, tot.sum, tot.cnt, Wo, Wi, tinit, sh.src_k, sh.src_c, passes, nord, osum);
        if (cont && fixed) cont = fixed_point(md, it, osum);
        if (passes) {
            // both workgroups read the whole grid alike; they meet before the
            // next iteration's sweep 2 overwrites Wi
            ok = rendezvous(cnt, role, phase++, ss.err);
            if (!ok) break;
        }
    }
    if (role != 0) return;
    const int nfield = md.nb * (int)md.bstride;
    if (last >= 0) {  // the last sweep wrote W[last & 1]; the adjoint reads TA and TB
        float *F = W[last & 1];
        copy_a_to_b<R>(w, F, TB, F == TA ? nullptr : TA);
        __syncthreads();
    } else {        // loop never entered: the field is the initial one in both layouts
        for (int p = threadIdx.x; p < nfield; p += blockDim.x) {
            const int b = p / (int)md.bstride, q = p - b * (int)md.bstride;
            const int key = q / (kWave * R), wv = q - key * (kWave * R), l = wv / R, r = wv - l * R;
            const int k = b * kWave * R + l * R + r;
            TA[p] = (k == sh.src_k && key - l == sh.src_c) ? 0.f : tinit;
            TB[p] = (k == sh.src_k && key + l - 63 == sh.src_c) ? 0.f : tinit;
        }
        __syncthreads();
    }
    const long long ob = rec_off[sh.model] + (long long)sh.shot * md.ntr;
    for (int i = threadIdx.x; i < md.ntr; i += blockDim.x) {
        const int k = md.rec_k[i], c = md.rec_c[i];
        const int kn = (k <= md.ny - 2) ? k + 1 : (k == md.ny - 1 ? k - 1 : k);
        out.tmod[ob + i] = TA[addr_a<R>(k, c, md.bstride)];
        out.tnb[ob + i] = TA[addr_a<R>(kn, c, md.bstride)];
    }
    __syncthreads();
    mark_kept_cells<R>(md, TA, TB);
    if (threadIdx.x == 0) {
        out.it_f[shot_base[sh.model] + sh.shot] = it;
        out.ord_f[shot_base[sh.model] + sh.shot] = nord;
        out.osum_f[shot_base[sh.model] + sh.shot] = osum;
        out.phases_f[shot_base[sh.model] + sh.shot] = phases;
    }
}

template <int R, int FD, bool CHKINF>
__global__ __launch_bounds__(max_bands_for(R) * kWave, max_bands_for(R) / 2) void k_adjoint_split(
    const ModelDev *__restrict__ mds, const ShotDev *__restrict__ shots, int nshots, Slots slots,
    const long long *__restrict__ rec_off, const int *__restrict__ shot_base, const float *__restrict__ tres,
    ShotOut out, SplitSync ss)
{
    constexpr int BR = kWave * R;
    __shared__ float ring[kMaxBands * kRingQ + kWave];
    __shared__ double red[kMaxBands];
    __shared__ float stage_all[max_bands_for(R) * Stage<R>::WAVE_FLOATS];
    int s, role;
    if (!split_role(nshots, s, role)) return;
    const ShotDev sh = shots[s];
    const ModelDev md = mds[sh.model];
    const int nwaves = blockDim.x >> 6;
    int *cnt = ss.cnt + (long long)s * kSyncInts;
    double *part = ss.part + (long long)s * kSyncPart;
    if (!pair_handshake(cnt, role, ss.err)) return;
    const WaveCtx w = make_wave_split(md, ring, role);
    float *stage = stage_all + (threadIdx.x >> 6) * Stage<R>::WAVE_FLOATS;
    Pace pc;
    pc.pre = pc.post = 0; pc.lead = cnt + kSyncLead; pc.err = ss.err; pc.base = 0;
    pc.D = pair_delay(md.nb, Stage<R>::HOLD);
    const long long so = (long long)sh.slot * slots.slot_stride;
    const float *TA = slots.TA + so, *TB = slots.TB + so;
    float *LA = slots.LA + so, *LB = slots.LB + so;
    const float big = (float)(10.0 * (double)md.time);
    if (role == 0) {   // init_travel_adj.c:22-58, before the leading sweep's first publication
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
    }
    float *const W[2] = {slots.LC + so, LA};                      // as in k_adjoint
    bool cont = big >= md.ttnorm;                                 // eikonal_adj.c:34,37
    int it = 0, last = -1, phase = 0, nord = 0;
    float osum = -1.f;
    bool ok = true;
    while (cont) {
        if (md.max_iter > 0 && it >= md.max_iter) { it = -1; break; }
        float *Wo = W[it & 1];
        const float *Wi = it > 0 ? W[(it - 1) & 1] : LA;
        double acc;
        pc.base = phase * kPhaseStride;
        if (role == 0)
            acc = adj_sweep<R, +1, +1, false, FD, CHKINF, kPaceLead>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wi, LB,
                                                                      stage, pc);
        else
            acc = adj_sweep<R, +1, -1, true, FD, CHKINF, kPaceTrail>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, LB,
                                                                      stage, pc);
        ok = rendezvous(cnt, role, phase++, ss.err);
        if (!ok) break;
        pc.base = phase * kPhaseStride;
        if (role == 0)
            acc += adj_sweep<R, -1, +1, false, FD, CHKINF, kPaceLead>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TB, LB, Wo,
                                                                       stage, pc);
        else
            acc += adj_sweep<R, -1, -1, true, FD, CHKINF, kPaceTrail>(w, md.dh, md.rdh, md.rdh_hi, md.rdh_lo, md.eps_adj, TA, Wo, Wo,
                                                                       stage, pc);
        const double tot = pair_sum(acc, red, nwaves, part, cnt, role, phase++, ss.err, ok);
        if (!ok) break;
        last = it++;
        bool passes;                                              // norm1.c:27, eikonal_adj.c:37
        cont = decide<R, CHKINF>(md, tot, -1.0, Wo, Wi, 0.f, -1, -1, passes, nord, osum);
        if (passes) {   // as in k_forward_split
            ok = rendezvous(cnt, role, phase++, ss.err);
            if (!ok) break;
        }
    }
    if (role == 0 && threadIdx.x == 0) {
        const int g = shot_base[sh.model] + sh.shot;
        out.it_a[g] = it;
        out.lsel[g] = (last >= 0 && (last & 1) == 0) ? 1 : 0;
        out.ord_a[g] = nord;
        out.osum_a[g] = osum;
    }
}

// Zero the split shots' counters before a launch (a kernel, so that graphs
// can hold it).
__global__ void k_split_reset(int *cnt, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        cnt[i] = 0;
}

// ---------------------------------------------------------------------------
// Residuals, adjoint sources and misfit (calc_FA_res.c:10-77) for shots
// [s0, s1) of one model.  carry (ntr) is the reference's Tres array: a muted
// pick keeps the previous value, so each receiver is a chain over the shots
// (one thread per receiver, loads issued ahead of the chain).
// direct != nullptr: the residual Tmod - Tobs is given (an adjoint source
// dL/dTmod); no pick is muted and the slope scaling is applied as usual.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float residual_of(const ModelDev &md, int k, int ny, float tm, float tn, float r)
{
    // calc_FA_res.c:52-69: divide by the vertical slope at the receiver
    // (k: its row, ny: the grid's rows)
    float g = 0.f;
    bool have = true;
    if (k <= ny - 2) g = (tn - tm) / md.dh;
    else if (k == ny - 1) g = (tm - tn) / md.dh;
    else have = false;
    return have ? ((fabsf(g) > 0.f) ? r / g : 0.f) : r;
}

__global__ void k_residual(const ModelDev *__restrict__ mds, int model, int ntr, int s0, int s1,
                           long long rec_off, const float *__restrict__ tmod, const float *__restrict__ tnb,
                           const float *__restrict__ tobs, const float *__restrict__ direct, float *carry,
                           float *tres)
{
    const ModelDev md = mds[model];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntr) return;
    const int k = md.trans ? md.rec_c[i] : md.rec_k[i], ny = md.trans ? md.nx : md.ny;   // the grid's row
    float cr = carry[i];
    constexpr int U = 8;
    for (int s = s0; s < s1; s += U) {
        float to[U], tm[U], tn[U], dd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long x = rec_off + (long long)min(s + u, s1 - 1) * ntr + i;
            to[u] = direct ? 0.f : tobs[x];
            tm[u] = tmod[x];
            tn[u] = tnb[x];
            dd[u] = direct ? direct[x] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (s + u >= s1) break;
            if (to[u] >= 0.f) {                                   // calc_FA_res.c:39
                const float r = direct ? dd[u] : tm[u] - to[u];   // calc_FA_res.c:42
                cr = residual_of(md, k, ny, tm[u], tn[u], r);
            }
            tres[rec_off + (long long)(s + u) * ntr + i] = cr;
        }
    }
    carry[i] = cr;
}

// Misfit of shot s0 + blockIdx.x: sum over receivers, in receiver order, of
// (double)(r*r) for unmuted picks (calc_FA_res.c:47).  One wave: the lanes
// form 64 products at a time, lane 0 adds them in order.
__global__ void k_misfit_shot(int ntr, int s0, long long rec_off, const float *__restrict__ tmod,
                              const float *__restrict__ tobs, const float *__restrict__ direct, double *l2_shot)
{
    const int s = s0 + blockIdx.x, l = threadIdx.x;
    double l2 = 0.0;
    for (int i0 = 0; i0 < ntr; i0 += kWave) {
        const int i = i0 + l;
        float rr = 0.f;
        bool use = false;
        if (i < ntr) {
            const long long x = rec_off + (long long)s * ntr + i;
            const float to = direct ? 0.f : tobs[x];
            if (to >= 0.f) {
                const float r = direct ? direct[x] : tmod[x] - to;
                rr = r * r;
                use = true;
            }
        }
        const unsigned long long m = __builtin_amdgcn_ballot_w64(use);
        const int n = min(kWave, ntr - i0);
        for (int j = 0; j < n; ++j) {
            const float v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rr), j));
            if ((m >> j) & 1ull) l2 += (double)v;
        }
    }
    if (l == 0) l2_shot[s] = l2;
}

// grad_obj.c:42: the evaluation's sum of per-shot misfits, in shot order.
__global__ void k_misfit_sum(int s0, int s1, const double *__restrict__ l2_shot, double *l2_sum)
{
    double t = *l2_sum;
    for (int s = s0; s < s1; ++s) t += l2_shot[s];
    *l2_sum = t;
}

// Start of an evaluation: the residual carry from its initial value, the
// misfit sums and the gradient to zero (kernels rather than memsets, so
// that an evaluation can be recorded in a HIP graph).
__global__ void k_eval_init(float *carry, const float *carry0, int ntr_total, double *l2_sum, int nmodels,
                            float *grad, long long ncells, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) {
        if (i < ncells) grad[i] = 0.f;
        if (i < ntr_total) carry[i] = carry0[i];
        if (i < nmodels) l2_sum[i] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Gradient assembly and shot sum (ass_grad.c:20-46 + cp_grad_frame.c):
// grad(k,c) += (float)(lambda_s(k',c') * S(k',c')^3) over the given slots in
// order, (k',c') = (k,c) clamped to rows 2..NY-3 and columns 2..NX-3.
// ---------------------------------------------------------------------------
template <int R>
__global__ void k_gradient(const ModelDev *__restrict__ mds, int model, const int *__restrict__ slot_of,
                           const int *__restrict__ lsel, int nsh, Slots slots, float *grad, int accumulate,
                           TaperDev tp)
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
        // the gradient is row-major in the grid's orientation: (ko, co) is
        // the cell's row and column there
        const int ko = md.trans ? c : k, co = md.trans ? k : c;
        const long long cell = (long long)ko * (md.trans ? md.ny : md.nx) + co;
        float g = accumulate ? grad[cell] : 0.f;
        if (tp.mode == 0) {
            // unrolled so that several shots' loads are in flight; the sum
            // stays in shot order
#pragma unroll 8
            for (int j = 0; j < nsh; ++j) {
                const float *L = lsel[j] ? slots.LC : slots.LA;   // where the shot's lambda converged
                const float lam = L[(long long)slot_of[j] * slots.slot_stride + pos];
                g += (float)((double)lam * s3);
            }
        } else {
            // the taper acts on the frame-copied value at the cell's own position
            const unsigned r0 = tp.start[cell], r1 = tp.start[cell + 1];
            for (int j = 0; j < nsh; ++j) {
                const float *L = lsel[j] ? slots.LC : slots.LA;
                const float lam = L[(long long)slot_of[j] * slots.slot_stride + pos];
                float v = (float)((double)lam * s3);
                if (tp.mode == 2) {
                    if (r1 > r0) v = 1e-5f;
                } else {
                    const int *bx = tp.box + 4 * (tp.src0 + j);
                    if (ko >= bx[0] && ko <= bx[1] && co >= bx[2] && co <= bx[3])
                        v *= tp.fsrc[tp.off[tp.src0 + j] + (long long)(ko - bx[0]) * (bx[3] - bx[2] + 1) +
                                     (co - bx[2])];
                    for (unsigned i = r0; i < r1; ++i) v *= tp.frec[i];
                }
                g += v;
            }
        }
        grad[cell] = g;
    }
}

// cp_grad_frame.c:10-53: rows 0-1 and ny-2..ny-1 take row 2 / ny-3, then
// columns 0-1 and nx-2..nx-1 take column 2 / nx-3; for nx, ny >= 5 every
// frame cell thus copies the interior cell at its clamped position, which
// the copy does not change.
__global__ void k_grad_frame(int nx, int ny, float *grad)
{
    const long long n = (long long)nx * ny;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nx), c = (int)(i - (long long)k * nx);
        const int kk = clampi(k, 2, ny - 3), cc = clampi(c, 2, nx - 3);
        if (kk != k || cc != c) grad[i] = grad[(long long)kk * nx + cc];
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
            if (va < md.s_min_ft) atomicOr(err_flag, 2);   // too small for the float T-only quotients
        }
        if (k < md.ny && cb >= 0 && cb < md.nx) vb = md.S[(long long)k * md.nx + cb];
        SA[p] = va;
        SB[p] = vb;
    }
}

// Layout-A field -> row-major (optionally clearing the receiver sign bits).
template <int R>
__global__ void k_field_to_rowmajor(int nx, int ny, long long bstride, const float *field, const float *alt,
                                    const int *sel, float *out, int strip)
{
    if (sel && *sel) field = alt;
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

// out (cols x rows) = the transpose of in (rows x cols), row-major: the
// fields of a grid stored transposed, to and from the caller's orientation.
__global__ void k_transpose(const float *__restrict__ in, float *__restrict__ out, int rows, int cols)
{
    const long long n = (long long)rows * cols;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / cols), c = (int)(i - (long long)r * cols);
        out[(long long)c * rows + r] = in[i];
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

hipError_t launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st)
{
    hipLaunchKernelGGL(k_transpose, dim3(grid_for((long long)rows * cols, 256)), dim3(256), 0, st, in, out, rows,
                       cols);
    return hipGetLastError();
}

hipError_t launch_prepare_slowness(const ModelDev *mh, int model, const ModelDev *md, float *SA, float *SB,
                                   int *err_flag, hipStream_t st)
{
    const long long n = (long long)mh->nb * mh->bstride;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_prepare_slowness<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st,
                                            md, model, SA, SB, err_flag));
    return hipGetLastError();
}

hipError_t launch_forward(int R, int nb_max, bool paired, const ModelDev *md, const ShotDev *shots, int nshots,
                          Slots slots, const long long *rec_off, const int *shot_base, ShotOut out, hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    if (paired && !paired_for(R, nb_max)) return hipErrorInvalidValue;
    if (!paired) {   // unpaired: only at R = 8, one wave per band
        if (R != 8 || nb_max > max_bands_unpaired(R)) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_forward<8, false>), dim3(nshots), dim3(nb_max * kWave), 0, st, md, shots, slots, rec_off,
                           shot_base, out);
        return hipGetLastError();
    }
    RZ_DISPATCH_R(R, hipLaunchKernelGGL((k_forward<RR, true>), dim3(nshots), dim3(kGroups * nb_max * kWave), 0, st,
                                        md, shots, slots, rec_off, shot_base, out));
    return hipGetLastError();
}

hipError_t launch_forward_split(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots,
                                Slots slots, const long long *rec_off, const int *shot_base, ShotOut out,
                                SplitSync ss, hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    const long long n = (long long)nshots * kSyncInts;
    hipLaunchKernelGGL(k_split_reset, dim3((unsigned)std::min<long long>((n + 255) / 256, 1024)), dim3(256), 0, st,
                       ss.cnt, n);
    const dim3 g(16 * ((nshots + 7) / 8)), b(nb_max * kWave);
    RZ_DISPATCH_R(R, hipLaunchKernelGGL(k_forward_split<RR>, g, b, 0, st, md, shots, nshots, slots, rec_off,
                                        shot_base, out, ss));
    return hipGetLastError();
}

hipError_t launch_adjoint_split(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots,
                                Slots slots, const long long *rec_off, const int *shot_base, const float *tres,
                                ShotOut out, int variant, SplitSync ss, hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    const long long n = (long long)nshots * kSyncInts;
    hipLaunchKernelGGL(k_split_reset, dim3((unsigned)std::min<long long>((n + 255) / 256, 1024)), dim3(256), 0, st,
                       ss.cnt, n);
    const dim3 g(16 * ((nshots + 7) / 8)), b(nb_max * kWave);
#define RZ_ADJ_LAUNCH(FD_, CI_)                                                                           \
    RZ_DISPATCH_R(R, hipLaunchKernelGGL((k_adjoint_split<RR, FD_, CI_>), g, b, 0, st, md, shots, nshots, \
                                        slots, rec_off, shot_base, tres, out, ss))
    switch ((variant & kAdjFastDiv) ? ((variant & kAdjFloatT) ? 2 : 1) : 0) {   // DH division mode
    case 0:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(0, true); } else { RZ_ADJ_LAUNCH(0, false); }
        break;
    case 1:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(1, true); } else { RZ_ADJ_LAUNCH(1, false); }
        break;
    default:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(2, true); } else { RZ_ADJ_LAUNCH(2, false); }
        break;
    }
#undef RZ_ADJ_LAUNCH
    return hipGetLastError();
}

hipError_t launch_adjoint(int R, int nb_max, bool paired, const ModelDev *md, const ShotDev *shots, int nshots,
                          Slots slots, const long long *rec_off, const int *shot_base, const float *tres, ShotOut out,
                          int variant, hipStream_t st)
{
    if (nshots <= 0) return hipSuccess;
    if (paired && !paired_for(R, nb_max)) return hipErrorInvalidValue;
    if (!paired && (R != 8 || nb_max > max_bands_unpaired(R))) return hipErrorInvalidValue;
    const dim3 g(nshots), b((paired ? kGroups : 1) * nb_max * kWave);
    // unpaired: only at R = 8, one wave per band
#define RZ_ADJ_LAUNCH(FD_, CI_)                                                                         \
    if (!paired) {                                                                                      \
        hipLaunchKernelGGL((k_adjoint<8, FD_, CI_, false>), g, b, 0, st, md, shots, slots, rec_off,    \
                           shot_base, tres, out);                                                       \
    } else {                                                                                            \
        RZ_DISPATCH_R(R, hipLaunchKernelGGL((k_adjoint<RR, FD_, CI_, true>), g, b, 0, st, md, shots,   \
                                            slots, rec_off, shot_base, tres, out))                      \
    }
    switch ((variant & kAdjFastDiv) ? ((variant & kAdjFloatT) ? 2 : 1) : 0) {   // DH division mode
    case 0:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(0, true); } else { RZ_ADJ_LAUNCH(0, false); }
        break;
    case 1:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(1, true); } else { RZ_ADJ_LAUNCH(1, false); }
        break;
    default:
        if (variant & kAdjInfGuard) { RZ_ADJ_LAUNCH(2, true); } else { RZ_ADJ_LAUNCH(2, false); }
        break;
    }
#undef RZ_ADJ_LAUNCH
    return hipGetLastError();
}

hipError_t launch_residual(const ModelDev *md, int model, int ntr, int s0, int s1, long long rec_off,
                           const float *tmod, const float *tnb, const float *tobs, const float *direct,
                           float *carry, float *tres, double *l2_shot, double *l2_sum, hipStream_t st)
{
    if (s1 <= s0) return hipSuccess;
    if (ntr > 0) {
        hipLaunchKernelGGL(k_residual, dim3((ntr + 63) / 64), dim3(64), 0, st, md, model, ntr, s0, s1, rec_off, tmod,
                           tnb, tobs, direct, carry, tres);
        hipLaunchKernelGGL(k_misfit_shot, dim3(s1 - s0), dim3(kWave), 0, st, ntr, s0, rec_off, tmod, tobs, direct,
                           l2_shot);
    }
    hipLaunchKernelGGL(k_misfit_sum, dim3(1), dim3(1), 0, st, s0, s1, l2_shot, l2_sum);
    return hipGetLastError();
}

hipError_t launch_eval_init(float *carry, const float *carry0, int ntr_total, double *l2_sum, int nmodels,
                            float *grad, long long ncells, hipStream_t st)
{
    const long long n = std::max<long long>(std::max<long long>(ncells, ntr_total), nmodels);
    hipLaunchKernelGGL(k_eval_init, dim3(grid_for(n, 256)), dim3(256), 0, st, carry, carry0, ntr_total, l2_sum,
                       nmodels, grad, ncells, n);
    return hipGetLastError();
}

hipError_t launch_gradient(const ModelDev *mh, int model, const ModelDev *md, const int *slot_of, const int *lsel,
                           int nsh, Slots slots, float *grad, int accumulate, const TaperDev &tp, hipStream_t st)
{
    const long long n = (long long)mh->nb * mh->bstride;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_gradient<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st, md, model,
                                            slot_of, lsel, nsh, slots, grad, accumulate, tp));
    return hipGetLastError();
}

hipError_t launch_grad_frame(int nx, int ny, float *grad, hipStream_t st)
{
    hipLaunchKernelGGL(k_grad_frame, dim3(grid_for((long long)nx * ny, 256)), dim3(256), 0, st, nx, ny, grad);
    return hipGetLastError();
}

hipError_t launch_field_to_rowmajor(const ModelDev *mh, const float *field, const float *alt, const int *sel,
                                    float *out, int strip, hipStream_t st)
{
    const long long n = (long long)mh->nx * mh->ny;
    RZ_DISPATCH_R(mh->R, hipLaunchKernelGGL(k_field_to_rowmajor<RR>, dim3(grid_for(n, 256)), dim3(256), 0, st,
                                            mh->nx, mh->ny, mh->bstride, field, alt, sel, out, strip));
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
