// The l-BFGS two-loop recursion on the device (LBFGS.c:12-226, fd.h:88),
// with the reference's float arithmetic, and dotp.c's serial float dot
// product (dotp.c:22-36) computed exactly.
//
// dotp() sums tmp += a[i]*b[i] in float, i ascending: a chain of n dependent
// float additions.  While the running sum s keeps its sign and stays in one
// binade [2^e, 2^(e+1)) (below 2^-125 the floats form one grid of step
// 2^-149), the floats there are the multiples of u = 2^(e-23), and s + d is
// s plus d rounded to a multiple of u: floor(d/u), plus one where the
// remainder exceeds one half, and a remainder of exactly one half (a tie)
// rounds to the even neighbour, which depends only on the parity of s/u.
// A run of terms is therefore, for a given binade, a function of the
// incoming parity (increment, outgoing parity, and the least and largest
// partial sums, which tell whether the run stays in the binade at a given
// incoming sum), and those functions compose.
//
// Phase A, on every CU: the double prefix sums of the terms predict the
// binade the float sum has at the start of each run of 64 terms; each
// run's function at that binade is computed, and composed into runs of
// 1024 and 8192 terms where all parts predict the same binade.  Phase B,
// one workgroup: thread 0 carries the float sum through the runs in order;
// a run whose function applies (right binade, and the sum stays in it) is
// one step, exactly; otherwise its parts are tried, and a 64-term run that
// leaves the binade is added term by term as floats, the reference's own
// operation.  Every partial sum is therefore the serial one; the
// predictions only decide how fast.  A sum that becomes infinite or NaN
// stays so (a NaN result is a NaN; its payload bits are not the x86
// build's).
//
// The vectors are held in the reference's h order (x outer, y inner; the
// index h of LBFGS.c:45-58), each padded with zeros to whole 8192-term runs
// (adding +0 changes no float sum, and the running sum is never -0).
// Vector updates (LBFGS.c:119-124,128-131,143-148) are elementwise over all
// CUs with separate multiplies and adds, as the reference's build has no
// FMA contraction.  The scalars (gamma, rho, alpha, beta) are computed from
// the dot products where they are used, with the reference's operations:
// rho = (float)(1.0/(double)ys) when fabs(ys) > 0.0 (LBFGS.c:83-88).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rajzel_amd.h"
#include "rz_internal.h"

namespace rz {
namespace {

constexpr unsigned kTie = 1u, kUp = 2u, kBig = 4u;

typedef float vf4 __attribute__((ext_vector_type(4)));

// d relative to a running sum of sign sgn (0 or the sign bit) in the grid of
// u = 2^(e-23): q = sgn*d/u, f = floor(q) and the class of the remainder
// (kUp above one half, kTie exactly one half); kBig where |q| >= 2^24 or d is
// not finite (such a term always leaves the binade).
__device__ __forceinline__ void classify(float d, unsigned sgn, int e, int &f, unsigned &fl)
{
    const unsigned b = __float_as_uint(d) ^ sgn;
    const int E = (b >> 23) & 0xFF;
    const unsigned M = (b & 0x7FFFFFu) | (E ? 0x800000u : 0u);
    f = 0;
    fl = 0u;
    if (M == 0u) return;                              // +0 or -0
    if (E == 0xFF) { fl = kBig; return; }             // inf, NaN
    const int sh = (E ? E : 1) - 127 - e;             // q = M * 2^sh
    const bool neg = (b >> 31) != 0u;
    if (sh >= 0) {
        if (sh >= 24 || ((unsigned long long)M << sh) >= (1ull << 24)) { fl = kBig; return; }
        const int m = (int)(M << sh);
        f = neg ? -m : m;
        return;
    }
    const int k = min(-sh, 26);                       // k = 26: q < 1/4 (M < 2^24)
    const unsigned fa = M >> k, rem = M & ((1u << k) - 1u), half = 1u << (k - 1);
    if (!neg) {
        f = (int)fa;
        fl = (rem == half ? kTie : 0u) | (rem > half ? kUp : 0u);
    } else {
        f = -(int)fa - (rem ? 1 : 0);
        fl = rem == 0u ? 0u : rem == half ? kTie : rem < half ? kUp : 0u;
    }
}

// the increment of a classified term on a running sum of parity p
__device__ __forceinline__ unsigned inc_at(int f, unsigned fl, unsigned p)
{
    const unsigned uf = (unsigned)f;
    return uf + (((fl & kUp) || ((fl & kTie) && ((p + uf) & 1u))) ? 1u : 0u);
}

// ---------------------------------------------------------------------------
// Run functions.  For a run of terms and a candidate binade (a sign and an
// exponent field Es' = max(Es, 1), packed as key = sign << 8 | Es'), an NF
// holds, for an incoming sum of parity p = N & 1 (|s| = N * u): the
// increment t[p], the outgoing parity (bit p of q), and the bounds the walk
// must respect: the run stays in the binade iff N + lo[p] >= LO and
// N + hi[p] < 2^24, where lo is the least "sum before a term + floor(term)"
// (the exact sum's floor) and hi the largest rounded sum, both relative to
// the run's start.  key = -1: no such function (mixed candidates, or a term
// that always leaves); key = kZeroRun: every term is +-0, which leaves every
// sum as it is.  Values are kept within +-2^30: a run whose bounds
// reach that far can never stay, and saturation keeps it failing.
// ---------------------------------------------------------------------------
constexpr int kSat = 1 << 30;
constexpr int kZeroRun = -2;   // key of a run of zeros: it changes no sum, at any binade

struct alignas(16) NF {
    int key;
    unsigned q;
    int t[2], lo[2], hi[2];
};
static_assert(sizeof(NF) == 32, "NF layout");

__device__ __forceinline__ int sat(long long x)
{
    return (int)max(min(x, (long long)kSat), -(long long)kSat);
}

__device__ __forceinline__ NF nf_identity(int key)
{
    NF F;
    F.key = key;
    F.q = 2u;
    F.t[0] = F.t[1] = 0;
    F.lo[0] = F.lo[1] = kSat;
    F.hi[0] = F.hi[1] = -kSat;
    return F;
}

// first F, then G
__device__ __forceinline__ NF nf_compose(const NF &F, const NF &G)
{
    NF H;
    H.key = F.key == kZeroRun ? G.key : G.key == kZeroRun ? F.key : F.key == G.key ? F.key : -1;
    H.q = 0u;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int pm = (F.q >> p) & 1;
        H.t[p] = sat((long long)F.t[p] + G.t[pm]);
        H.lo[p] = min(F.lo[p], sat((long long)F.t[p] + G.lo[pm]));
        H.hi[p] = max(F.hi[p], sat((long long)F.t[p] + G.hi[pm]));
        H.q |= ((G.q >> pm) & 1u) << p;
    }
    return H;
}

// one classified term as a run function
__device__ __forceinline__ NF nf_leaf(int key, int f, unsigned fl)
{
    NF F;
    F.key = (fl & kBig) ? -1 : key;
    F.q = 0u;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int t = (int)inc_at(f, fl, (unsigned)p);
        F.t[p] = t;
        F.lo[p] = f;
        F.hi[p] = t;
        F.q |= (unsigned)((p + t) & 1) << p;
    }
    return F;
}

// key of a float sum: -1 for a zero or non-finite one
__device__ __forceinline__ int key_of(float s)
{
    const unsigned b = __float_as_uint(s);
    const int Es = (b >> 23) & 0xFF;
    if ((b & 0x7FFFFFFFu) == 0u || Es == 0xFF) return -1;
    return (int)((b >> 31) << 8) | max(Es, 1);
}

// ---------------------------------------------------------------------------
// One serial dot product (a job): terms a[h] (b == nullptr) or products
// a[h]*b[h] for h < npad (zero pads), summed in that order into *out.
// Runs: sub-runs of 64 terms, 16 of them a segment (1024), 8 segments a
// super-run (8192 terms).
// ---------------------------------------------------------------------------
constexpr int kSub = 64, kSegSubs = 16, kSupSegs = 8;
constexpr int kSeg = kSub * kSegSubs, kSup = kSeg * kSupSegs;

struct DotJob {
    const float *a, *b;
    double *dsub, *dpre;        // per sub-run: double sum; exclusive prefix of those
    NF *sub, *seg, *sup;        // run functions at the predicted binades
    float *out;
};

__device__ __forceinline__ float term_at(const DotJob &J, long long h)
{
    return J.b ? J.a[h] * J.b[h] : J.a[h];
}

// (A1) each sub-run's sum in double: only a predictor of the float sums
__global__ __launch_bounds__(256) void k_ds_sub(const DotJob *jobs, long long nsub)
{
    const DotJob J = jobs[blockIdx.y];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsub) return;
    double d = 0.0;
    for (int k = 0; k < kSub; ++k) d += (double)term_at(J, g * kSub + k);
    J.dsub[g] = d;
}

// (A2) exclusive prefix sums of the sub-run sums, one workgroup per job
__global__ __launch_bounds__(1024) void k_ds_scan(const DotJob *jobs, long long nsub)
{
    __shared__ double part[1024];
    const DotJob J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    const long long per = (nsub + 1023) / 1024, g0 = t * per, g1 = min(g0 + per, nsub);
    double acc = 0.0;
    for (long long g = g0; g < g1; ++g) acc += J.dsub[g];
    part[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {          // inclusive scan (Hillis-Steele)
        const double v = t >= o ? part[t - o] : 0.0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    double run = t ? part[t - 1] : 0.0;
    for (long long g = g0; g < g1; ++g) {
        J.dpre[g] = run;
        run += J.dsub[g];
    }
}

// (A3) each sub-run's function at the binade of its predicted incoming sum
__global__ __launch_bounds__(256) void k_nf_sub(const DotJob *jobs, long long nsub)
{
    const DotJob J = jobs[blockIdx.y];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nsub) return;
    bool zeros = true;
    for (int k = 0; k < kSub; ++k) zeros &= term_at(J, g * kSub + k) == 0.f;
    const int key = zeros ? kZeroRun : key_of((float)J.dpre[g]);
    NF F = nf_identity(key);
    if (key >= 0) {
        const unsigned sgn = (unsigned)(key >> 8) << 31;
        const int e = (key & 0xFF) - 127;
        for (int k = 0; k < kSub && F.key >= 0; ++k) {
            int f;
            unsigned fl;
            classify(term_at(J, g * kSub + k), sgn, e, f, fl);
            F = nf_compose(F, nf_leaf(key, f, fl));
        }
    }
    J.sub[g] = F;
}

// (A4, A5) a run's function from its `fan` children's (one key for all)
__global__ __launch_bounds__(256) void k_nf_up(const DotJob *jobs, int level, int fan, long long nparent)
{
    const DotJob J = jobs[blockIdx.y];
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nparent) return;
    const NF *child = level == 0 ? J.sub : J.seg;
    NF *parent = level == 0 ? J.seg : J.sup;
    NF F = child[g * fan];
    for (int k = 1; k < fan && F.key != -1; ++k) F = nf_compose(F, child[g * fan + k]);
    parent[g] = F;
}

// (B) the walk, one wave per job.  The wave carries the float sum through
// the super-runs in order (every lane holds the same state): a super-run
// whose function applies at the current sum is one step; otherwise its
// segments', then their sub-runs' functions are tried, and a sub-run that
// leaves the binade is added term by term as floats (the reference's own
// operation).  The segment and super-run functions of a window of super-runs
// are staged in LDS together; the sub-run functions and terms of a super-run
// are fetched only when the walk has to descend into it.
constexpr int kWalkThreads = 64;
constexpr int kWinSups = 128;                      // super-runs per LDS window

struct WalkLds {
    NF sup[kWinSups];
    NF seg[kWinSups * kSupSegs];
    NF sub[kSup / kSub];                           // the super-run being descended into
    alignas(16) float t[kSup];
};

// The walker's sum as its binade key and N (|s| = N * 2^(Es-150)) while it
// has one (key >= 0); the float itself is rebuilt only for serial runs.
struct WState {
    int key, N, LO;
    float s;
};

__device__ __forceinline__ void set_sum(WState &w, float s)
{
    w.s = s;
    w.key = key_of(s);
    const unsigned b = __float_as_uint(s);
    const int E = (b >> 23) & 0xFF;
    w.N = (int)((b & 0x7FFFFFu) | (E ? 0x800000u : 0u));
    w.LO = E <= 1 ? 1 : (1 << 23);
}

__device__ __forceinline__ float get_sum(const WState &w)
{
    if (w.key < 0) return w.s;
    const float m = __builtin_ldexpf((float)w.N, (w.key & 0xFF) - 150);
    return (w.key >> 8) ? -m : m;
}

// a whole function out of LDS in two 16-byte reads
__device__ __forceinline__ NF lds_nf(const NF &F)
{
    const vf4 a = ((const vf4 *)&F)[0], b = ((const vf4 *)&F)[1];
    NF G;
    G.key = __float_as_int(a.x);
    G.q = __float_as_uint(a.y);
    G.t[0] = __float_as_int(a.z);
    G.t[1] = __float_as_int(a.w);
    G.lo[0] = __float_as_int(b.x);
    G.lo[1] = __float_as_int(b.y);
    G.hi[0] = __float_as_int(b.z);
    G.hi[1] = __float_as_int(b.w);
    return G;
}

// the run F at the walker's sum if it stays in its binade
__device__ __forceinline__ bool jump(WState &w, const NF &F)
{
    if (F.key == kZeroRun) return true;
    if (w.key < 0 || F.key != w.key) return false;
    const int p = w.N & 1;
    if (w.N + F.lo[p] < w.LO || w.N + F.hi[p] >= (1 << 24)) return false;
    w.N += F.t[p];
    return true;
}

// n 16-byte words from global src to LDS dst by the wave's lanes
__device__ __forceinline__ void wave_copy(vf4 *dst, const vf4 *src, int n, int lane)
{
    for (int i = lane; i < n; i += kWalkThreads) dst[i] = src[i];
}

__global__ __launch_bounds__(kWalkThreads) void k_walk(const DotJob *jobs, long long nsup)
{
    __shared__ WalkLds L;
    const DotJob J = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    WState w;
    set_sum(w, 0.f);
    for (long long w0 = 0; w0 < nsup; w0 += kWinSups) {
        const int nw = (int)min((long long)kWinSups, nsup - w0);
        __syncthreads();
        wave_copy((vf4 *)L.sup, (const vf4 *)(J.sup + w0), nw * 2, lane);
        wave_copy((vf4 *)L.seg, (const vf4 *)(J.seg + w0 * kSupSegs), nw * kSupSegs * 2, lane);
        __syncthreads();
        for (int sp = 0; sp < nw; ++sp) {
            if (jump(w, lds_nf(L.sup[sp]))) continue;
            // descend: this super-run's sub-run functions and terms
            const long long h0 = (w0 + sp) * kSup;
            __syncthreads();
            wave_copy((vf4 *)L.sub, (const vf4 *)(J.sub + (w0 + sp) * (kSup / kSub)), (kSup / kSub) * 2, lane);
            const vf4 *a = (const vf4 *)(J.a + h0);
            const vf4 *b = J.b ? (const vf4 *)(J.b + h0) : nullptr;
            vf4 *t = (vf4 *)L.t;
            for (int i = lane; i < kSup / 4; i += kWalkThreads) {
                vf4 v = a[i];
                if (b) {
                    const vf4 u = b[i];
                    v.x = v.x * u.x; v.y = v.y * u.y; v.z = v.z * u.z; v.w = v.w * u.w;
                }
                t[i] = v;
            }
            __syncthreads();
            for (int g = 0; g < kSupSegs; ++g) {
                if (jump(w, lds_nf(L.seg[sp * kSupSegs + g]))) continue;
                for (int u = 0; u < kSegSubs; ++u) {
                    const int k = g * kSegSubs + u;
                    if (jump(w, lds_nf(L.sub[k]))) continue;
                    // the run's terms in registers first, then the float chain
                    const vf4 *t4 = (const vf4 *)(L.t + k * kSub);
                    vf4 v[kSub / 4];
#pragma unroll
                    for (int i = 0; i < kSub / 4; ++i) v[i] = t4[i];
                    float s = get_sum(w);
#pragma unroll
                    for (int i = 0; i < kSub / 4; ++i) {
                        s = s + v[i].x;
                        s = s + v[i].y;
                        s = s + v[i].z;
                        s = s + v[i].w;
                    }
                    set_sum(w, s);
                }
            }
        }
    }
    if (lane == 0) *J.out = get_sum(w);
}

__device__ __forceinline__ float rho_of(float ys)
{
    return fabs((double)ys) > 0.0 ? (float)(1.0 / (double)ys) : 0.0f;
}

// Scalars of the recursion, from the dot products: dys[k] = y_k.s_k,
// dyy = y_p.y_p, da[k] = s_k.q, db[k] = y_k.r (0-based slots).
struct Dots {
    const float *dys, *dyy, *da, *db;
    __device__ float alpha(int k) const { return rho_of(dys[k]) * da[k]; }   // LBFGS.c:116
    __device__ float beta(int k) const { return rho_of(dys[k]) * db[k]; }    // LBFGS.c:139
};

enum UpdMode { kAlphaFirst = 0, kAlpha = 1, kBetaFirst = 2, kBeta = 3, kFinal = 4 };

// One step of the two loops for slot k (0-based), elementwise over h < n:
//   kAlphaFirst  term = s_k q                       (k = K-1)
//   kAlpha       q -= alpha_{k+1} y_{k+1}; term = s_k q
//   kBetaFirst   q -= alpha_0 y_0; r = gamma q; term = y_0 r
//   kBeta        r += s_{k-1} (alpha_{k-1} - beta_{k-1}); term = y_k r
//   kFinal       r += s_k (alpha_k - beta_k)        (k = K-1)
__global__ __launch_bounds__(256) void k_lb_update(int mode, int k, int p, long long n, long long npad,
                                                   const float *__restrict__ y, const float *__restrict__ s,
                                                   float *__restrict__ q, float *__restrict__ r,
                                                   float *__restrict__ term, Dots D)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    float c = 0.f;
    const float *v = nullptr, *w = nullptr;
    switch (mode) {
    case kAlphaFirst: w = s + k * npad; break;
    case kAlpha: c = D.alpha(k + 1); v = y + (k + 1) * npad; w = s + k * npad; break;
    case kBetaFirst: c = D.alpha(0); v = y; w = y; break;
    case kBeta: c = D.alpha(k - 1) - D.beta(k - 1); v = s + (k - 1) * npad; w = y + k * npad; break;
    default: c = D.alpha(k) - D.beta(k); v = s + k * npad; break;
    }
    const float gamma = mode == kBetaFirst ? D.dys[p] / *D.dyy : 0.f;   // LBFGS.c:72
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < n; h += stride) {
        switch (mode) {
        case kAlphaFirst: term[h] = w[h] * q[h]; break;
        case kAlpha: {
            const float qn = q[h] - c * v[h];
            q[h] = qn;
            term[h] = w[h] * qn;
            break;
        }
        case kBetaFirst: {
            const float qn = q[h] - c * v[h];
            q[h] = qn;
            const float rn = gamma * qn;
            r[h] = rn;
            term[h] = w[h] * rn;
            break;
        }
        case kBeta: {
            const float rn = r[h] + v[h] * c;
            r[h] = rn;
            term[h] = w[h] * rn;
            break;
        }
        default: r[h] = r[h] + v[h] * c; break;
        }
    }
}

// y_p = g - gprev, s_p = v - vprev (LBFGS.c:45-58)
__global__ __launch_bounds__(256) void k_lb_pair(long long n, const float *__restrict__ g,
                                                 const float *__restrict__ gprev, const float *__restrict__ v,
                                                 const float *__restrict__ vprev, float *__restrict__ y,
                                                 float *__restrict__ s)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long h = (long long)blockIdx.x * blockDim.x + threadIdx.x; h < n; h += stride) {
        y[h] = g[h] - gprev[h];
        s[h] = v[h] - vprev[h];
    }
}

// Row-major NY x NX field <-> h order (h = x * NY + y), through 32 x 32 LDS
// tiles.  to_h: dst[x*ny + y] = src[y*nx + x]; else the inverse.
constexpr int kTile = 32;
__global__ __launch_bounds__(kTile * 8) void k_lb_transpose(const float *__restrict__ src, float *__restrict__ dst,
                                                            int nx, int ny, int to_h)
{
    __shared__ float tile[kTile][kTile + 1];
    // the source's rows are y (to_h) or x (from h)
    const int rows = to_h ? ny : nx, cols = to_h ? nx : ny;
    const int c0 = blockIdx.x * kTile, r0 = blockIdx.y * kTile;
    for (int rr = threadIdx.y; rr < kTile; rr += 8) {
        const int r = r0 + rr, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[rr][threadIdx.x] = src[(long long)r * cols + c];
    }
    __syncthreads();
    for (int cc = threadIdx.y; cc < kTile; cc += 8) {
        const int c = c0 + cc, r = r0 + threadIdx.x;
        if (r < rows && c < cols) dst[(long long)c * rows + r] = tile[threadIdx.x][cc];
    }
}

int grid_for(long long n)
{
    return (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 8192);
}

}  // namespace
}  // namespace rz

using namespace rz;

namespace rz {
namespace {

// Device scratch for njobs concurrent dot products over npad terms (npad a
// multiple of kSup); job descriptors live in a device array beside it.
struct DotScratch {
    long long nsub = 0, nseg = 0, nsup = 0;
    int njobs = 0, ndesc = 0;
    double *dsub = nullptr, *dpre = nullptr;
    NF *sub = nullptr, *seg = nullptr, *sup = nullptr;
    DotJob *desc = nullptr;
};

hipError_t dots_alloc(DotScratch &S, long long npad, int njobs, int ndesc)
{
    S.nsub = npad / kSub;
    S.nseg = npad / kSeg;
    S.nsup = npad / kSup;
    S.njobs = njobs;
    S.ndesc = ndesc;
    hipError_t e;
    if ((e = hipMalloc(&S.dsub, sizeof(double) * S.nsub * njobs)) != hipSuccess) return e;
    if ((e = hipMalloc(&S.dpre, sizeof(double) * S.nsub * njobs)) != hipSuccess) return e;
    if ((e = hipMalloc(&S.sub, sizeof(NF) * S.nsub * njobs)) != hipSuccess) return e;
    if ((e = hipMalloc(&S.seg, sizeof(NF) * S.nseg * njobs)) != hipSuccess) return e;
    if ((e = hipMalloc(&S.sup, sizeof(NF) * S.nsup * njobs)) != hipSuccess) return e;
    return hipMalloc(&S.desc, sizeof(DotJob) * ndesc);
}

void dots_free(DotScratch &S)
{
    void *p[] = {S.dsub, S.dpre, S.sub, S.seg, S.sup, S.desc};
    for (void *q : p)
        if (q) (void)hipFree(q);
    S = DotScratch();
}

// A descriptor for job slot j of the scratch: terms a (b null) or a*b, sum to out
DotJob dot_job(const DotScratch &S, int j, const float *a, const float *b, float *out)
{
    DotJob J;
    J.a = a;
    J.b = b;
    J.dsub = S.dsub + j * S.nsub;
    J.dpre = S.dpre + j * S.nsub;
    J.sub = S.sub + j * S.nsub;
    J.seg = S.seg + j * S.nseg;
    J.sup = S.sup + j * S.nsup;
    J.out = out;
    return J;
}

// the dot products of the n device descriptors at d (distinct scratch slots)
hipError_t dots_run(const DotScratch &S, const DotJob *d, int n, hipStream_t st)
{
    const dim3 gs((unsigned)((S.nsub + 255) / 256), n), gg((unsigned)((S.nseg + 255) / 256), n),
        gp((unsigned)((S.nsup + 255) / 256), n);
    k_ds_sub<<<gs, 256, 0, st>>>(d, S.nsub);
    k_ds_scan<<<n, 1024, 0, st>>>(d, S.nsub);
    k_nf_sub<<<gs, 256, 0, st>>>(d, S.nsub);
    k_nf_up<<<gg, 256, 0, st>>>(d, 0, kSegSubs, S.nseg);
    k_nf_up<<<gp, 256, 0, st>>>(d, 1, kSupSegs, S.nsup);
    k_walk<<<n, kWalkThreads, 0, st>>>(d, S.nsup);
    return hipGetLastError();
}

}  // namespace
}  // namespace rz

struct rz_lbfgs {
    int device = 0, nx = 0, ny = 0, K = 0;
    long long n = 0, npad = 0;
    hipStream_t stream = nullptr;
    float *y = nullptr, *s = nullptr;          // K slots of npad floats each
    float *q = nullptr, *r = nullptr, *term = nullptr;
    float *gh = nullptr, *vh = nullptr, *gprev = nullptr, *vprev = nullptr;
    float *in = nullptr;                       // a row-major field on its way in or out
    float *sc = nullptr;                       // dys[K], dyy, da[K], db[K]
    rz::DotScratch dots;                       // K + 1 jobs; descriptors: K + 1 slot dots, 2K passes
    int has_prev = 0;
    std::string err;
};

namespace {

int lb_fail(rz_lbfgs *lb, int code, const char *what, hipError_t e = hipSuccess)
{
    char buf[256];
    if (e != hipSuccess)
        snprintf(buf, sizeof buf, "lbfgs: %s: %s", what, hipGetErrorString(e));
    else
        snprintf(buf, sizeof buf, "lbfgs: %s", what);
    if (lb) lb->err = buf;
    else set_global_error(buf);
    return code;
}

#define LBCHK(lb, expr)                                                       \
    do {                                                                      \
        const hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) return lb_fail(lb, RZ_ERR_HIP, #expr, e_);      \
    } while (0)

float host_rho(float ys)
{
    return std::fabs((double)ys) > 0.0 ? (float)(1.0 / (double)ys) : 0.0f;
}

hipError_t lb_transpose(rz_lbfgs *lb, const float *src, float *dst, int to_h)
{
    const int rows = to_h ? lb->ny : lb->nx, cols = to_h ? lb->nx : lb->ny;
    dim3 grid((cols + kTile - 1) / kTile, (rows + kTile - 1) / kTile), block(kTile, 8);
    k_lb_transpose<<<grid, block, 0, lb->stream>>>(src, dst, lb->nx, lb->ny, to_h);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int rz_lbfgs_create(rz_lbfgs **out, int device, int nx, int ny, int nslots)
{
    if (!out) return lb_fail(nullptr, RZ_ERR_INVALID, "null handle pointer");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nslots < 1) return lb_fail(nullptr, RZ_ERR_INVALID, "NX, NY and NLBFGS must be >= 1");
    const long long n = (long long)nx * ny;
    if (n > (1ll << 31) / 2) return lb_fail(nullptr, RZ_ERR_INVALID, "grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return lb_fail(nullptr, RZ_ERR_INVALID, "no such device");
    rz_lbfgs *lb = new rz_lbfgs;
    lb->device = device; lb->nx = nx; lb->ny = ny; lb->K = nslots; lb->n = n;
    lb->npad = (n + kSup - 1) / kSup * kSup;
    const size_t vb = (size_t)lb->npad * sizeof(float);
    auto fail = [&](const char *what, hipError_t e) {
        const int rc = lb_fail(nullptr, e == hipErrorOutOfMemory ? RZ_ERR_NOMEM : RZ_ERR_HIP, what, e);
        rz_lbfgs_destroy(lb);
        return rc;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail("hipSetDevice", e);
    if ((e = hipStreamCreateWithFlags(&lb->stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
    float **bufs[] = {&lb->q, &lb->r, &lb->term, &lb->gh, &lb->vh, &lb->gprev, &lb->vprev, &lb->in};
    if ((e = hipMalloc(&lb->y, vb * nslots)) != hipSuccess || (e = hipMalloc(&lb->s, vb * nslots)) != hipSuccess)
        return fail("history allocation", e);
    for (float **b : bufs)
        if ((e = hipMalloc(b, vb)) != hipSuccess) return fail("vector allocation", e);
    if ((e = hipMalloc(&lb->sc, (3 * nslots + 1) * sizeof(float))) != hipSuccess) return fail("scalars", e);
    if ((e = dots_alloc(lb->dots, lb->npad, nslots + 1, 3 * nslots + 1)) != hipSuccess) return fail("dot scratch", e);
    // zero pads and an empty history (fatt.c:74-82: vector() starts at 0)
    if ((e = hipMemsetAsync(lb->y, 0, vb * nslots, lb->stream)) != hipSuccess ||
        (e = hipMemsetAsync(lb->s, 0, vb * nslots, lb->stream)) != hipSuccess)
        return fail("history clear", e);
    for (float **b : bufs)
        if ((e = hipMemsetAsync(*b, 0, vb, lb->stream)) != hipSuccess) return fail("vector clear", e);
    if ((e = hipStreamSynchronize(lb->stream)) != hipSuccess) return fail("clear", e);
    *out = lb;
    return RZ_OK;
}

void rz_lbfgs_destroy(rz_lbfgs *lb)
{
    if (!lb) return;
    (void)hipSetDevice(lb->device);
    if (lb->stream) (void)hipStreamSynchronize(lb->stream);
    float *bufs[] = {lb->y, lb->s, lb->q, lb->r, lb->term, lb->gh, lb->vh, lb->gprev, lb->vprev, lb->in, lb->sc};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    dots_free(lb->dots);
    if (lb->stream) (void)hipStreamDestroy(lb->stream);
    delete lb;
}

const char *rz_lbfgs_error(const rz_lbfgs *lb)
{
    return lb ? lb->err.c_str() : rz_last_error(nullptr);
}

int rz_lbfgs_zero(rz_lbfgs *lb)
{
    if (!lb) return lb_fail(nullptr, RZ_ERR_INVALID, "null handle");
    LBCHK(lb, hipSetDevice(lb->device));
    const size_t vb = (size_t)lb->npad * sizeof(float);
    LBCHK(lb, hipMemsetAsync(lb->y, 0, vb * lb->K, lb->stream));
    LBCHK(lb, hipMemsetAsync(lb->s, 0, vb * lb->K, lb->stream));
    LBCHK(lb, hipStreamSynchronize(lb->stream));
    return RZ_OK;
}

int rz_lbfgs_has_previous(const rz_lbfgs *lb)
{
    return lb ? lb->has_prev : 0;
}

int rz_lbfgs_set_previous(rz_lbfgs *lb, const float *grad_h, const float *vp_h)
{
    if (!lb || !grad_h || !vp_h) return lb_fail(lb, RZ_ERR_INVALID, "set_previous: null pointer");
    LBCHK(lb, hipSetDevice(lb->device));
    const size_t b = (size_t)lb->n * sizeof(float);
    LBCHK(lb, hipMemcpyAsync(lb->gprev, grad_h, b, hipMemcpyHostToDevice, lb->stream));
    LBCHK(lb, hipMemcpyAsync(lb->vprev, vp_h, b, hipMemcpyHostToDevice, lb->stream));
    LBCHK(lb, hipStreamSynchronize(lb->stream));
    lb->has_prev = 1;
    return RZ_OK;
}

int rz_lbfgs_step(rz_lbfgs *lb, int iter, int pointer, const float *grad, const float *gradm, const float *vp,
                  float *hgrad, float *grad_h, float *vp_h, float *hgrad_h, float *gamma, float *rho, float *alpha)
{
    if (!lb) return lb_fail(nullptr, RZ_ERR_INVALID, "null handle");
    if (!grad || !gradm || !vp || !hgrad) return lb_fail(lb, RZ_ERR_INVALID, "step: null field");
    if (iter < 1) return lb_fail(lb, RZ_ERR_INVALID, "step: iter must be >= 1");
    if (pointer < 1 || pointer > lb->K) return lb_fail(lb, RZ_ERR_INVALID, "step: LBFGS_pointer outside 1..NLBFGS");
    if (iter > 1 && !lb->has_prev) return lb_fail(lb, RZ_ERR_STATE, "step: no previous model and gradient");
    LBCHK(lb, hipSetDevice(lb->device));
    hipStream_t st = lb->stream;
    const long long n = lb->n, np = lb->npad;
    const size_t b = (size_t)n * sizeof(float);
    const int K = lb->K, p = pointer - 1, G = grid_for(n);
    // this iteration's gradient and model in h order (LBFGS.c:45-58,178-199)
    LBCHK(lb, hipMemcpyAsync(lb->in, grad, b, hipMemcpyHostToDevice, st));
    LBCHK(lb, lb_transpose(lb, lb->in, lb->gh, 1));
    LBCHK(lb, hipMemcpyAsync(lb->in, vp, b, hipMemcpyHostToDevice, st));   // after the transpose (stream order)
    LBCHK(lb, lb_transpose(lb, lb->in, lb->vh, 1));
    LBCHK(lb, hipMemcpyAsync(lb->in, gradm, b, hipMemcpyHostToDevice, st));
    LBCHK(lb, lb_transpose(lb, lb->in, lb->q, 1));                             // q = gradm (LBFGS.c:101-110)
    float *sc = lb->sc;
    const Dots D{sc, sc + K, sc + K + 1, sc + 2 * K + 1};
    if (iter > 1) {
        float *yp = lb->y + p * np, *sp = lb->s + p * np;
        k_lb_pair<<<G, 256, 0, st>>>(n, lb->gh, lb->gprev, lb->vh, lb->vprev, yp, sp);
        LBCHK(lb, hipGetLastError());
        // descriptors: y_k.s_k (k < K) and y_p.y_p side by side, then one per pass
        std::vector<DotJob> jobs;
        for (int k = 0; k < K; ++k) jobs.push_back(dot_job(lb->dots, k, lb->y + k * np, lb->s + k * np, sc + k));
        jobs.push_back(dot_job(lb->dots, K, yp, yp, sc + K));
        for (int k = K - 1; k >= 0; --k) jobs.push_back(dot_job(lb->dots, 0, lb->term, nullptr, sc + K + 1 + k));
        for (int k = 0; k < K; ++k) jobs.push_back(dot_job(lb->dots, 0, lb->term, nullptr, sc + 2 * K + 1 + k));
        const DotJob *dj = lb->dots.desc;
        LBCHK(lb, hipMemcpyAsync(lb->dots.desc, jobs.data(), jobs.size() * sizeof(DotJob), hipMemcpyHostToDevice, st));
        LBCHK(lb, hipStreamSynchronize(st));                                    // jobs is a local
        LBCHK(lb, dots_run(lb->dots, dj, K + 1, st));                           // LBFGS.c:68-96
        for (int k = K - 1, j = K + 1; k >= 0; --k, ++j) {                      // LBFGS.c:113-125
            k_lb_update<<<G, 256, 0, st>>>(k == K - 1 ? kAlphaFirst : kAlpha, k, p, n, np, lb->y, lb->s, lb->q,
                                           lb->r, lb->term, D);
            LBCHK(lb, dots_run(lb->dots, dj + j, 1, st));
        }
        for (int k = 0, j = 2 * K + 1; k < K; ++k, ++j) {                       // LBFGS.c:128-149
            k_lb_update<<<G, 256, 0, st>>>(k == 0 ? kBetaFirst : kBeta, k, p, n, np, lb->y, lb->s, lb->q, lb->r,
                                           lb->term, D);
            LBCHK(lb, dots_run(lb->dots, dj + j, 1, st));
        }
        k_lb_update<<<G, 256, 0, st>>>(kFinal, K - 1, p, n, np, lb->y, lb->s, lb->q, lb->r, lb->term, D);
        LBCHK(lb, hipGetLastError());
        LBCHK(lb, lb_transpose(lb, lb->r, lb->in, 0));                         // LBFGS.c:151-161
        LBCHK(lb, hipMemcpyAsync(hgrad, lb->in, b, hipMemcpyDeviceToHost, st));
        if (hgrad_h) LBCHK(lb, hipMemcpyAsync(hgrad_h, lb->r, b, hipMemcpyDeviceToHost, st));
    } else {
        // iter == 1: the direction is gradm itself (LBFGS.c:165-174)
        std::memcpy(hgrad, gradm, b);
        if (hgrad_h) LBCHK(lb, hipMemcpyAsync(hgrad_h, lb->q, b, hipMemcpyDeviceToHost, st));
    }
    if (grad_h) LBCHK(lb, hipMemcpyAsync(grad_h, lb->gh, b, hipMemcpyDeviceToHost, st));
    if (vp_h) LBCHK(lb, hipMemcpyAsync(vp_h, lb->vh, b, hipMemcpyDeviceToHost, st));
    std::vector<float> hs(3 * K + 1);
    if (iter > 1) LBCHK(lb, hipMemcpyAsync(hs.data(), sc, hs.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    LBCHK(lb, hipStreamSynchronize(st));
    if (iter > 1) {
        // the reference's scalars from the same dot products (LBFGS.c:72,83-88,116)
        if (gamma) *gamma = hs[p] / hs[K];
        for (int k = 0; k < K; ++k) {
            if (rho) rho[k] = host_rho(hs[k]);
            if (alpha) alpha[k] = host_rho(hs[k]) * hs[K + 1 + k];
        }
    }
    std::swap(lb->gh, lb->gprev);
    std::swap(lb->vh, lb->vprev);
    lb->has_prev = 1;
    return RZ_OK;
}

int rz_dotp_ordered(int device, const float *a, const float *b, long long n, int on_device, float *out)
{
    if (!a || !b || !out || n < 0) return lb_fail(nullptr, RZ_ERR_INVALID, "dotp: bad argument");
    if (n > (1ll << 40)) return lb_fail(nullptr, RZ_ERR_INVALID, "dotp: too many terms");
    if (hipSetDevice(device) != hipSuccess) return lb_fail(nullptr, RZ_ERR_INVALID, "dotp: no such device");
    const long long npad = std::max<long long>((n + kSup - 1) / kSup, 1) * kSup;
    const size_t nb = (size_t)n * sizeof(float), pb = (size_t)npad * sizeof(float);
    float *da = nullptr, *db = nullptr, *dout = nullptr;
    DotScratch S;
    auto done = [&](int rc) {
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dout);
        dots_free(S);
        return rc;
    };
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    hipError_t e;
    if ((e = hipMalloc(&da, pb)) != hipSuccess || (e = hipMalloc(&db, pb)) != hipSuccess ||
        (e = hipMalloc(&dout, sizeof(float))) != hipSuccess || (e = dots_alloc(S, npad, 1, 1)) != hipSuccess)
        return done(lb_fail(nullptr, RZ_ERR_NOMEM, "dotp: allocation", e));
    // zero-padded copies: +0 terms change no float sum
    if ((e = hipMemset(da, 0, pb)) != hipSuccess || (e = hipMemset(db, 0, pb)) != hipSuccess ||
        (n && ((e = hipMemcpy(da, a, nb, kind)) != hipSuccess || (e = hipMemcpy(db, b, nb, kind)) != hipSuccess)))
        return done(lb_fail(nullptr, RZ_ERR_HIP, "dotp: upload", e));
    const DotJob J = dot_job(S, 0, da, db, dout);
    if ((e = hipMemcpy(S.desc, &J, sizeof J, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = dots_run(S, S.desc, 1, nullptr)) != hipSuccess ||
        (e = hipMemcpy(out, dout, sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
        return done(lb_fail(nullptr, RZ_ERR_HIP, "dotp", e));
    return done(RZ_OK);
}

}  // extern "C"
