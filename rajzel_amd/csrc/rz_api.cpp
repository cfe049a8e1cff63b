// C ABI of include/rajzel_amd.h: context, device buffers, batching and the
// per-evaluation launch sequence.  Kernels live in rz_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <string>
#include <mutex>
#include <unordered_map>
#include <chrono>
#include <vector>

#include "../../include/rajzel_amd.h"
#include "rz_internal.h"

using namespace rz;

namespace {

struct ModelHost {
    rz_model_desc d{};
    int nb = 0;
    long long bstride = 0;
    std::vector<int> src_k, src_c, rec_k, rec_c, uq_k, uq_c, uq_r;
    long long rec_off = 0;   // offset of this model's (nshots, ntr) blocks
    int shot_base = 0;       // global index of shot 0
    long long grad_off = 0;  // offset in the gradient buffer
    int carry_off = 0;       // offset in the carry buffer
    bool have_S = false, have_picks = false, have_res = false;
    bool fastdiv = false;
    // stored transposed (NY above the rows a workgroup holds, NX within
    // them): d.nx and d.ny, the sources' and receivers' rows (k) and columns
    // (c) are those of the transposed grid
    bool trans = false;
    bool ft_dh = false;     // the float FMA quotient is exact for this DH (ft_div)
    bool ft_s = false;      // ... and the slowness keeps its inputs clear of underflow
    float rdh_hi = 0.f, rdh_lo = 0.f, s_min_ft = 0.f;
    double t_min_ft = 0.0;  // smallest nonzero T the float quotient allows
    double stop_s = 0.0, band_lo = 0.0, band_hi = 0.0;   // stopping test (set_stop)
    int stop_mode = 0;      // tests of the slower decisions (RAJZEL_AMD_ORDERED_NORM)
    // device
    float *S = nullptr, *SA = nullptr, *SB = nullptr;
    int *rec_k_d = nullptr, *rec_c_d = nullptr, *uq_k_d = nullptr, *uq_c_d = nullptr, *uq_r_d = nullptr;
    // per-shot gradient taper (rz_set_shot_taper): settings, and the device
    // tables built from them
    int tap_mode = 0, tap_filt = 0, tap_first = 0, tap_nsrc = 0;
    float tap_radius = 0.f;
    std::vector<float> tap_sx, tap_sz;
    int *tap_box = nullptr;
    long long *tap_off = nullptr;
    float *tap_fsrc = nullptr, *tap_frec = nullptr;
    unsigned *tap_start = nullptr;
};

}  // namespace

struct rz_ctx {
    int device = 0;
    hipStream_t own = nullptr, stream = nullptr;
    int R = 4, nb_max = 1;
    bool paired = true;    // two sweeps at a time; false: tall or transposed grids (R = 8)
    bool fastdiv = true;
    bool chkinf = false;   // INF guard of sweep_adj.c:68 can fire only if 2*TIME >= INF
    bool ft = false;       // every model allows the float T-only quotients (adjoint variant 2)
    bool ft_allowed = true;   // RAJZEL_AMD_FLOAT_T=0 turns them off (tests compare both)
    std::vector<ModelHost> m;
    int total_shots = 0;
    long long total_rec = 0, total_cells = 0;
    int total_tr = 0;
    // device buffers
    ModelDev *md_d = nullptr;
    ShotDev *shots_d = nullptr;
    int *slot_of_d = nullptr;
    long long *rec_off_d = nullptr, *rec_off0_d = nullptr;
    int *shot_base_d = nullptr, *shot_base0_d = nullptr;
    float *tobs_d = nullptr, *tmod_d = nullptr, *tnb_d = nullptr, *tres_d = nullptr, *carry_d = nullptr;
    float *dres_d = nullptr;            // residuals given by rz_set_residuals
    float *carry0_d = nullptr;          // initial residual carry (rz_set_tres_carry)
    // whole evaluations recorded as HIP graphs (rz_set_graph): [0] forward, [1] gradient
    bool graph = false;
    // split shots (rz_set_split): two workgroups per shot while a launch has
    // at most `ncu` shots; off for good once a launch reports that it cannot
    bool split = true;
    int ncu = 0;
    int split_fail = 0;                 // 1: workgroups of a shot on different XCDs, 2: a wait timed out
    int split_shots = -1;               // split shots per mixed launch: -1 all the CUs allow, 0 none
    // after a failed split launch: runs until split is tried again, and the
    // wait after the next failure (doubled each time: 16, 32, 64, ...)
    long long split_wait = 0, split_backoff = 16;
    SplitSync ss{nullptr, nullptr, nullptr};
    // Mixed launches: with more shots than half the CUs (and fewer than
    // the CUs), the shots that took the most outer iterations in the
    // previous evaluation run split on a side stream, the rest beside them
    // with one workgroup each, so that every workgroup has a CU.
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<int> hist_f, hist_a;    // outer iterations per shot in the last run (-1 unknown)
    std::vector<int> hist_pf, hist_pa;  // stopping-test passes per shot in the last run (ord_f/ord_a words)
    std::vector<int> hist_wf;           // sweep phases of each shot's forward in the last run
    ShotDev *plan_d = nullptr;          // [4][total_shots]: forward split / rest, adjoint split / rest
    int plan_n[4] = {0, 0, 0, 0};
    std::vector<int> plan_ids;          // shot order of the plan now on the device
    bool plan_on = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    long long s_epoch = 0;              // bumped by every rz_set_slowness
    long long fields_epoch = -1;        // slowness epoch of resident forward fields of all shots
    float *grad_d = nullptr;
    double *l2_shot_d = nullptr, *l2_sum_d = nullptr;
    int *it_f_d = nullptr, *it_a_d = nullptr;
    int *lsel_d = nullptr;              // per shot: the converged lambda is in LC (1) or LA (0)
    int *ord_f_d = nullptr, *ord_a_d = nullptr;   // per shot: passes of the stopping test over the fields
    float *osum_f_d = nullptr, *osum_a_d = nullptr; // per shot: ordered sum of the last decision
    int *phases_f_d = nullptr;          // per shot: sweep phases of the forward
    int *err_d = nullptr;
    // single-shot scratch
    float *s_tmod = nullptr, *s_tnb = nullptr, *s_tres = nullptr, *s_row = nullptr, *s_row2 = nullptr;
    int *s_it = nullptr;                // iterations forward, adjoint; lambda location; stop passes f, a;
                                        // forward phases
    float *s_osum = nullptr;            // ordered sums of the last decisions, forward and adjoint
    ShotDev *s_shot = nullptr;
    int *s_slot0 = nullptr;
    // field slots
    int nslots = 0;
    long long slot_floats = 0;
    float *TA = nullptr, *TB = nullptr, *LA = nullptr, *LB = nullptr, *LC = nullptr;
    // host mirror of the shot table
    std::vector<ShotDev> shots_h;
    int last_b0 = -1, last_b1 = -1;   // global shot range resident in the slots
    bool have_grad = false, have_fwd = false;
    size_t bytes = 0;
    std::vector<void *> allocs;
    std::vector<size_t> alloc_bytes;
    std::string err;
    // phase timing with events on the work stream (rz_set_profiling)
    bool prof = false;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double prof_ms[4] = {0, 0, 0, 0};   // forward, residual, adjoint, gradient
    long long prof_launches[4] = {0, 0, 0, 0};
};

namespace {

int fail(rz_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), RZ_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                 \
    } while (0)

template <typename T>
int dalloc(rz_ctx *c, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, RZ_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    }
    c->allocs.push_back(q);
    c->alloc_bytes.push_back(n * sizeof(T));
    c->bytes += n * sizeof(T);
    *p = static_cast<T *>(q);
    return RZ_OK;
}

// Free one allocation of the context (the stream must be idle).
void dfree(rz_ctx *c, void *p)
{
    if (!p) return;
    for (size_t i = 0; i < c->allocs.size(); ++i)
        if (c->allocs[i] == p) {
            c->allocs.erase(c->allocs.begin() + i);
            c->bytes -= c->alloc_bytes[i];
            c->alloc_bytes.erase(c->alloc_bytes.begin() + i);
            (void)hipFree(p);
            return;
        }
}

// Factors the per-shot taper tables may hold (1 GiB of floats).
constexpr size_t kTaperMaxEntries = (size_t)1 << 28;

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != RZ_OK) return rc_; \
    } while (0)

int bands_for(int ny, int R) { return (ny + kWave * R - 1) / (kWave * R); }

// Rows a workgroup holds (8 bands of 512 rows, R = 8).  A taller grid whose
// NX fits them is stored and solved transposed: the Godunov and adjoint
// updates treat x and y alike, so with the sweep directions transposed too
// (and norm1's sum still taken in the grid's order) every value is the same.
constexpr int kMaxRowsStored = kWave * 8 * max_bands_unpaired(8);
bool transposed(const rz_model_desc &d) { return d.ny > kMaxRowsStored && d.nx <= kMaxRowsStored; }

// NX and NY of the grid as the caller sees it
int grid_nx(const ModelHost &h) { return h.trans ? h.d.ny : h.d.nx; }
int grid_ny(const ModelHost &h) { return h.trans ? h.d.nx : h.d.ny; }

// The adjoint divides by DH as (float)(x * RN64(1/DH)).  That equals the IEEE
// quotient x/DH for every float x (a float quotient is never within 2^-49
// relative of a rounding tie; the double product errs by < 2^-52).  Confirm
// on 2^20 inputs spread over all exponents, including subnormal ones, before
// the kernels rely on it; otherwise they divide.
bool fast_div_exact(float dh)
{
    const double r = 1.0 / (double)dh;
    uint32_t state = 0x9E3779B9u;
    for (int i = 0; i < (1 << 20); ++i) {
        state ^= state << 13; state ^= state >> 17; state ^= state << 5;
        uint32_t bits = state & 0x7F7FFFFFu;   // finite, both signs via the loop below
        float x;
        std::memcpy(&x, &bits, 4);
        for (int sg = 0; sg < 2; ++sg, x = -x) {
            const volatile float q = x / dh;
            const float f = (float)((double)x * r);
            if (std::memcmp((const void *)&q, &f, 4) != 0) return false;
        }
    }
    return true;
}

// The adjoint's T-only quotients (slopes and the terms of lhs, sweep_adj.c:
// 33-54) as fma(x, rh, x * rl) in float, rh = RN(1/DH), rl = RN(1/DH - rh):
// two instructions instead of three.  Where x/DH keeps away from float
// rounding ties (a DH with a short significand: 5 = 101b puts every
// quotient at least 2^-27 relative from a tie) and the sum rh + rl errs by
// ~2^-48, this is RN(x/DH) -- provided x * rl is a normal float (|x| >=
// xmin).  Confirmed for every significand of one binade; other binades
// follow by scaling with powers of 2 while nothing underflows or overflows.
struct FtDiv {
    bool ok;
    float rh, rl;
    double xmin;   // smallest |x| != 0 for which the identity holds
};

FtDiv ft_div(float dh)
{
    static std::mutex mu;
    static std::unordered_map<uint32_t, FtDiv> cache;
    uint32_t key;
    std::memcpy(&key, &dh, 4);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    FtDiv r{false, 0.f, 0.f, 0.0};
    const volatile float one = 1.f;
    r.rh = one / dh;
    r.rl = (float)(1.0 / (double)dh - (double)r.rh);
    r.xmin = r.rl != 0.f ? std::ldexp(1.0, -126) / std::fabs((double)r.rl) : 0.0;
    bool ok = std::isfinite(r.rh) && r.rh > 0.f && std::isfinite(r.rl) && r.xmin < std::ldexp(1.0, 60);
    for (uint32_t i = 0; ok && i < (1u << 23); ++i) {
        const uint32_t bits = 0x3F800000u | i;   // x in [1, 2)
        float x;
        std::memcpy(&x, &bits, 4);
        const volatile float q = x / dh;
        const float t = x * r.rl;
        const float f = std::fma(x, r.rh, t);
        ok = std::memcmp((const void *)&q, &f, 4) == 0;
    }
    r.ok = ok;
    cache[key] = r;
    return r;
}

// The float T-only quotients of a model: its DH passes ft_div and every
// input of those quotients is 0 or at least xmin.  The inputs are
// differences of traveltimes and quotients of such differences by DH; a
// nonzero difference of two values >= t_min is at least t_min * 2^-24.  The
// forward's traveltimes are 0 (source), the initial 2*TIME, or at least
// S*DH/2 (a 1-D update adds S*DH, a 2-D one half of a root above S*DH), so
// the host requires 2*TIME >= t_min and the slowness >= t_min/(0.48*DH)
// (checked with every slowness upload); a caller's T (rz_eikonal_adj_shot)
// is checked directly.
void set_ft(ModelHost &h)
{
    const FtDiv d = ft_div(h.d.dh);
    h.rdh_hi = d.rh;
    h.rdh_lo = d.rl;
    const double dh = h.d.dh;
    h.t_min_ft = d.xmin * std::ldexp(1.0, 25) * std::max(1.0, dh);
    const double smin = h.t_min_ft / (0.48 * dh);
    h.s_min_ft = (float)std::min(smin, 1e30);
    const double tmax = std::max(2.0 * (double)h.d.time, (double)kInf);
    h.ft_dh = d.ok && smin < 1e30 && 2.0 * (double)h.d.time >= h.t_min_ft && tmax / dh < std::ldexp(1.0, 120);
}

// The stopping test of eikonal.c:44 / eikonal_adj.c:37 continues iff
// (float)sum / (float)(NX*NY) >= TTNORM, sum being norm1's float sum of
// |T - Told| (norm1.c:21-27).  A correctly rounded float division is
// monotone in the dividend, so this holds iff sum >= stop_s, the least
// non-negative float with that property (0 if every sum continues, +inf if
// none does).  The kernels' double sum G of per-sweep decreases is within
// 16u of the exact sum E (u = 2^-24: one float difference per cell and a
// float sum over at most 8 rows per lane; the double sums add ~1e-12), and
// norm1's float sum of NX*NY non-negative terms is within (1+u)(1+gamma) - 1
// of E, gamma = m u / (1 - m u) with m = NX*NY.  band_lo and band_hi turn G
// into bounds on norm1's sum.  Where they straddle stop_s the kernels count
// the cells that changed (norm1's nonzero terms) and bound again with that
// count, and only where that cannot decide either take the sum in the
// reference's order.  With NX*NY*u >= 1/2 the first band is useless and
// every nonzero decision goes to the count.
void set_stop(ModelHost &h)
{
    const float denom = (float)(h.d.nx * h.d.ny);
    const float ttn = h.d.ttnorm;
    auto cont = [&](uint32_t bits) {
        float s;
        std::memcpy(&s, &bits, 4);
        const volatile float q = s / denom;
        return q >= ttn;
    };
    const uint32_t top = 0x7F7FFFFFu;   // FLT_MAX
    if (cont(0)) {
        h.stop_s = 0.0;
    } else if (!cont(top)) {
        h.stop_s = INFINITY;
    } else {
        uint32_t lo = 0, hi = top;      // cont(lo) false, cont(hi) true; bit order = value order here
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            (cont(mid) ? hi : lo) = mid;
        }
        float s;
        std::memcpy(&s, &hi, 4);
        h.stop_s = s;
    }
    const double u = std::ldexp(1.0, -24), mu = (double)h.d.nx * (double)h.d.ny * u;
    // tests of the slower decisions: 1 = every nonzero decision takes the
    // ordered sum, 2 = every one counts the changed cells first
    const char *env = getenv("RAJZEL_AMD_ORDERED_NORM");
    h.stop_mode = (env && env[0] == '1') ? 1 : (env && env[0] == '2') ? 2 : 0;
    if (mu >= 0.5 || h.stop_mode != 0) {
        h.band_lo = 0.0;
        h.band_hi = INFINITY;
        return;
    }
    const double gam = mu / (1.0 - mu), er = (1.0 + u) * (1.0 + gam) - 1.0, eg = 16.0 * u + 1e-9;
    h.band_lo = (1.0 - er) / (1.0 + eg) * (1.0 - 1e-9);
    h.band_hi = (1.0 + er) / (1.0 - eg) * (1.0 + 1e-9);
}

int choose_R(int ny_max)
{
    for (int R : {1, 2, 4, 8})
        if (bands_for(ny_max, R) <= 4) return R;
    return 8;
}

// The taper of model h for a launch whose first shot uses source src0.
TaperDev taper_of(const ModelHost &h, int src0)
{
    TaperDev t{};
    t.mode = h.tap_mode;
    t.src0 = src0;
    t.box = h.tap_box; t.off = h.tap_off; t.fsrc = h.tap_fsrc;
    t.start = h.tap_start; t.frec = h.tap_frec;
    return t;
}

ModelDev make_dev(const rz_ctx *c, const ModelHost &h)
{
    ModelDev d{};
    d.nx = h.d.nx; d.ny = h.d.ny; d.nb = h.nb; d.R = c->R;
    d.dh = h.d.dh; d.time = h.d.time; d.ttnorm = h.d.ttnorm; d.eps_adj = h.d.eps_adj;
    d.two_dh2 = 2.0 * ((double)h.d.dh * (double)h.d.dh);
    d.rdh = 1.0 / (double)h.d.dh;
    d.bstride = h.bstride;
    d.fstride = (long long)h.nb * h.bstride;
    d.SA = h.SA; d.SB = h.SB; d.S = h.S;
    d.ntr = h.d.ntr; d.nuq = (int)h.uq_k.size();
    d.rec_k = h.rec_k_d; d.rec_c = h.rec_c_d;
    d.uq_k = h.uq_k_d; d.uq_c = h.uq_c_d; d.uq_r = h.uq_r_d;
    d.max_iter = h.d.max_iter > 0 ? h.d.max_iter : 1000;
    d.trans = h.trans ? 1 : 0;
    d.rdh_hi = h.rdh_hi; d.rdh_lo = h.rdh_lo; d.s_min_ft = h.s_min_ft;
    d.stop_s = h.stop_s; d.band_lo = h.band_lo; d.band_hi = h.band_hi; d.stop_mode = h.stop_mode;
    return d;
}

int check_model(rz_ctx *c, int model)
{
    if (!c) return RZ_ERR_INVALID;
    if (model < 0 || model >= (int)c->m.size()) return fail(c, RZ_ERR_INVALID, "model %d out of range", model);
    HIPCHK(c, hipSetDevice(c->device));
    return RZ_OK;
}

int adj_variant(const rz_ctx *c)
{
    return (c->fastdiv ? kAdjFastDiv : 0) | (c->chkinf ? kAdjInfGuard : 0) |
           (c->fastdiv && c->ft ? kAdjFloatT : 0);
}

Slots slots_of(const rz_ctx *c)
{
    Slots s;
    s.TA = c->TA; s.TB = c->TB; s.LA = c->LA; s.LB = c->LB; s.LC = c->LC;
    s.slot_stride = c->slot_floats;
    return s;
}

ShotOut outs_of(const rz_ctx *c)
{
    ShotOut o;
    o.tmod = c->tmod_d; o.tnb = c->tnb_d; o.it_f = c->it_f_d; o.it_a = c->it_a_d; o.lsel = c->lsel_d;
    o.ord_f = c->ord_f_d; o.ord_a = c->ord_a_d; o.osum_f = c->osum_f_d; o.osum_a = c->osum_a_d;
    o.phases_f = c->phases_f_d;
    return o;
}

int ready(rz_ctx *c, bool need_picks)
{
    for (size_t i = 0; i < c->m.size(); ++i) {
        if (!c->m[i].have_S) return fail(c, RZ_ERR_STATE, "model %zu: slowness not set", i);
        if (need_picks && !c->m[i].have_picks && c->m[i].d.nshots > 0 && c->m[i].d.ntr > 0)
            return fail(c, RZ_ERR_STATE, "model %zu: picks not set", i);
    }
    return RZ_OK;
}

// A launch with split shots reported that it could not run split.
constexpr int kSplitRetry = -100;

int check_convergence(rz_ctx *c, bool adj)
{
    std::vector<int> itf(c->total_shots), ita(c->total_shots);
    if (c->total_shots == 0) return RZ_OK;
    int serr = 0;
    HIPCHK(c, hipMemcpyAsync(&serr, c->ss.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    std::vector<int> pf(c->total_shots), pa(c->total_shots), wf(c->total_shots);
    HIPCHK(c, hipMemcpyAsync(itf.data(), c->it_f_d, itf.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(wf.data(), c->phases_f_d, wf.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(pf.data(), c->ord_f_d, pf.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (adj) {
        HIPCHK(c, hipMemcpyAsync(ita.data(), c->it_a_d, ita.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(pa.data(), c->ord_a_d, pa.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (serr != 0) {
        c->split = false;
        c->split_fail = serr;
        c->split_wait = c->split_backoff;
        c->split_backoff = std::min<long long>(2 * c->split_backoff, 1 << 20);
        HIPCHK(c, hipMemsetAsync(c->ss.err, 0, sizeof(int), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return kSplitRetry;
    }
    c->hist_f = itf;
    c->hist_pf = pf;
    c->hist_wf = wf;
    if (adj) {
        c->hist_a = ita;
        c->hist_pa = pa;
    } else if ((int)c->hist_a.size() != c->total_shots) {
        c->hist_a.assign(c->total_shots, -1);
        c->hist_pa.assign(c->total_shots, 0);
    }
    for (int i = 0; i < c->total_shots; ++i)
        if (itf[i] < 0 || (adj && ita[i] < 0))
            return fail(c, RZ_ERR_NOT_CONVERGED, "shot %d did not converge within max_iter", i);
    return RZ_OK;
}

// Record an event on the work stream when profiling (index into c->ev).
int mark(rz_ctx *c, int *idx)
{
    *idx = -1;
    if (!c->prof) return RZ_OK;
    if (c->ev_used == (int)c->ev.size()) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->ev.push_back(e);
    }
    *idx = c->ev_used++;
    HIPCHK(c, hipEventRecord(c->ev[*idx], c->stream));
    return RZ_OK;
}

struct PhaseMarks {
    int e[5];
};

// Add the elapsed times of recorded batches to the phase totals.
int collect(rz_ctx *c, const std::vector<PhaseMarks> &marks, bool grad)
{
    if (!c->prof) return RZ_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (const PhaseMarks &pm : marks) {
        const int nph = grad ? 4 : 1;
        for (int p = 0; p < nph; ++p) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev[pm.e[p]], c->ev[pm.e[p + 1]]));
            c->prof_ms[p] += ms;
            c->prof_launches[p] += 1;
        }
    }
    c->ev_used = 0;
    return RZ_OK;
}

void drop_graphs(rz_ctx *c);

// The side stream follows the work stream from here on / the work stream
// waits for everything the side stream has so far.
int fork_side(rz_ctx *c)
{
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    return RZ_OK;
}

int join_side(rz_ctx *c)
{
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return RZ_OK;
}

// Decide the mixed launches of the next single-batch run from the last
// run's iteration counts (stable order: more iterations first, then shot
// order) and put the shot lists on the device when they change.
int plan_mixed(rz_ctx *c, int per)
{
    const int T = c->total_shots;
    c->plan_on = false;
    if (!c->split || c->split_shots == 0 || per < T || T <= 0 || 2 * T <= c->ncu || T >= c->ncu || !c->paired)
        return RZ_OK;
    if ((int)c->hist_f.size() != T || (int)c->hist_a.size() != T) return RZ_OK;
    for (int i = 0; i < T; ++i)
        if (c->hist_f[i] < 0 || c->hist_a[i] < 0) return RZ_OK;
    const int ns = c->split_shots < 0 ? c->ncu - T : std::min(c->split_shots, c->ncu - T);
    std::vector<int> ids;
    ids.reserve(4 * (size_t)T);
    for (int ph = 0; ph < 2; ++ph) {
        const std::vector<int> &h = ph ? c->hist_a : c->hist_f, &hp = ph ? c->hist_pa : c->hist_pf;
        // a shot's expected time in sixteenths of an outer iteration: an
        // ordered-sum pass of the stopping test costs about half an
        // iteration, a count pass about a sixteenth
        // (the forward's sweeps after a sweep that changed no cell are
        // skipped: its time is its sweep phases, two per iteration)
        std::vector<int> cost(T);
        for (int i = 0; i < T; ++i) {
            const int w = (int)hp.size() == T ? hp[i] : 0;
            const int run = (ph == 0 && (int)c->hist_wf.size() == T) ? 8 * c->hist_wf[i] : 16 * h[i];
            cost[i] = run + 8 * (w >> 16) + (w & 0xFFFF);
        }
        std::vector<int> ord(T);
        for (int i = 0; i < T; ++i) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        // Workgroup w of a launch runs on XCD w mod 8 (a split shot's two at
        // list position j run on XCD j mod 8).  The slowest shot of the phase
        // goes first, to XCD 0, and the one-workgroup list puts its fastest
        // shots on the positions that land on XCD 0: they finish early, and
        // the slowest shot shares that XCD's L2 with less work for the rest
        // of its run (measured 0.5-1 ms per evaluation at 200 shots).  The
        // order of shots in a launch does not change any result.
        for (int i = 0; i < ns; ++i) ids.push_back(ord[i]);            // split, slowest first
        const int nr = T - ns;
        std::vector<int> pos(nr, -1);
        int fast = T - 1;                                              // ord[ns, T): slowest to fastest
        for (int p = 0; p < nr; p += 8) pos[p] = ord[fast--];
        std::vector<int> others(ord.begin() + ns, ord.begin() + fast + 1);
        std::sort(others.begin(), others.end());                       // the others in shot order
        size_t o = 0;
        for (int p = 0; p < nr; ++p) ids.push_back(pos[p] >= 0 ? pos[p] : others[o++]);
    }
    if (!c->side) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    if (!c->plan_d) TRY(dalloc(c, &c->plan_d, 4 * (size_t)T));
    if (ids != c->plan_ids) {
        std::vector<ShotDev> lists(4 * (size_t)T);
        for (int q = 0; q < 2; ++q)
            for (int i = 0; i < T; ++i) {
                const int shot = ids[(size_t)q * T + i];
                // forward lists at [0, T) and [T, 2T), adjoint at [2T, 3T) and [3T, 4T)
                const size_t at = i < ns ? (size_t)(2 * q) * T + i : (size_t)(2 * q + 1) * T + (i - ns);
                lists[at] = c->shots_h[shot];
            }
        HIPCHK(c, hipMemcpy(c->plan_d, lists.data(), lists.size() * sizeof(ShotDev), hipMemcpyHostToDevice));
        c->plan_ids = ids;
        drop_graphs(c);
    }
    c->plan_n[0] = ns; c->plan_n[1] = T - ns; c->plan_n[2] = ns; c->plan_n[3] = T - ns;
    c->plan_on = true;
    return RZ_OK;
}

// Split shots (two workgroups each) while every workgroup can have a CU of
// its own: at most one shot per two CUs.  With more shots the workgroups
// would share CUs (two fit one), which measured slower than one workgroup
// per shot.  Unpaired grids (more bands than a pair holds, or transposed)
// have no pair to split.
bool use_split(const rz_ctx *c, int nshots)
{
    return c->split && c->split_shots != 0 && nshots > 0 && 2 * nshots <= c->ncu && c->paired;
}

// Enqueue the launches of one batch of shots [b0, b1) on the work stream:
// forward (unless its fields are reused), and with grad the residuals, the
// adjoint and the gradient.  pm (optional) receives the phase events.
int enqueue_batch(rz_ctx *c, bool grad, bool direct, bool reuse, int b0, int b1, PhaseMarks *pm)
{
    const Slots sl = slots_of(c);
    const ShotOut so = outs_of(c);
    PhaseMarks dummy;
    PhaseMarks &m = pm ? *pm : dummy;
    TRY(mark(c, &m.e[0]));
    const bool split = use_split(c, b1 - b0);
    const bool mixed = !split && c->plan_on && b0 == 0 && b1 == c->total_shots;
    if (!reuse && mixed) {
        const long long T = c->total_shots;
        TRY(fork_side(c));
        HIPCHK(c, launch_forward_split(c->R, c->nb_max, c->md_d, c->plan_d, c->plan_n[0], sl, c->rec_off_d,
                                       c->shot_base_d, so, c->ss, c->side));
        HIPCHK(c, launch_forward(c->R, c->nb_max, c->paired, c->md_d, c->plan_d + T, c->plan_n[1], sl, c->rec_off_d,
                                 c->shot_base_d, so, c->stream));
        TRY(join_side(c));
    } else if (!reuse) {
        if (split)
            HIPCHK(c, launch_forward_split(c->R, c->nb_max, c->md_d, c->shots_d + b0, b1 - b0, sl, c->rec_off_d,
                                           c->shot_base_d, so, c->ss, c->stream));
        else
            HIPCHK(c, launch_forward(c->R, c->nb_max, c->paired, c->md_d, c->shots_d + b0, b1 - b0, sl, c->rec_off_d,
                                     c->shot_base_d, so, c->stream));
    }
    TRY(mark(c, &m.e[1]));
    if (!grad) return RZ_OK;
    for (size_t mi = 0; mi < c->m.size(); ++mi) {
        const ModelHost &h = c->m[mi];
        const int s0 = std::max(b0, h.shot_base) - h.shot_base;
        const int s1 = std::min(b1, h.shot_base + h.d.nshots) - h.shot_base;
        if (s1 <= s0 || h.d.ntr <= 0) continue;
        HIPCHK(c, launch_residual(c->md_d, (int)mi, h.d.ntr, s0, s1, h.rec_off, c->tmod_d, c->tnb_d, c->tobs_d,
                                  direct ? c->dres_d : nullptr, c->carry_d + h.carry_off, c->tres_d,
                                  c->l2_shot_d + h.shot_base, c->l2_sum_d + mi, c->stream));
    }
    TRY(mark(c, &m.e[2]));
    if (mixed) {
        const long long T = c->total_shots;
        TRY(fork_side(c));
        HIPCHK(c, launch_adjoint_split(c->R, c->nb_max, c->md_d, c->plan_d + 2 * T, c->plan_n[2], sl, c->rec_off_d,
                                       c->shot_base_d, c->tres_d, so, adj_variant(c), c->ss, c->side));
        HIPCHK(c, launch_adjoint(c->R, c->nb_max, c->paired, c->md_d, c->plan_d + 3 * T, c->plan_n[3], sl, c->rec_off_d,
                                 c->shot_base_d, c->tres_d, so, adj_variant(c), c->stream));
        TRY(join_side(c));
    } else if (split)
        HIPCHK(c, launch_adjoint_split(c->R, c->nb_max, c->md_d, c->shots_d + b0, b1 - b0, sl, c->rec_off_d,
                                       c->shot_base_d, c->tres_d, so, adj_variant(c), c->ss, c->stream));
    else
        HIPCHK(c, launch_adjoint(c->R, c->nb_max, c->paired, c->md_d, c->shots_d + b0, b1 - b0, sl, c->rec_off_d,
                                 c->shot_base_d, c->tres_d, so, adj_variant(c), c->stream));
    TRY(mark(c, &m.e[3]));
    for (size_t mi = 0; mi < c->m.size(); ++mi) {
        const ModelHost &h = c->m[mi];
        const int s0 = std::max(b0, h.shot_base) - h.shot_base;
        const int s1 = std::min(b1, h.shot_base + h.d.nshots) - h.shot_base;
        if (s1 <= s0) continue;
        ModelDev mh = make_dev(c, h);
        HIPCHK(c, launch_gradient(&mh, (int)mi, c->md_d, c->slot_of_d + h.shot_base + s0,
                                  c->lsel_d + h.shot_base + s0, s1 - s0, sl,
                                  c->grad_d + h.grad_off, s0 > 0 ? 1 : 0, taper_of(h, h.tap_first + s0),
                                  c->stream));
    }
    if (b1 == c->total_shots)   // grad_obj.c:70: with a per-shot taper the sum's frame is copied again
        for (const ModelHost &h : c->m)
            if (h.tap_mode != 0)
                HIPCHK(c, launch_grad_frame(grid_nx(h), grid_ny(h), c->grad_d + h.grad_off, c->stream));
    TRY(mark(c, &m.e[4]));
    return RZ_OK;
}

// Record a whole single-batch evaluation (forward only, or the gradient) as
// a HIP graph on the work stream.  It holds kernels only: fixed pointers
// into the context's buffers, no memset or host copy nodes.
int record_graph(rz_ctx *c, bool grad, hipGraphExec_t *exec)
{
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = RZ_OK;
    if (grad) {
        const hipError_t e = launch_eval_init(c->carry_d, c->carry0_d, c->total_tr, c->l2_sum_d, (int)c->m.size(),
                                              c->grad_d, c->total_cells, c->stream);
        if (e != hipSuccess) rc = fail(c, RZ_ERR_HIP, "evaluation init: %s", hipGetErrorString(e));
    }
    if (rc == RZ_OK) rc = enqueue_batch(c, grad, false, false, 0, c->total_shots, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != RZ_OK) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc;
    }
    HIPCHK(c, e);
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(c, ei);
    return RZ_OK;
}

// Run forward (and, with grad, residual + adjoint + gradient) over the
// global shot order in batches of at most `cap` shots.
int run(rz_ctx *c, bool grad, int cap, bool direct = false)
{
    if (direct) {
        for (size_t i = 0; i < c->m.size(); ++i)
            if (!c->m[i].have_res && c->m[i].d.nshots > 0 && c->m[i].d.ntr > 0)
                return fail(c, RZ_ERR_STATE, "model %zu: residuals not set", i);
    }
    TRY(ready(c, grad && !direct));
    // a context unsplit by a failed launch (its two workgroups of a shot
    // apart, or a wait timed out, e.g. with another process on the GPU)
    // tries split launches again after a back-off
    if (!c->split && c->split_fail && c->split_wait > 0 && --c->split_wait == 0) {
        c->split = true;
        drop_graphs(c);
    }
    // forward fields of every shot are still resident and current: the
    // adjoint-only run does not recompute them
    const bool reuse = direct && c->fields_epoch == c->s_epoch && c->nslots >= c->total_shots;
    HIPCHK(c, hipSetDevice(c->device));
    // Shot g lives in slot g % nslots; any batch of at most nslots
    // consecutive shots therefore uses distinct slots.
    const int per = std::max(1, std::min(cap, c->nslots));
    TRY(plan_mixed(c, per));
    const bool use_graph = c->graph && !direct && !c->prof && c->stream != nullptr && c->total_shots > 0 &&
                           per >= c->total_shots;
    if (use_graph) {
        hipGraphExec_t &ge = c->gexec[grad ? 1 : 0];
        if (!ge) TRY(record_graph(c, grad, &ge));
        HIPCHK(c, hipGraphLaunch(ge, c->stream));
        c->last_b0 = 0;
        c->last_b1 = c->total_shots;
    } else {
        // models without shots still report a zero gradient
        if (grad)
            HIPCHK(c, launch_eval_init(c->carry_d, c->carry0_d, c->total_tr, c->l2_sum_d, (int)c->m.size(),
                                       c->grad_d, c->total_cells, c->stream));
        std::vector<PhaseMarks> marks;
        for (int b0 = 0; b0 < c->total_shots; b0 += per) {
            const int b1 = std::min(c->total_shots, b0 + per);
            PhaseMarks pm;
            TRY(enqueue_batch(c, grad, direct, reuse, b0, b1, &pm));
            if (c->prof) marks.push_back(pm);
            c->last_b0 = b0;
            c->last_b1 = b1;
        }
        TRY(collect(c, marks, grad));
    }
    c->have_fwd = true;
    c->have_grad = grad;
    // all shots' forward fields stay in their slots only without batching
    if (c->nslots >= c->total_shots && !(grad && cap < c->nslots)) c->fields_epoch = c->s_epoch;
    else if (!reuse) c->fields_epoch = -1;
    const int rc = check_convergence(c, grad);
    if (rc == kSplitRetry) {   // split shots cannot run here: repeat everything unsplit
        drop_graphs(c);
        c->fields_epoch = -1;
        return run(c, grad, cap, direct);
    }
    return rc;
}

void drop_graphs(rz_ctx *c)
{
    for (hipGraphExec_t &g : c->gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

int copy_out(rz_ctx *c, void *dst, const void *src, size_t bytes, int on_device)
{
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RZ_OK;
}

// A row-major field of model h as stored (f): in the caller's orientation,
// transposed into s_row2 if the grid is stored transposed (null if that
// launch fails).
const float *to_grid(rz_ctx *c, const ModelHost &h, const float *f)
{
    if (!h.trans) return f;
    if (launch_transpose(f, c->s_row2, h.d.ny, h.d.nx, c->stream) != hipSuccess) return nullptr;
    return c->s_row2;
}

}  // namespace

extern "C" {

int rz_grid_index(float pos_m, float dh)
{
    const float q = pos_m / dh;
    return (int)std::floor((double)q + 0.5) - 1;
}

static thread_local std::string g_create_err;

const char *rz_last_error(const rz_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

}  // extern "C"

// Errors of calls without a context (rz_last_error(NULL)).
void rz::set_global_error(const char *msg) { g_create_err = msg; }

extern "C" {

void rz_destroy(rz_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);   // also fine for the null stream
    if (c->own && c->own != c->stream) (void)hipStreamSynchronize(c->own);
    drop_graphs(c);
    for (void *p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

int rz_create(rz_ctx **out, int device, int nmodels, const rz_model_desc *models, int rows_per_lane,
              size_t hbm_budget)
{
    if (!out) return RZ_ERR_INVALID;
    *out = nullptr;
    rz_ctx *c = new rz_ctx();
    auto bail = [&](int rc) {
        // rz_last_error(NULL) reports why creation failed
        g_create_err = c->err;
        rz_destroy(c);
        return rc;
    };
    c->device = device;
    if (nmodels <= 0 || !models) return bail(fail(c, RZ_ERR_INVALID, "need at least one model"));
    int ny_max = 0;          // rows of the tallest grid as stored
    bool any_trans = false;
    for (int i = 0; i < nmodels; ++i) {
        const rz_model_desc &d = models[i];
        if (d.nx < 5 || d.ny < 5) return bail(fail(c, RZ_ERR_INVALID, "model %d: NX, NY must be >= 5", i));
        if (!(d.dh > 0.f) || !std::isfinite(d.dh)) return bail(fail(c, RZ_ERR_INVALID, "model %d: bad DH", i));
        if (!std::isfinite(d.time) || !std::isfinite(d.ttnorm) || !std::isfinite(d.eps_adj))
            return bail(fail(c, RZ_ERR_INVALID, "model %d: TIME/TTNORM/EPS_ADJ must be finite", i));
        if (d.nshots < 0 || d.ntr < 0 || d.max_iter < 0)
            return bail(fail(c, RZ_ERR_INVALID, "model %d: negative count", i));
        if ((long long)d.nx * d.ny > (1LL << 28))
            return bail(fail(c, RZ_ERR_INVALID, "model %d: grid too large", i));
        ny_max = std::max(ny_max, transposed(d) ? d.nx : d.ny);
        any_trans = any_trans || transposed(d);
    }
    if (rows_per_lane == RZ_ROWS_AUTOTUNE) {
        int best = 0;
        if (rz_autotune_rows(device, nmodels, models, hbm_budget, &best, nullptr) != RZ_OK)
            return bail(fail(c, RZ_ERR_INVALID, "autotune: %s", g_create_err.c_str()));
        rows_per_lane = best;
    }
    c->R = rows_per_lane ? rows_per_lane : (any_trans ? 8 : choose_R(ny_max));
    if (c->R != 1 && c->R != 2 && c->R != 4 && c->R != 8)
        return bail(fail(c, RZ_ERR_INVALID, "rows_per_lane must be 1, 2, 4 or 8"));
    if (any_trans && c->R != 8)
        return bail(fail(c, RZ_ERR_INVALID, "NY above %d is solved transposed, at 8 rows per lane only (not %d)",
                         kMaxRowsStored, c->R));
    if (bands_for(ny_max, c->R) > max_bands_unpaired(c->R))
        return bail(fail(c, RZ_ERR_INVALID, "NY=%d needs more than %d bands at %d rows per lane", ny_max,
                         max_bands_unpaired(c->R), c->R));

    long long rec_off = 0, cells = 0;
    int shots = 0, carry = 0;
    long long field_max = 0;
    for (int i = 0; i < nmodels; ++i) {
        const rz_model_desc &d = models[i];
        ModelHost h;
        h.d = d;
        h.d.src_ix = h.d.src_iy = h.d.rec_ix = h.d.rec_iy = nullptr;
        h.trans = transposed(d);
        if (h.trans) std::swap(h.d.nx, h.d.ny);
        const int nx = h.d.nx, ny = h.d.ny;   // as stored
        h.nb = bands_for(ny, c->R);
        h.bstride = (long long)(nx + 64) * kWave * c->R;   // keys 0..nx+62 + one dummy block
        field_max = std::max(field_max, (long long)h.nb * h.bstride);
        // the adjoint reads a field through one buffer descriptor, whose
        // byte range is a 31-bit count
        if (4 * (long long)h.nb * h.bstride > 0x7fffffffLL)
            return bail(fail(c, RZ_ERR_INVALID, "model %d: a %d x %d field exceeds 2 GiB in the band layout", i,
                             d.nx, d.ny));
        // a cell (x, y) of the grid is (row y, column x) as stored, or (row
        // x, column y) transposed
        for (int s = 0; s < d.nshots; ++s) {
            const int x = d.src_ix[s], y = d.src_iy[s];
            if (x < 0 || x >= d.nx || y < 0 || y >= d.ny)
                return bail(fail(c, RZ_ERR_INVALID, "model %d shot %d: source cell (%d,%d) outside the grid", i, s,
                                 x, y));
            h.src_c.push_back(h.trans ? y : x);
            h.src_k.push_back(h.trans ? x : y);
        }
        std::unordered_map<long long, int> last;
        for (int r = 0; r < d.ntr; ++r) {
            const int x = d.rec_ix[r], y = d.rec_iy[r];
            if (x < 0 || x >= d.nx || y < 0 || y >= d.ny)
                return bail(fail(c, RZ_ERR_INVALID, "model %d receiver %d: cell (%d,%d) outside the grid", i, r,
                                 x, y));
            const int kk = h.trans ? x : y, cc = h.trans ? y : x;
            h.rec_c.push_back(cc);
            h.rec_k.push_back(kk);
            if (cc > 0 && cc < nx - 1 && kk > 0 && kk < ny - 1) last[(long long)kk * nx + cc] = r;
        }
        // interior receiver cells, last receiver wins (init_travel_adj.c:49-55)
        std::vector<std::pair<long long, int>> uq(last.begin(), last.end());
        std::sort(uq.begin(), uq.end());
        for (auto &p : uq) {
            h.uq_k.push_back((int)(p.first / nx));
            h.uq_c.push_back((int)(p.first % nx));
            h.uq_r.push_back(p.second);
        }
        set_stop(h);
        h.fastdiv = fast_div_exact(d.dh);
        c->fastdiv = c->fastdiv && h.fastdiv;
        set_ft(h);
        // T starts at 2*TIME and only decreases, so the guard TT >= INF
        // (INF = 9999999, fd.h:29-31) is dead unless 2*TIME reaches it.
        c->chkinf = c->chkinf || (float)(2.0 * (double)d.time) >= kInf;
        h.rec_off = rec_off;
        h.shot_base = shots;
        h.grad_off = cells;
        h.carry_off = carry;
        rec_off += (long long)d.nshots * d.ntr;
        shots += d.nshots;
        cells += (long long)d.nx * d.ny;
        carry += d.ntr;
        c->nb_max = std::max(c->nb_max, h.nb);
        c->m.push_back(std::move(h));
    }
    c->paired = paired_for(c->R, c->nb_max) && !any_trans;
    c->total_shots = shots;
    c->total_rec = rec_off;
    c->total_cells = cells;
    c->total_tr = carry;
    c->slot_floats = (field_max + 63) / 64 * 64;

    int rc;
    if (hipSetDevice(device) != hipSuccess) return bail(fail(c, RZ_ERR_HIP, "hipSetDevice(%d) failed", device));
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, RZ_ERR_HIP, "hipStreamCreate failed"));
    c->stream = c->own;

    // per-model buffers
    for (auto &h : c->m) {
        const size_t n = (size_t)h.d.nx * h.d.ny, f = (size_t)h.nb * h.bstride;
        if ((rc = dalloc(c, &h.S, n)) || (rc = dalloc(c, &h.SA, f)) || (rc = dalloc(c, &h.SB, f)) ||
            (rc = dalloc(c, &h.rec_k_d, h.rec_k.size())) || (rc = dalloc(c, &h.rec_c_d, h.rec_c.size())) ||
            (rc = dalloc(c, &h.uq_k_d, h.uq_k.size())) || (rc = dalloc(c, &h.uq_c_d, h.uq_c.size())) ||
            (rc = dalloc(c, &h.uq_r_d, h.uq_r.size())))
            return bail(rc);
        auto up = [&](int *dst, const std::vector<int> &v) {
            return v.empty() ? hipSuccess : hipMemcpy(dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice);
        };
        if (up(h.rec_k_d, h.rec_k) || up(h.rec_c_d, h.rec_c) || up(h.uq_k_d, h.uq_k) || up(h.uq_c_d, h.uq_c) ||
            up(h.uq_r_d, h.uq_r))
            return bail(fail(c, RZ_ERR_HIP, "geometry upload failed"));
    }
    const size_t nm = c->m.size();
    if ((rc = dalloc(c, &c->md_d, nm)) || (rc = dalloc(c, &c->shots_d, shots)) ||
        (rc = dalloc(c, &c->slot_of_d, shots)) || (rc = dalloc(c, &c->rec_off_d, nm)) ||
        (rc = dalloc(c, &c->rec_off0_d, nm)) || (rc = dalloc(c, &c->shot_base_d, nm)) ||
        (rc = dalloc(c, &c->shot_base0_d, nm)) || (rc = dalloc(c, &c->tobs_d, rec_off)) ||
        (rc = dalloc(c, &c->tmod_d, rec_off)) || (rc = dalloc(c, &c->tnb_d, rec_off)) ||
        (rc = dalloc(c, &c->tres_d, rec_off)) || (rc = dalloc(c, &c->carry_d, carry)) ||
        (rc = dalloc(c, &c->carry0_d, carry)) ||
        (rc = dalloc(c, &c->dres_d, rec_off)) ||
        (rc = dalloc(c, &c->grad_d, cells)) || (rc = dalloc(c, &c->l2_shot_d, shots)) ||
        (rc = dalloc(c, &c->l2_sum_d, nm)) || (rc = dalloc(c, &c->it_f_d, shots)) ||
        (rc = dalloc(c, &c->it_a_d, shots)) || (rc = dalloc(c, &c->lsel_d, shots)) ||
        (rc = dalloc(c, &c->ord_f_d, shots)) || (rc = dalloc(c, &c->ord_a_d, shots)) ||
        (rc = dalloc(c, &c->osum_f_d, shots)) || (rc = dalloc(c, &c->osum_a_d, shots)) ||
        (rc = dalloc(c, &c->phases_f_d, shots)) ||
        (rc = dalloc(c, &c->err_d, 1)))
        return bail(rc);
    int ntr_max = 0, ncell_max = 0;
    for (auto &h : c->m) {
        ntr_max = std::max(ntr_max, h.d.ntr);
        ncell_max = std::max(ncell_max, h.d.nx * h.d.ny);
    }
    if ((rc = dalloc(c, &c->s_tmod, ntr_max)) || (rc = dalloc(c, &c->s_tnb, ntr_max)) ||
        (rc = dalloc(c, &c->s_tres, ntr_max)) || (rc = dalloc(c, &c->s_row, ncell_max)) ||
        (rc = dalloc(c, &c->s_row2, ncell_max)) || (rc = dalloc(c, &c->s_it, 6)) || (rc = dalloc(c, &c->s_osum, 2)) ||
        (rc = dalloc(c, &c->s_shot, 1)) || (rc = dalloc(c, &c->s_slot0, 1)))
        return bail(rc);

    // field slots (T in two layouts, lambda in three buffers): all shots, or
    // as many as the budget allows
    const size_t per_slot = (size_t)c->slot_floats * 5 * sizeof(float);
    int nslots = std::max(1, shots);
    if (hbm_budget == 0) {
        // no budget given: every shot if the fields fit the device's free
        // memory less a reserve of 1/16 of the device (the tables below,
        // the caller's own buffers), else as many shots as do fit
        size_t mfree = 0, mtotal = 0;
        if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess && mtotal > 0) {
            const size_t reserve = mtotal / 16;
            const size_t room = mfree > reserve ? mfree - reserve : 0;
            if ((size_t)nslots * per_slot > room) hbm_budget = std::max(room, per_slot);
        }
    }
    if (hbm_budget > 0) {
        const size_t fit = hbm_budget / per_slot;
        if (fit < 1) return bail(fail(c, RZ_ERR_NOMEM, "HBM budget below one shot's fields (%zu bytes)", per_slot));
        nslots = (int)std::min<size_t>((size_t)nslots, fit);
    }
    c->nslots = nslots;
    const size_t sf = (size_t)c->slot_floats * nslots;
    if ((rc = dalloc(c, &c->TA, sf)) || (rc = dalloc(c, &c->TB, sf)) || (rc = dalloc(c, &c->LA, sf)) ||
        (rc = dalloc(c, &c->LB, sf)) || (rc = dalloc(c, &c->LC, sf)))
        return bail(rc);
    {   // RAJZEL_AMD_POISON=1: every field starts as NaN, so that a grid cell
        // that a kernel reads before any kernel wrote it shows in the results
        const char *env = getenv("RAJZEL_AMD_POISON");
        if (env && env[0] == '1')
            for (float *f : {c->TA, c->TB, c->LA, c->LB, c->LC})
                if (hipMemset(f, 0xFF, sf * sizeof(float)) != hipSuccess)
                    return bail(fail(c, RZ_ERR_HIP, "poisoning the fields failed"));
    }
    {
        const size_t nsh = (size_t)std::max(1, shots);
        if ((rc = dalloc(c, &c->ss.cnt, nsh * kSyncInts)) || (rc = dalloc(c, &c->ss.part, nsh * kSyncPart)) ||
            (rc = dalloc(c, &c->ss.err, 1)))
            return bail(rc);
        if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
            c->ncu = 0;
        const char *env = getenv("RAJZEL_AMD_SPLIT");
        c->split = !(env && env[0] == '0');
        const char *eft = getenv("RAJZEL_AMD_FLOAT_T");
        c->ft_allowed = !(eft && eft[0] == '0');
    }

    // tables
    std::vector<ModelDev> mdh;
    std::vector<long long> ro, ro0(nm, 0);
    std::vector<int> sb, sb0(nm, 0);
    for (size_t i = 0; i < nm; ++i) {
        const ModelHost &h = c->m[i];
        mdh.push_back(make_dev(c, h));
        ro.push_back(h.rec_off);
        sb.push_back(h.shot_base);
        for (int s = 0; s < h.d.nshots; ++s) {
            ShotDev sd{};
            sd.model = (int)i;
            sd.shot = s;
            sd.src_k = h.src_k[s];
            sd.src_c = h.src_c[s];
            sd.slot = (h.shot_base + s) % nslots;
            c->shots_h.push_back(sd);
        }
    }
    std::vector<int> slot_of(shots);
    for (int g = 0; g < shots; ++g) slot_of[g] = c->shots_h[g].slot;
    const int zero = 0;
    ShotDev s1{};
    if (hipMemcpy(c->md_d, mdh.data(), nm * sizeof(ModelDev), hipMemcpyHostToDevice) ||
        (shots && hipMemcpy(c->shots_d, c->shots_h.data(), shots * sizeof(ShotDev), hipMemcpyHostToDevice)) ||
        (shots && hipMemcpy(c->slot_of_d, slot_of.data(), shots * sizeof(int), hipMemcpyHostToDevice)) ||
        hipMemcpy(c->rec_off_d, ro.data(), nm * sizeof(long long), hipMemcpyHostToDevice) ||
        hipMemcpy(c->rec_off0_d, ro0.data(), nm * sizeof(long long), hipMemcpyHostToDevice) ||
        hipMemcpy(c->shot_base_d, sb.data(), nm * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(c->shot_base0_d, sb0.data(), nm * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(c->s_slot0, &zero, sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(c->s_shot, &s1, sizeof(ShotDev), hipMemcpyHostToDevice) ||
        hipMemset(c->carry0_d, 0, (size_t)std::max(c->total_tr, 1) * sizeof(float)) ||
        hipMemset(c->ss.err, 0, sizeof(int)) || hipDeviceSynchronize())
        return bail(fail(c, RZ_ERR_HIP, "table upload failed"));
    *out = c;
    return RZ_OK;
}

int rz_set_graph(rz_ctx *c, int on)
{
    if (!c) return RZ_ERR_INVALID;
    c->graph = on != 0;
    if (!c->graph) {
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        drop_graphs(c);
    }
    return RZ_OK;
}

int rz_set_stream(rz_ctx *c, void *s)
{
    if (!c) return RZ_ERR_INVALID;
    c->stream = static_cast<hipStream_t>(s);   // NULL: HIP's null stream
    return RZ_OK;
}

void *rz_get_stream(const rz_ctx *c) { return c ? static_cast<void *>(c->stream) : nullptr; }

int rz_set_slowness(rz_ctx *c, int model, const float *S, int on_device)
{
    TRY(check_model(c, model));
    if (!S) return fail(c, RZ_ERR_INVALID, "null slowness");
    ModelHost &h = c->m[model];
    const size_t n = (size_t)h.d.nx * h.d.ny;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (h.trans) {   // the caller's NY x NX rows become the stored grid's columns
        HIPCHK(c, hipMemcpyAsync(c->s_row2, S, n * sizeof(float), kind, c->stream));
        HIPCHK(c, launch_transpose(c->s_row2, h.S, grid_ny(h), grid_nx(h), c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(h.S, S, n * sizeof(float), kind, c->stream));
    }
    HIPCHK(c, hipMemsetAsync(c->err_d, 0, sizeof(int), c->stream));
    ModelDev mh = make_dev(c, h);
    HIPCHK(c, launch_prepare_slowness(&mh, model, c->md_d, h.SA, h.SB, c->err_d, c->stream));
    int bad = 0;
    HIPCHK(c, hipMemcpyAsync(&bad, c->err_d, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bad & 1) {
        h.have_S = false;
        return fail(c, RZ_ERR_INVALID, "model %d: slowness must be finite and > 0 everywhere", model);
    }
    h.have_S = true;
    h.ft_s = !(bad & 2);
    bool ft = c->ft_allowed;
    for (const ModelHost &x : c->m) ft = ft && x.ft_dh && (x.ft_s || !x.have_S);
    if (ft != c->ft) {
        c->ft = ft;
        drop_graphs(c);   // recorded adjoint launches carry the variant
    }
    ++c->s_epoch;
    return RZ_OK;
}

int rz_set_picks(rz_ctx *c, int model, const float *tobs, int on_device)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    const size_t n = (size_t)h.d.nshots * h.d.ntr;
    if (n == 0) { h.have_picks = true; return RZ_OK; }
    if (!tobs) return fail(c, RZ_ERR_INVALID, "null picks");
    HIPCHK(c, hipMemcpyAsync(c->tobs_d + h.rec_off, tobs, n * sizeof(float),
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    h.have_picks = true;
    return RZ_OK;
}

int rz_set_tres_carry(rz_ctx *c, int model, const float *tres0)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    if (h.d.ntr == 0) return RZ_OK;
    std::vector<float> v(tres0 ? tres0 : nullptr, tres0 ? tres0 + h.d.ntr : nullptr);
    v.resize(h.d.ntr, 0.f);
    HIPCHK(c, hipMemcpyAsync(c->carry0_d + h.carry_off, v.data(), h.d.ntr * sizeof(float), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RZ_OK;
}

int rz_forward(rz_ctx *c)
{
    if (!c) return RZ_ERR_INVALID;
    return run(c, false, c->nslots);
}

int rz_gradient(rz_ctx *c)
{
    if (!c) return RZ_ERR_INVALID;
    return run(c, true, c->nslots);
}

int rz_gradient_per_shot(rz_ctx *c)
{
    if (!c) return RZ_ERR_INVALID;
    return run(c, true, 1);
}

int rz_set_residuals(rz_ctx *c, int model, const float *dT, int on_device)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    const size_t n = (size_t)h.d.nshots * h.d.ntr;
    if (n == 0) { h.have_res = true; return RZ_OK; }
    if (!dT) return fail(c, RZ_ERR_INVALID, "null residuals");
    HIPCHK(c, hipMemcpyAsync(c->dres_d + h.rec_off, dT, n * sizeof(float),
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    h.have_res = true;
    return RZ_OK;
}

int rz_adjoint(rz_ctx *c)
{
    if (!c) return RZ_ERR_INVALID;
    return run(c, true, c->nslots, true);
}

int rz_get_tmod(rz_ctx *c, int model, float *out, int on_device)
{
    TRY(check_model(c, model));
    if (!c->have_fwd) return fail(c, RZ_ERR_STATE, "no forward has run");
    const ModelHost &h = c->m[model];
    return copy_out(c, out, c->tmod_d + h.rec_off, (size_t)h.d.nshots * h.d.ntr * sizeof(float), on_device);
}

int rz_get_tnb(rz_ctx *c, int model, float *out, int on_device)
{
    TRY(check_model(c, model));
    if (!c->have_fwd) return fail(c, RZ_ERR_STATE, "no forward has run");
    const ModelHost &h = c->m[model];
    return copy_out(c, out, c->tnb_d + h.rec_off, (size_t)h.d.nshots * h.d.ntr * sizeof(float), on_device);
}

int rz_get_tres(rz_ctx *c, int model, float *out, int on_device)
{
    TRY(check_model(c, model));
    if (!c->have_grad) return fail(c, RZ_ERR_STATE, "no gradient has run");
    const ModelHost &h = c->m[model];
    return copy_out(c, out, c->tres_d + h.rec_off, (size_t)h.d.nshots * h.d.ntr * sizeof(float), on_device);
}

int rz_get_grad(rz_ctx *c, int model, float *out, int on_device)
{
    TRY(check_model(c, model));
    if (!c->have_grad) return fail(c, RZ_ERR_STATE, "no gradient has run");
    const ModelHost &h = c->m[model];
    return copy_out(c, out, c->grad_d + h.grad_off, (size_t)h.d.nx * h.d.ny * sizeof(float), on_device);
}

int rz_get_misfit(rz_ctx *c, int model, double *L2, double *l2_shot)
{
    TRY(check_model(c, model));
    if (!c->have_grad) return fail(c, RZ_ERR_STATE, "no gradient has run");
    const ModelHost &h = c->m[model];
    double sum = 0.0;
    HIPCHK(c, hipMemcpyAsync(&sum, c->l2_sum_d + model, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (l2_shot && h.d.nshots)
        HIPCHK(c, hipMemcpyAsync(l2_shot, c->l2_shot_d + h.shot_base, h.d.nshots * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (L2) *L2 = 0.5 * sum;   // grad_obj.c:61
    return RZ_OK;
}

int rz_get_iterations(rz_ctx *c, int model, int *it_f, int *it_a)
{
    TRY(check_model(c, model));
    const ModelHost &h = c->m[model];
    if (it_f && h.d.nshots) TRY(copy_out(c, it_f, c->it_f_d + h.shot_base, h.d.nshots * sizeof(int), 0));
    if (it_a && h.d.nshots) TRY(copy_out(c, it_a, c->it_a_d + h.shot_base, h.d.nshots * sizeof(int), 0));
    return RZ_OK;
}

int rz_get_stop_passes(rz_ctx *c, int model, int *count_f, int *count_a, int *ordered_f, int *ordered_a,
                       float *sum_f, float *sum_a)
{
    TRY(check_model(c, model));
    const ModelHost &h = c->m[model];
    const int ns = h.d.nshots;
    std::vector<int> f(ns), a(ns);
    if (ns) {
        TRY(copy_out(c, f.data(), c->ord_f_d + h.shot_base, ns * sizeof(int), 0));
        TRY(copy_out(c, a.data(), c->ord_a_d + h.shot_base, ns * sizeof(int), 0));
        if (sum_f) TRY(copy_out(c, sum_f, c->osum_f_d + h.shot_base, ns * sizeof(float), 0));
        if (sum_a) TRY(copy_out(c, sum_a, c->osum_a_d + h.shot_base, ns * sizeof(float), 0));
    }
    for (int i = 0; i < ns; ++i) {
        if (count_f) count_f[i] = f[i] & 0xFFFF;
        if (count_a) count_a[i] = a[i] & 0xFFFF;
        if (ordered_f) ordered_f[i] = f[i] >> 16;
        if (ordered_a) ordered_a[i] = a[i] >> 16;
    }
    return RZ_OK;
}

int rz_get_tres_carry(rz_ctx *c, int model, float *out)
{
    TRY(check_model(c, model));
    const ModelHost &h = c->m[model];
    if (h.d.ntr == 0) return RZ_OK;
    return copy_out(c, out, c->carry_d + h.carry_off, h.d.ntr * sizeof(float), 0);
}

int rz_get_field(rz_ctx *c, int model, int shot, int which, float *out, int on_device)
{
    TRY(check_model(c, model));
    const ModelHost &h = c->m[model];
    if (shot < 0 || shot >= h.d.nshots) return fail(c, RZ_ERR_INVALID, "shot out of range");
    if (which != 0 && which != 1) return fail(c, RZ_ERR_INVALID, "which must be 0 (T) or 1 (lambda)");
    if (which == 1 && !c->have_grad) return fail(c, RZ_ERR_STATE, "no gradient has run");
    const int g = h.shot_base + shot;
    if (g < c->last_b0 || g >= c->last_b1)
        return fail(c, RZ_ERR_STATE, "fields of shot %d are not resident (HBM budget batching)", shot);
    const int slot = g % c->nslots;
    ModelDev mh = make_dev(c, h);
    const long long so = (long long)slot * c->slot_floats;
    const size_t n = (size_t)h.d.nx * h.d.ny;
    if (which == 0)
        HIPCHK(c, launch_field_to_rowmajor(&mh, c->TA + so, nullptr, nullptr, c->s_row, 1, c->stream));
    else
        HIPCHK(c, launch_field_to_rowmajor(&mh, c->LA + so, c->LC + so, c->lsel_d + g, c->s_row, 0, c->stream));
    return copy_out(c, out, to_grid(c, h, c->s_row), n * sizeof(float), on_device);
}

int rz_eikonal_shot(rz_ctx *c, int model, int src_ix, int src_iy, float *T, float *tmod, int *iters)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    if (!h.have_S) return fail(c, RZ_ERR_STATE, "slowness not set");
    if (src_ix < 0 || src_ix >= grid_nx(h) || src_iy < 0 || src_iy >= grid_ny(h))
        return fail(c, RZ_ERR_INVALID, "source cell (%d,%d) outside the grid", src_ix, src_iy);
    ShotDev sd{};
    sd.model = model; sd.shot = 0; sd.slot = 0;
    sd.src_k = h.trans ? src_ix : src_iy;
    sd.src_c = h.trans ? src_iy : src_ix;
    HIPCHK(c, hipMemcpyAsync(c->s_shot, &sd, sizeof sd, hipMemcpyHostToDevice, c->stream));
    ShotOut o;
    o.tmod = c->s_tmod; o.tnb = c->s_tnb; o.it_f = c->s_it; o.it_a = c->s_it + 1; o.lsel = c->s_it + 2;
    o.ord_f = c->s_it + 3; o.ord_a = c->s_it + 4;
    o.osum_f = c->s_osum; o.osum_a = c->s_osum + 1; o.phases_f = c->s_it + 5;
    HIPCHK(c, launch_forward(c->R, c->nb_max, c->paired, c->md_d, c->s_shot, 1, slots_of(c), c->rec_off0_d, c->shot_base0_d, o,
                             c->stream));
    ModelDev mh = make_dev(c, h);
    HIPCHK(c, launch_field_to_rowmajor(&mh, c->TA, nullptr, nullptr, c->s_row, 1, c->stream));
    int it = 0;
    const size_t n = (size_t)h.d.nx * h.d.ny;
    if (T) HIPCHK(c, hipMemcpyAsync(T, to_grid(c, h, c->s_row), n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (tmod && h.d.ntr)
        HIPCHK(c, hipMemcpyAsync(tmod, c->s_tmod, h.d.ntr * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&it, c->s_it, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->last_b0 = c->last_b1 = -1;
    c->fields_epoch = -1;
    if (iters) *iters = it;
    if (it < 0) return fail(c, RZ_ERR_NOT_CONVERGED, "forward did not converge within max_iter");
    return RZ_OK;
}

int rz_eikonal_adj_shot(rz_ctx *c, int model, const float *T, const float *tres, float *lam, int *iters)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    if (!T) return fail(c, RZ_ERR_INVALID, "null T");
    const size_t n = (size_t)h.d.nx * h.d.ny;
    std::vector<float> Tt;   // T as stored, for a grid stored transposed
    if (h.trans) {
        const int gx = grid_nx(h), gy = grid_ny(h);
        Tt.resize(n);
        for (int y = 0; y < gy; ++y)
            for (int x = 0; x < gx; ++x) Tt[(size_t)x * gy + y] = T[(size_t)y * gx + x];
        T = Tt.data();
    }
    bool big = false;
    float tmin = INFINITY, tmax = 0.f;   // smallest nonzero and largest T (float T-only quotients)
    for (size_t i = 0; i < n; ++i) {
        if (!(T[i] >= 0.f)) return fail(c, RZ_ERR_INVALID, "T must be >= 0 (cell %zu)", i);
        big = big || T[i] >= kInf;
        if (T[i] > 0.f && T[i] < tmin) tmin = T[i];
        tmax = std::max(tmax, T[i]);
    }
    // the caller's T, not the slowness, bounds the quotients' inputs here
    int variant = (adj_variant(c) & ~kAdjFloatT) | (big ? kAdjInfGuard : 0);
    if (c->ft_allowed && c->fastdiv && h.ft_dh && (double)tmin >= h.t_min_ft &&
        (double)tmax / h.d.dh < std::ldexp(1.0, 120))
        variant |= kAdjFloatT;
    ModelDev mh = make_dev(c, h);
    HIPCHK(c, hipMemcpyAsync(c->s_row, T, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    // receiver and frame cells carry the sign bit in the adjoint's copy of T
    // (as mark_kept_cells sets it after a forward)
    // |T| everywhere (a -0 passes the check above and must not read as a
    // mark), then the sign bit on the cells that keep their value
    std::vector<float> marked(n);
    for (size_t i = 0; i < n; ++i) marked[i] = std::fabs(T[i]);
    const int nx = h.d.nx, ny = h.d.ny;
    // -|T| (-0 for 0); idempotent, since the frame corners are marked twice
    auto keep = [&](size_t i) { marked[i] = -std::fabs(marked[i]); };
    for (size_t u = 0; u < h.uq_k.size(); ++u) keep((size_t)h.uq_k[u] * nx + h.uq_c[u]);
    for (int c = 0; c < nx; ++c) { keep((size_t)c); keep((size_t)(ny - 1) * nx + c); }
    for (int k = 0; k < ny; ++k) { keep((size_t)k * nx); keep((size_t)k * nx + nx - 1); }
    HIPCHK(c, hipMemcpyAsync(c->s_row2, marked.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_rowmajor_to_field(&mh, c->s_row2, c->TA, 0, c->stream));
    HIPCHK(c, launch_rowmajor_to_field(&mh, c->s_row2, c->TB, 1, c->stream));
    if (h.d.ntr)
        HIPCHK(c, hipMemcpyAsync(c->s_tres, tres, h.d.ntr * sizeof(float), hipMemcpyHostToDevice, c->stream));
    ShotDev sd{};
    sd.model = model; sd.slot = 0;
    HIPCHK(c, hipMemcpyAsync(c->s_shot, &sd, sizeof sd, hipMemcpyHostToDevice, c->stream));
    ShotOut o;
    o.tmod = c->s_tmod; o.tnb = c->s_tnb; o.it_f = c->s_it; o.it_a = c->s_it + 1; o.lsel = c->s_it + 2;
    o.ord_f = c->s_it + 3; o.ord_a = c->s_it + 4;
    o.osum_f = c->s_osum; o.osum_a = c->s_osum + 1; o.phases_f = c->s_it + 5;
    HIPCHK(c, launch_adjoint(c->R, c->nb_max, c->paired, c->md_d, c->s_shot, 1, slots_of(c), c->rec_off0_d, c->shot_base0_d,
                             c->s_tres, o, variant, c->stream));
    HIPCHK(c, launch_field_to_rowmajor(&mh, c->LA, c->LC, c->s_it + 2, c->s_row, 0, c->stream));
    int it = 0;
    if (lam)
        HIPCHK(c, hipMemcpyAsync(lam, to_grid(c, h, c->s_row), n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&it, c->s_it + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->last_b0 = c->last_b1 = -1;
    c->fields_epoch = -1;
    if (iters) *iters = it;
    if (it < 0) return fail(c, RZ_ERR_NOT_CONVERGED, "adjoint did not converge within max_iter");
    return RZ_OK;
}

static int ass_grad_shot(rz_ctx *c, int model, int src, const float *lam, float *grad_shot, float *grad)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    if (src >= 0 && h.tap_mode == 1 && src >= h.tap_nsrc)
        return fail(c, RZ_ERR_INVALID, "source %d out of range: the taper holds %d sources", src, h.tap_nsrc);
    if (!h.have_S) return fail(c, RZ_ERR_STATE, "slowness not set");
    if (!lam) return fail(c, RZ_ERR_INVALID, "null lambda");
    const size_t n = (size_t)h.d.nx * h.d.ny;
    ModelDev mh = make_dev(c, h);
    if (h.trans) {   // lambda as stored; the gradient comes out in the caller's orientation
        HIPCHK(c, hipMemcpyAsync(c->s_row2, lam, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_transpose(c->s_row2, c->s_row, grid_ny(h), grid_nx(h), c->stream));
    } else {
        HIPCHK(c, hipMemcpyAsync(c->s_row, lam, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, launch_rowmajor_to_field(&mh, c->s_row, c->LA, 0, c->stream));
    // slot 0, lambda in LA: s_slot0 holds the 0 for both
    HIPCHK(c, launch_gradient(&mh, model, c->md_d, c->s_slot0, c->s_slot0, 1, slots_of(c), c->s_row2, 0,
                              src >= 0 ? taper_of(h, src) : TaperDev{}, c->stream));
    std::vector<float> gs(n);
    HIPCHK(c, hipMemcpyAsync(gs.data(), c->s_row2, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->last_b0 = c->last_b1 = -1;
    c->fields_epoch = -1;
    if (grad_shot) std::memcpy(grad_shot, gs.data(), n * sizeof(float));
    if (grad)
        for (size_t i = 0; i < n; ++i) grad[i] += gs[i];   // ass_grad.c:40-46
    return RZ_OK;
}

int rz_ass_grad_shot(rz_ctx *c, int model, const float *lam, float *grad_shot, float *grad)
{
    return ass_grad_shot(c, model, -1, lam, grad_shot, grad);
}

int rz_ass_grad_shot_at(rz_ctx *c, int model, int src, const float *lam, float *grad_shot, float *grad)
{
    if (src < 0) return fail(c, RZ_ERR_INVALID, "negative source index %d", src);
    return ass_grad_shot(c, model, src, lam, grad_shot, grad);
}

int rz_set_shot_taper(rz_ctx *c, int model, int mode, float radius_m, int filtsize, int nsrc, const float *src_x_m,
                      const float *src_z_m, int first_shot)
{
    TRY(check_model(c, model));
    ModelHost &h = c->m[model];
    if (mode < 0 || mode > 2) return fail(c, RZ_ERR_INVALID, "taper mode %d (0, 1 or 2)", mode);
    if (nsrc < 0 || (nsrc > 0 && (!src_x_m || !src_z_m)))
        return fail(c, RZ_ERR_INVALID, "source positions: %d sources, null pointer", nsrc);
    if (mode == 1) {
        if (!(radius_m > 0.f) || !std::isfinite(radius_m))
            return fail(c, RZ_ERR_INVALID, "taper radius %g m: must be finite and > 0", (double)radius_m);
        if (first_shot < 0 || (long long)first_shot + h.d.nshots > nsrc)
            return fail(c, RZ_ERR_INVALID, "shots %d..%d of model %d are not in the %d sources", first_shot,
                        first_shot + h.d.nshots - 1, model, nsrc);
    }
    if (mode == 0) nsrc = 0;
    const std::vector<float> sx(src_x_m, src_x_m + nsrc), sz(src_z_m, src_z_m + nsrc);
    if (mode == h.tap_mode && (mode == 0 || (sx == h.tap_sx && sz == h.tap_sz && first_shot == h.tap_first &&
                                             (mode == 1 ? radius_m == h.tap_radius : filtsize == h.tap_filt))))
        return RZ_OK;   // unchanged geometry: keep the tables
    TaperTables t;
    if (mode != 0) {
        // in the grid's orientation, as k_gradient applies the tables
        const std::string why = build_taper(mode, grid_nx(h), grid_ny(h), h.d.dh, radius_m, filtsize, nsrc, sx.data(),
                                            sz.data(), h.d.ntr, h.trans ? h.rec_k.data() : h.rec_c.data(),
                                            h.trans ? h.rec_c.data() : h.rec_k.data(), kTaperMaxEntries, t);
        if (!why.empty()) return fail(c, RZ_ERR_INVALID, "per-shot taper: %s", why.c_str());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (void *p : {(void *)h.tap_box, (void *)h.tap_off, (void *)h.tap_fsrc, (void *)h.tap_frec,
                    (void *)h.tap_start})
        dfree(c, p);
    h.tap_box = nullptr; h.tap_off = nullptr; h.tap_fsrc = nullptr; h.tap_frec = nullptr; h.tap_start = nullptr;
    h.tap_mode = 0;
    drop_graphs(c);   // a recorded gradient has no final frame copy, or one too many
    if (mode != 0) {
        TRY(dalloc(c, &h.tap_start, t.start.size()));
        HIPCHK(c, hipMemcpy(h.tap_start, t.start.data(), t.start.size() * sizeof(unsigned), hipMemcpyHostToDevice));
        if (mode == 1) {
            TRY(dalloc(c, &h.tap_box, t.box.size()));
            TRY(dalloc(c, &h.tap_off, t.off.size()));
            TRY(dalloc(c, &h.tap_fsrc, t.fsrc.size()));
            TRY(dalloc(c, &h.tap_frec, t.frec.size()));
            auto up = [&](void *dst, const void *src, size_t bytes) {
                return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
            };
            HIPCHK(c, up(h.tap_box, t.box.data(), t.box.size() * sizeof(int)));
            HIPCHK(c, up(h.tap_off, t.off.data(), t.off.size() * sizeof(long long)));
            HIPCHK(c, up(h.tap_fsrc, t.fsrc.data(), t.fsrc.size() * sizeof(float)));
            HIPCHK(c, up(h.tap_frec, t.frec.data(), t.frec.size() * sizeof(float)));
        }
    }
    h.tap_mode = mode;
    h.tap_radius = radius_m;
    h.tap_filt = filtsize;
    h.tap_first = first_shot;
    h.tap_nsrc = nsrc;
    h.tap_sx = sx;
    h.tap_sz = sz;
    c->have_grad = false;
    return RZ_OK;
}

int rz_sync(rz_ctx *c)
{
    if (!c) return RZ_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RZ_OK;
}

int rz_set_profiling(rz_ctx *c, int on)
{
    if (!c) return RZ_ERR_INVALID;
    c->prof = on != 0;
    for (int p = 0; p < 4; ++p) {
        c->prof_ms[p] = 0.0;
        c->prof_launches[p] = 0;
    }
    return RZ_OK;
}

int rz_get_profile(const rz_ctx *c, double *ms, long long *launches)
{
    if (!c) return RZ_ERR_INVALID;
    for (int p = 0; p < 4; ++p) {
        if (ms) ms[p] = c->prof_ms[p];
        if (launches) launches[p] = c->prof_launches[p];
    }
    return RZ_OK;
}

int rz_rows_per_lane(const rz_ctx *c) { return c ? c->R : 0; }
int rz_compute_units(const rz_ctx *c) { return c ? c->ncu : 0; }

int rz_set_split(rz_ctx *c, int on)
{
    if (!c || on < 0 || on > 2) return RZ_ERR_INVALID;
    if (c->split != (on != 0)) drop_graphs(c);
    c->split = on != 0;
    c->split_wait = 0;
    c->split_backoff = 16;
    if (on) c->split_fail = 0;
    if (on == 2) {   // the next split launch finds its failure flag already raised
        const int two = 2;
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipMemcpy(c->ss.err, &two, sizeof(int), hipMemcpyHostToDevice));
    }
    return RZ_OK;
}

int rz_get_split(const rz_ctx *c, int *active, int *failure)
{
    if (!c) return RZ_ERR_INVALID;
    if (active) *active = c->split ? 1 : 0;
    if (failure) *failure = c->split_fail;
    return RZ_OK;
}

int rz_set_split_shots(rz_ctx *c, int n)
{
    if (!c || n < -1) return RZ_ERR_INVALID;
    if (n != c->split_shots) drop_graphs(c);
    c->split_shots = n;
    return RZ_OK;
}

int rz_get_split_shots(const rz_ctx *c) { return c ? c->split_shots : 0; }

int rz_tune_split(rz_ctx *c, int *split_shots, double *ms)
{
    if (!c) return RZ_ERR_INVALID;
    if (const int rc = ready(c, true)) return rc;
    const int T = c->total_shots, F = c->ncu - T;
    std::vector<int> cand;
    if (T > 0 && 2 * T <= c->ncu) cand = {-1, 0};                          // every shot split, or none
    else if (T > 0 && F > 0) cand = {0, F / 4, F / 2, (3 * F) / 4, -1};   // how many of the slowest run split
    const int keep = c->split_shots;
    int best = keep;
    double best_ms = -1.0;
    HIPCHK(c, hipSetDevice(c->device));
    auto run = [&](double *el) {
        const auto t0 = std::chrono::steady_clock::now();
        int rc = rz_gradient(c);
        if (rc == RZ_OK || rc == RZ_ERR_NOT_CONVERGED) rc = rz_sync(c);
        *el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    };
    double el = 0.0;
    if (cand.empty()) {                          // nothing to choose: the time of the setting in use
        const int rc = run(&el);
        if (rc != RZ_OK && rc != RZ_ERR_NOT_CONVERGED) return rc;
        best_ms = el;
    }
    for (int n : cand) {
        rz_set_split_shots(c, n);
        // the first run plans from the previous run's iteration counts; time the second
        for (int rep = 0; rep < 2; ++rep) {
            const int rc = run(&el);
            if (rc != RZ_OK && rc != RZ_ERR_NOT_CONVERGED) {
                rz_set_split_shots(c, keep);
                return rc;
            }
        }
        if (best_ms < 0.0 || el < best_ms) {
            best = n;
            best_ms = el;
        }
    }
    rz_set_split_shots(c, best);
    if (split_shots) *split_shots = best;
    if (ms) *ms = best_ms;
    return RZ_OK;
}

int rz_autotune_rows(int device, int nmodels, const rz_model_desc *models, size_t hbm_budget, int *rows_per_lane,
                     double *ms)
{
    if (!rows_per_lane || nmodels <= 0 || !models) return RZ_ERR_INVALID;
    // One outer iteration of forward and adjoint per candidate, on a smooth
    // synthetic model (1500 + 0.5 z m/s) with picks of 1 s: the cost per
    // iteration is what the rows per lane change.
    std::vector<rz_model_desc> d(models, models + nmodels);
    for (auto &x : d) x.max_iter = 1;
    double best_ms = 0.0;
    int best = 0;
    for (int R : {1, 2, 4, 8}) {
        rz_ctx *t = nullptr;
        if (rz_create(&t, device, nmodels, d.data(), R, hbm_budget) != RZ_OK) continue;
        bool ok = true;
        for (int m = 0; m < nmodels && ok; ++m) {
            std::vector<float> S((size_t)d[m].nx * d[m].ny);
            for (int k = 0; k < d[m].ny; ++k)
                for (int i = 0; i < d[m].nx; ++i)
                    S[(size_t)k * d[m].nx + i] = (float)(1.0 / (1500.0 + 0.5 * (k + 1) * (double)d[m].dh));
            std::vector<float> picks((size_t)std::max(1, d[m].nshots * d[m].ntr), 1.f);
            ok = rz_set_slowness(t, m, S.data(), 0) == RZ_OK && rz_set_picks(t, m, picks.data(), 0) == RZ_OK;
        }
        double el = 0.0;
        for (int rep = 0; rep < 2 && ok; ++rep) {   // the first run warms up
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = rz_gradient(t);
            ok = rc == RZ_OK || rc == RZ_ERR_NOT_CONVERGED;
            el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        rz_destroy(t);
        if (ok && (best == 0 || el < best_ms)) {
            best = R;
            best_ms = el;
        }
    }
    if (best == 0) {
        g_create_err = "no rows-per-lane setting can hold these models";
        return RZ_ERR_INVALID;
    }
    *rows_per_lane = best;
    if (ms) *ms = best_ms;
    return RZ_OK;
}
size_t rz_device_bytes(const rz_ctx *c) { return c ? c->bytes : 0; }
int rz_batch_shots(const rz_ctx *c) { return c ? c->nslots : 0; }
size_t rz_slot_bytes(const rz_ctx *c) { return c ? (size_t)c->slot_floats * 5 * sizeof(float) : 0; }

}  // extern "C"
