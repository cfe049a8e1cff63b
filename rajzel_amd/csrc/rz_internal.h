// Internal declarations shared by the HIP kernels and the C-ABI layer.
//
// Band-skewed field layout (one per shot field, and for the slowness):
//   rows are split into bands of 64*R rows; band b holds rows [64Rb, 64R(b+1)).
//   Inside a band, lane l (0..63) owns R consecutive rows 64Rb + lR + r.
//   Family A stores cell (k, c) of lane l at key  c + l,
//   family B stores it at key                     c - l + 63,
//   address = b*bstride + key*64R + l*R + r,  bstride = (NX+64)*64R floats
//   (keys 0..NX+62 hold cells; key NX+63 is a dummy block that absorbs the
//   stores of lanes outside the grid).
// A sweep along (x+,y+) or (x-,y-) visits, at each step, the cells of one
// A-key (one anti-diagonal of the band), a sweep along (x+,y-) or (x-,y+)
// those of one B-key; each step of a wavefront therefore reads and writes
// 64*R contiguous floats.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace rz {

constexpr int kWave = 64;
constexpr int kPF = 8;                // steps between barriers (an epoch); prefetch depth at R <= 4
constexpr int kLag = kPF;             // extra lag between consecutive bands (>= kPF-1)
constexpr int kBandOff = kWave + kLag;// step offset between consecutive bands
constexpr int kRingQ = 32;            // LDS ring slots per band (>= kLag+1+kPF, power of 2)
// A workgroup runs two sweeps of one shot at a time (two wave groups, one
// wave per band each): the second trails the first by pair_delay() epochs.
constexpr int kGroups = 2;
constexpr int kMaxBands = 8;          // bands (waves) per group
// Bands a group may hold at R rows per lane: 8 for R <= 2, 4 above (with
// two groups: 16 resp. 8 waves, a budget of 128 resp. 256 VGPRs per lane).
constexpr int max_bands_for(int R) { return R <= 2 ? 8 : 4; }
// Taller grids at R = 8 (5 to 8 bands: 2049 to 4096 rows) run unpaired: all
// 8 waves form one group, one wave per band, and the four sweeps of an
// iteration run one after the other.
constexpr int max_bands_unpaired(int R) { return R == 8 ? 8 : max_bands_for(R); }
constexpr bool paired_for(int R, int nb) { return nb <= max_bands_for(R); }
// Prefetch slots of a sweep: a step's blocks are loaded this many steps
// before it (one less in the adjoint sweep).  At R = 8 a step holds twice
// the rows of one at R = 4, so four slots keep the bytes in flight per wave
// and about the same lead in time, and the sweep within 256 VGPRs (eight
// slots spilled 32 to 60 registers to scratch).
constexpr int pf_slots(int R) { return R >= 8 ? 4 : kPF; }
static_assert(kPF % pf_slots(8) == 0 && kPF % pf_slots(4) == 0, "an epoch holds whole rounds of slots");

// INF of the reference (fd.h:29-31): an int compared against float fields.
constexpr float kInf = 9999999.0f;

struct ModelDev {
    int nx, ny, nb, R;
    float dh, time, ttnorm, eps_adj;
    double two_dh2;                // 2*DH^2 (exact in double)
    double rdh;                    // RN64(1/DH)
    long long bstride;             // floats per band
    long long fstride;             // floats per field (nb * bstride)
    const float *SA, *SB;          // slowness in layouts A and B
    const float *S;                // slowness row-major (ny, nx)
    int ntr, nuq;
    const int *rec_k, *rec_c;      // receivers, 0-based
    const int *uq_k, *uq_c, *uq_r; // interior receiver cells and their winning receiver
    int max_iter;
    int trans;                     // stored transposed: rows here are the grid's columns (x), columns its rows
    // T-only quotients of the adjoint as fma(x, rdh_hi, x * rdh_lo) in float
    // (adjoint variant 2); slowness below s_min_ft turns that variant off.
    float rdh_hi, rdh_lo;
    float s_min_ft;
    int stop_mode;                 // 0; tests: 1 every nonzero stop takes the ordered sum, 2 the count first
    // Stopping test (norm1.c:17-31, eikonal.c:44, eikonal_adj.c:37): the
    // reference continues iff its float sum of |T - Told| is >= stop_s.  A
    // block's double sum G of per-sweep decreases brackets that float sum
    // by [G*band_lo, G*band_hi]; when the bracket straddles stop_s the sum
    // is taken in the reference's order (ordered_norm1).
    double stop_s;
    double band_lo, band_hi;
};

struct ShotDev {
    int model;          // index into models
    int shot;           // shot index within the model
    int src_k, src_c;   // 0-based source cell
    int slot;           // field slot
    int pad_;
};

// Per-shot fields.  A-family outputs alternate between two buffers from one
// outer iteration to the next, so that the field at the start of an
// iteration and the one at its end are both intact at the stopping test:
// forward T in TA / LA (LA is free until the adjoint starts), adjoint
// lambda in LC / LA (LA holds the initial lambda).
struct Slots {
    float *TA, *TB, *LA, *LB, *LC; // base of slot 0; slot s at + s*slot_stride
    long long slot_stride;         // elements
};

// Per-shot outputs.  tmod/tnb are (nshots, ntr) blocks packed per model at
// rec_off[model]; it_f/it_a are indexed shot_base[model] + shot.
struct ShotOut {
    float *tmod;     // T at each receiver (calc_FA.c:15-17)
    float *tnb;      // T at the receiver's slope neighbour row (calc_FA_res.c:52-69)
    int *it_f, *it_a;
    int *lsel;       // where the converged lambda is: 0 LA, 1 LC (indexed as it_a)
    int *ord_f, *ord_a;  // passes of the stopping test over the fields (indexed as it_f): the
                         // counts of changed cells in bits 0-15, the ordered sums above
    float *osum_f, *osum_a;  // norm1's ordered float sum of the last decision, -1 if not taken
    int *phases_f;   // sweep phases the forward ran (a pair of sweeps or a lone sweep each;
                     // indexed as it_f): its time, in sweeps, for planning the next launch
};

// Coupling of the two workgroups of split shots (rz_kernels.hip): per shot
// kSyncInts ints of counters and kSyncPart doubles of partial sums; err
// becomes nonzero when a launch cannot run split (1: the two workgroups of
// a shot were on different XCDs, 2: a wait timed out).
constexpr int kSyncInts = 64;
// doubles per split shot: partial sums, counts, then the forward's flags of
// sweeps that changed no cell (two per pair phase)
constexpr int kPartFixed = 4 * kMaxBands;
constexpr int kSyncPart = kPartFixed + 4;
struct SplitSync {
    int *cnt;
    double *part;
    int *err;
};

// Launchers (rz_kernels.hip).  All return hipError_t of the launch.
hipError_t launch_prepare_slowness(const ModelDev *md_host, int model, const ModelDev *md_dev, float *SA,
                                   float *SB, int *err_flag, hipStream_t st);
// paired: two sweeps at a time (two wave groups); else one group of waves
// runs the four sweeps in turn (R = 8 only; tall and transposed grids)
hipError_t launch_forward(int R, int nb_max, bool paired, const ModelDev *md, const ShotDev *shots, int nshots,
                          Slots slots, const long long *rec_off, const int *shot_base, ShotOut out, hipStream_t st);
hipError_t launch_adjoint(int R, int nb_max, bool paired, const ModelDev *md, const ShotDev *shots, int nshots,
                          Slots slots, const long long *rec_off, const int *shot_base, const float *tres, ShotOut out,
                          int variant, hipStream_t st);
// out (cols x rows) = the transpose of in (rows x cols), both row-major
hipError_t launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t st);
// adjoint variant bits: 0 DH division by a double multiply, 1 INF guard,
// 2 (with bit 0) T-only quotients by the float FMA of ModelDev::rdh_hi/lo
constexpr int kAdjFastDiv = 1, kAdjInfGuard = 2, kAdjFloatT = 4;
hipError_t launch_forward_split(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots,
                                Slots slots, const long long *rec_off, const int *shot_base, ShotOut out,
                                SplitSync ss, hipStream_t st);
hipError_t launch_adjoint_split(int R, int nb_max, const ModelDev *md, const ShotDev *shots, int nshots,
                                Slots slots, const long long *rec_off, const int *shot_base, const float *tres,
                                ShotOut out, int variant, SplitSync ss, hipStream_t st);
hipError_t launch_residual(const ModelDev *md, int model, int ntr, int s0, int s1, long long rec_off,
                           const float *tmod, const float *tnb, const float *tobs, const float *direct,
                           float *carry, float *tres, double *l2_shot, double *l2_sum, hipStream_t st);
hipError_t launch_eval_init(float *carry, const float *carry0, int ntr_total, double *l2_sum, int nmodels,
                            float *grad, long long ncells, hipStream_t st);
// Per-shot gradient taper of ass_grad.c:31-37, applied to each shot's
// frame-copied lambda*S^3 before it joins the sum.
//   mode 1  taper_grad_shot.c:43-84: the value is multiplied by the factor of
//           the shot's source disc, then by the factor of every receiver
//           disc that holds the cell, in receiver order (float products in
//           the reference's order; factors of exactly 1 are left out).
//   mode 2  taper_grad_shot1.c:32-62: cells of any source or receiver window
//           take 1e-5.
// Factors are computed on the host with the C library's pow and cos, as the
// reference computes them.
struct TaperDev {
    int mode;                  // 0 none, 1, 2
    int src0;                  // source index of the launch's first shot
    const int *box;            // mode 1, per source: k0, k1, c0, c1 of its window (k0 > k1: empty)
    const long long *off;      // mode 1, per source: offset of its window in fsrc
    const float *fsrc;         // mode 1: window factors, row-major per window
    const unsigned *start;     // per row-major cell, nx*ny + 1 entries: [start[i], start[i+1])
                               // indexes frec (mode 1); non-empty marks the cell (mode 2)
    const float *frec;         // mode 1: receiver factors of each cell, in receiver order
};
hipError_t launch_gradient(const ModelDev *md_host, int model, const ModelDev *md_dev, const int *slot_of,
                           const int *lsel, int nsh, Slots slots, float *grad, int accumulate, const TaperDev &tp,
                           hipStream_t st);
// cp_grad_frame.c:10-53 on a row-major NY x NX field on the device (the
// copy grad_obj.c:70 applies to the shot sum).
hipError_t launch_grad_frame(int nx, int ny, float *grad, hipStream_t st);
// Layout-A field -> row-major; with sel != nullptr the field is `alt` where
// *sel is nonzero (a lambda that ended in LC).
hipError_t launch_field_to_rowmajor(const ModelDev *md_host, const float *field, const float *alt, const int *sel,
                                    float *out, int strip, hipStream_t st);
hipError_t launch_rowmajor_to_field(const ModelDev *md_host, const float *in, float *field, int family,
                                    hipStream_t st);

// Host tables of the per-shot taper (rz_taper.cpp), the contents of a
// TaperDev.  build_taper returns an empty string, or why it cannot build
// them (more than max_entries factors).
struct TaperTables {
    std::vector<int> box;         // mode 1: k0, k1, c0, c1 per source
    std::vector<long long> off;   // mode 1: per source
    std::vector<float> fsrc;      // mode 1
    std::vector<unsigned> start;  // nx*ny + 1
    std::vector<float> frec;      // mode 1
};
std::string build_taper(int mode, int nx, int ny, float dh, float radius, int filtsize, int nsrc,
                        const float *src_x_m, const float *src_z_m, int ntr, const int *rec_ix, const int *rec_iy,
                        size_t max_entries, TaperTables &t);
// The product of the radial tapers of taper_grad_shot.c:100-110 around n
// centres (cx, cy) in metres, in centre order: per row-major cell the
// factors other than 1, [start[i], start[i+1]) of fac.  An empty string, or
// why not (more than max_entries factors).
std::string disc_csr(int nx, int ny, float dh, float R, int n, const float *cx, const float *cy, size_t max_entries,
                     std::vector<unsigned> &start, std::vector<float> &fac);
// taper_grad.c's coefficients: per column of the vertical taper (sws 1) and
// per row of the horizontal one (sws 2).
std::vector<float> grad_taper_columns(int nx, int g1, int g2, int g3, int g4);
std::vector<float> grad_taper_rows(int ny, int g1, int g2, float dh, float expo);

// Message of the last failed call that has no context (rz_last_error(NULL)).
void set_global_error(const char *msg);

}  // namespace rz
