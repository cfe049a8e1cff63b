// Host tables of the gradient tapers: the per-shot ones of ass_grad.c:31-37
// (SWS_TAPER_CIRCULAR_PER_SHOT), which k_gradient applies on the device, and
// those of taper_grad.c (precond.c:43-53), which rz_taper_grad applies.
//
// Mode 1 (taper_grad_shot.c:43-84) multiplies each shot's gradient by a
// radial cosine taper around its source and then around every receiver in
// turn.  The factors depend only on the geometry, so they are computed once
// here with the C library's sqrt, pow and cos in the reference's mix of
// float and double, and kept as
//   - one dense window per source (the bounding box of the cells whose
//     factor is not exactly 1), and
//   - per cell, the factors of the receiver discs that hold it, in receiver
//     order (a CSR over row-major cells), so that the device multiplies in
//     the reference's order.
// A factor of exactly 1 leaves a float unchanged under multiplication, so
// cells outside every disc carry nothing.
//
// Mode 2 (taper_grad_shot1.c:32-62) sets the (2F+1) x (2F+1) window around
// every source (iround(pos/DH), fd.h:18) and every receiver to 1e-5.  The
// reference writes these windows without a bounds check through the row
// pointers of a contiguous matrix (util.c:177-204): a column outside 1..NX
// lands on the neighbouring row (linear index (j-1)*NX + i-1), and this is
// reproduced; rows outside 1..NY and linear indices before the first or past
// the last cell are writes outside the matrix (undefined behaviour in the
// reference) and are left out.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rz_internal.h"

namespace rz {

namespace {

constexpr double kPi = 3.141592653589793;   // fd.h:23

// taper_grad_shot.c:100-110 at cell (ix, iy) (1-based) for a disc centred at
// (cx, cy) metres: x, y and the differences in float, r from the double
// squares and sqrt stored as float, the taper evaluated in double and stored
// as float.
float disc_factor(int ix, int iy, float dh, float cx, float cy, float R, float a)
{
    const float f0 = 1e-3f;
    const float x = (float)ix * dh, y = (float)iy * dh;
    const float r = (float)std::sqrt(std::pow((double)(x - cx), 2.0) + std::pow((double)(y - cy), 2.0));
    float taper = 1.0f;
    if (r <= R) taper = (float)(f0 + (1.0 - f0) * std::pow(0.5 - 0.5 * std::cos(kPi * r / R), (double)a));
    return taper;
}

// Cells (1-based, clipped to the grid) that can lie within R of (cx, cy):
// the disc's bounding box widened by one cell against rounding.
bool disc_box(int nx, int ny, float dh, float cx, float cy, float R, int &i0, int &i1, int &j0, int &j1)
{
    auto lo = [&](double v, int n) { return (int)std::fmax(1.0, std::fmin((double)n + 1, std::floor(v / dh) - 1)); };
    auto hi = [&](double v, int n) { return (int)std::fmax(0.0, std::fmin((double)n, std::ceil(v / dh) + 1)); };
    i0 = lo((double)cx - R, nx); i1 = hi((double)cx + R, nx);
    j0 = lo((double)cy - R, ny); j1 = hi((double)cy + R, ny);
    return i0 <= i1 && j0 <= j1;
}

}  // namespace

float taper_exponent()
{
    const float frh = 0.1f, f0 = 1e-3f;
    // taper_grad_shot.c:38, taper_grad.c:204: from float frh - f0, in double, stored as float
    return (float)(-(std::log(frh - f0) - std::log(1.0 - f0)) / std::log(2.0));
}

std::string disc_csr(int nx, int ny, float dh, float R, int n, const float *cx, const float *cy, size_t max_entries,
                     std::vector<unsigned> &start, std::vector<float> &fac_out)
{
    const float a = taper_exponent();
    const size_t ncell = (size_t)nx * ny;
    start.assign(ncell + 1, 0u);
    std::vector<unsigned> cell;
    std::vector<float> fac;
    for (int q = 0; q < n; ++q) {
        int i0, i1, j0, j1;
        if (!disc_box(nx, ny, dh, cx[q], cy[q], R, i0, i1, j0, j1)) continue;
        for (int iy = j0; iy <= j1; ++iy)
            for (int ix = i0; ix <= i1; ++ix) {
                const float f = disc_factor(ix, iy, dh, cx[q], cy[q], R, a);
                if (f == 1.f) continue;
                if (cell.size() >= max_entries)
                    return "the discs cover too many cells (more than " + std::to_string(max_entries) +
                           " factors); reduce the radius";
                cell.push_back((unsigned)((size_t)(iy - 1) * nx + (ix - 1)));
                fac.push_back(f);
            }
    }
    // a stable bucket by cell keeps each cell's factors in disc order
    for (unsigned q : cell) ++start[q + 1];
    for (size_t i = 0; i < ncell; ++i) start[i + 1] += start[i];
    fac_out.resize(fac.size());
    std::vector<unsigned> fill(start.begin(), start.end() - 1);
    for (size_t e = 0; e < cell.size(); ++e) fac_out[fill[cell[e]]++] = fac[e];
    return std::string();
}

// taper_grad.c:62-75 and 137-146: a value of the Gaussian window, a = 3
// (float), exp(-(1/2) * A * A) in double with A = (a * x) / (L/2.0) (a float
// product divided in double), stored as float; x is i - L, or
// i - (ifw - L2) with L = L2 for the far side of the vertical taper.
static float gauss_window(int x, int L)
{
    const float a = 3.f;
    const double A = (double)(a * (float)x) / (L / 2.0);
    return (float)std::exp(-(1.0 / 2.0) * A * A);
}

std::vector<float> grad_taper_columns(int nx, int g1, int g2, int g3, int g4)
{
    // taper_grad.c:44-102: window[1..ifw], then column i (1-based) in
    // (GRADT1, GRADT4) takes window[h], h counting those columns from 1
    const int L = g2 - g1, L2 = g4 - g3, ifw = g4 - g1;
    std::vector<float> col(nx, 0.f);
    for (int i = 1; i <= nx; ++i) {
        if (!(i > g1 && i < g4)) continue;
        const int h = i - std::max(g1, 0);
        float w = 1.0f;
        if (h <= L) w = gauss_window(h - L, L);
        if (h >= ifw - L2) w = gauss_window(h - (ifw - L2), L2);
        col[i - 1] = w;
    }
    return col;
}

std::vector<float> grad_taper_rows(int ny, int g1, int g2, float dh, float expo)
{
    // taper_grad.c:118-187: rows GRADT1..GRADT2 take window[h] (h counting
    // rows from GRADT1, or from 1), deeper rows pow((float)(j*DH), EXP) in
    // double stored as float, shallower rows 0
    const int L = g2 - g1;
    std::vector<float> row(ny, 0.f);
    for (int j = 1; j <= ny; ++j) {
        float g = 0.f;
        if (j >= g1 && j <= g2) {
            const int h = 1 + j - std::max(g1, 1);
            g = h <= L ? gauss_window(h - L, L) : 1.0f;
        }
        if (j > g2) g = (float)std::pow((double)(float)((float)j * dh), (double)expo);
        row[j - 1] = g;
    }
    return row;
}

std::string build_taper(int mode, int nx, int ny, float dh, float radius, int filtsize, int nsrc,
                        const float *src_x_m, const float *src_z_m, int ntr, const int *rec_ix, const int *rec_iy,
                        size_t max_entries, TaperTables &t)
{
    t = TaperTables{};
    const size_t ncell = (size_t)nx * ny;
    t.start.assign(ncell + 1, 0u);
    if (mode == 1) {
        const float a = taper_exponent();
        const float R = radius;
        // source windows
        t.box.assign(4 * (size_t)nsrc, 0);
        t.off.assign((size_t)nsrc, 0);
        std::vector<float> win;
        for (int n = 0; n < nsrc; ++n) {
            int *bx = &t.box[4 * (size_t)n];
            bx[0] = 1; bx[1] = 0; bx[2] = 1; bx[3] = 0;   // empty
            t.off[n] = (long long)t.fsrc.size();
            int i0, i1, j0, j1;
            if (!disc_box(nx, ny, dh, src_x_m[n], src_z_m[n], R, i0, i1, j0, j1)) continue;
            const int w = i1 - i0 + 1, h = j1 - j0 + 1;
            win.assign((size_t)w * h, 1.f);
            int ki0 = ny, ki1 = -1, ci0 = nx, ci1 = -1;   // tight box of factors != 1 (0-based)
            for (int iy = j0; iy <= j1; ++iy)
                for (int ix = i0; ix <= i1; ++ix) {
                    const float f = disc_factor(ix, iy, dh, src_x_m[n], src_z_m[n], R, a);
                    win[(size_t)(iy - j0) * w + (ix - i0)] = f;
                    if (f != 1.f) {
                        ki0 = std::min(ki0, iy - 1); ki1 = std::max(ki1, iy - 1);
                        ci0 = std::min(ci0, ix - 1); ci1 = std::max(ci1, ix - 1);
                    }
                }
            if (ki1 < 0) continue;
            if (t.fsrc.size() + (size_t)(ki1 - ki0 + 1) * (ci1 - ci0 + 1) > max_entries)
                return "the source discs cover too many cells (more than " + std::to_string(max_entries) +
                       " factors); reduce the radius";
            bx[0] = ki0; bx[1] = ki1; bx[2] = ci0; bx[3] = ci1;
            for (int k = ki0; k <= ki1; ++k)
                for (int c = ci0; c <= ci1; ++c) t.fsrc.push_back(win[(size_t)(k + 1 - j0) * w + (c + 1 - i0)]);
        }
        // receiver discs, in receiver order
        std::vector<float> cx(ntr), cy(ntr);
        for (int n = 0; n < ntr; ++n) {
            cx[n] = (float)(rec_ix[n] + 1) * dh;   // recpos * DH, a float product
            cy[n] = (float)(rec_iy[n] + 1) * dh;
        }
        const std::string why = disc_csr(nx, ny, dh, R, ntr, cx.data(), cy.data(), max_entries, t.start, t.frec);
        if (!why.empty()) return why;
    } else if (mode == 2) {
        // windows as linear ranges of the contiguous matrix, one per row:
        // a difference array over the cells
        const int F = filtsize;
        std::vector<int> diff(ncell + 1, 0);
        auto window = [&](long long i1, long long j1) {
            if (F < 0) return;
            for (long long j = std::max(j1 - F, 1LL); j <= std::min(j1 + F, (long long)ny); ++j) {
                long long lo = (j - 1) * nx + (i1 - F - 1), hi = (j - 1) * nx + (i1 + F - 1);
                lo = std::max(lo, 0LL);
                hi = std::min(hi, (long long)ncell - 1);
                if (lo > hi) continue;
                ++diff[lo];
                --diff[hi + 1];
            }
        };
        for (int n = 0; n < nsrc; ++n)   // iround(srcpos/DH), fd.h:18: float quotient, + 0.5 in double
            window((long long)std::floor((double)(src_x_m[n] / dh) + 0.5),
                   (long long)std::floor((double)(src_z_m[n] / dh) + 0.5));
        for (int n = 0; n < ntr; ++n) window(rec_ix[n] + 1, rec_iy[n] + 1);
        long long run = 0;
        for (size_t i = 0; i < ncell; ++i) {
            run += diff[i];
            t.start[i + 1] = t.start[i] + (run > 0 ? 1u : 0u);
        }
    }
    return std::string();
}

}  // namespace rz
