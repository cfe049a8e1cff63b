// Gaussian smoothing on the device (gauss_filt.c:10-139): the SPATFILTER = 4
// branch of precond.c:35-37 on the gradient, and fatt.c:179 on the search
// direction.  Below it, the gradient tapers of precond.c:43-53
// (taper_grad.c): elementwise products with factors computed on the host.
//
// The weights are computed on the host exactly as the reference computes
// them (float arguments, libm exp in double, float sum and division).  Each
// output is then one thread's float sum over the window in the reference's
// order, x outer and y inner, of (replicated-edge input) * weight, a product
// and a sum each rounded to float (no contraction): the result is the
// reference's bit for bit.  A thread computes QY outputs stacked along y; for
// a fixed x offset they read a sliding column of QY inputs, one new load per
// y offset.  The weights are wave-uniform (scalar loads).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rajzel_amd.h"
#include "rz_internal.h"

namespace rz {
namespace {

constexpr int kGX = 64;   // threads along x per block
constexpr int kGY = 4;    // threads along y per block
constexpr int kQY = 4;    // outputs per thread along y

__global__ __launch_bounds__(kGX * kGY) void k_gauss_filt(const float *__restrict__ src, float *__restrict__ dst,
                                                          const float *__restrict__ w, int nx, int ny, int hx,
                                                          int hy)
{
    const int i = blockIdx.x * kGX + threadIdx.x;
    const int j0 = (blockIdx.y * kGY + threadIdx.y) * kQY;
    const int wx = 2 * hx + 1;
    float conv[kQY];
#pragma unroll
    for (int q = 0; q < kQY; ++q) conv[q] = 0.0f;
    for (int ii = -hx; ii <= hx; ++ii) {
        const int col = min(max(i + ii, 0), nx - 1);
        const float *cp = src + col;
        float v[kQY];
#pragma unroll
        for (int q = 0; q < kQY; ++q) v[q] = cp[(long long)min(max(j0 + q - hy, 0), ny - 1) * nx];
        const float *wcol = w + (ii + hx);
        for (int jj = -hy; jj <= hy; ++jj) {
            const float wt = wcol[(jj + hy) * wx];
#pragma unroll
            for (int q = 0; q < kQY; ++q) conv[q] += v[q] * wt;
#pragma unroll
            for (int q = 0; q < kQY - 1; ++q) v[q] = v[q + 1];
            v[kQY - 1] = cp[(long long)min(max(j0 + kQY + jj, 0), ny - 1) * nx];
        }
    }
    if (i < nx) {
#pragma unroll
        for (int q = 0; q < kQY; ++q)
            if (j0 + q < ny) dst[(long long)(j0 + q) * nx + i] = conv[q];
    }
}

int fail_global(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_global_error(buf);
    return code;
}

// gauss_filt.c:27-50: a size of 0 leaves the field as it is (returns 0);
// an even size is made odd.  Returns 1 with the half sizes set otherwise.
int normalise_sizes(int *fsx, int *fsy, int &hx, int &hy)
{
    if (*fsx == 0) return 0;
    if (!(*fsx % 2)) *fsx += *fsx > 0 ? 1 : -1;
    if (*fsy == 0) return 0;
    if (!(*fsy % 2)) *fsy += *fsy > 0 ? 1 : -1;
    hx = std::abs(*fsx) / 2;
    hy = std::abs(*fsy) / 2;
    return 1;
}

// gauss_filt.c:44-50, 101-120: sigma = hfs/2 (integer halves), weights
// exp(-ii^2/sx - jj^2/sy) in float, their float sum, and the division by it
// over 1..FILT_SIZE (none for negative sizes).  w holds (2hx+1)*(2hy+1)
// floats, row jj+hy, column ii+hx.
void gauss_weights(int fsx, int fsy, int hx, int hy, float *w)
{
    const float sigx = hx / 2, sigy = hy / 2;
    const float sx = 2.0 * sigx * sigx, sy = 2.0 * sigy * sigy;
    const int wx = 2 * hx + 1;
    float sum = 0.0;
    for (int ii = -hx; ii <= hx; ++ii)
        for (int jj = -hy; jj <= hy; ++jj) {
            float &k = w[(size_t)(jj + hy) * wx + (ii + hx)];
            k = ::exp((double)(-((ii * ii) / sx) - ((jj * jj) / sy)));
            sum += k;
        }
    for (int i = 0; i < fsx; ++i)
        for (int j = 0; j < fsy; ++j) w[(size_t)j * wx + i] /= sum;
}

hipError_t launch_gauss_filt(const float *src, float *dst, const float *w, int nx, int ny, int hx, int hy,
                             hipStream_t st)
{
    const dim3 block(kGX, kGY);
    const dim3 grid((nx + kGX - 1) / kGX, (ny + kGY * kQY - 1) / (kGY * kQY));
    hipLaunchKernelGGL(k_gauss_filt, grid, block, 0, st, src, dst, w, nx, ny, hx, hy);
    return hipGetLastError();
}

int check_grid(const void *field, int nx, int ny, const int *fsx, const int *fsy)
{
    if (!field || !fsx || !fsy) return fail_global(RZ_ERR_INVALID, "gauss_filt: null pointer");
    if (nx < 1 || ny < 1) return fail_global(RZ_ERR_INVALID, "gauss_filt: NX, NY must be >= 1");
    if (ny > 65535 * kGY * kQY || (long long)nx * ny > (1LL << 31))
        return fail_global(RZ_ERR_INVALID, "gauss_filt: grid too large");
    return RZ_OK;
}

// Scratch of rz_gauss_filt_noalloc: the filter weights at [0, kNaWeights)
// and the copy of the input field after them.  Every call on a device uses
// that device's copy of this array; calls that overlap on one device race.
constexpr size_t kNaFloats = RZ_GAUSS_NOALLOC_FLOATS;
constexpr size_t kNaWeights = RZ_GAUSS_NOALLOC_WEIGHTS;
static __device__ __align__(256) float scratch[kNaFloats];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
float *g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on `device` (the current device), looked up
// once per device.
int scratch_on(int device, float **out)
{
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail_global(RZ_ERR_HIP, "gauss_filt_noalloc: scratch address on device %d: %s", device,
                               hipGetErrorString(e));
        g_na_ptr[device] = static_cast<float *>(p);
    }
    *out = g_na_ptr[device];
    return RZ_OK;
}

// taper_grad.c's multiplications (waveconv[j][i] *= taper_coeff[j][i]):
// by a per-column (sws 1), per-row (sws 2) or per-cell (sws 4..6) factor.
__global__ void k_taper_scale(float *f, int nx, int ny, const float *col, const float *row, const float *cell)
{
    const long long n = (long long)nx * ny;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nx), c = (int)(i - (long long)k * nx);
        f[i] *= col ? col[c] : (row ? row[k] : cell[i]);
    }
}

// sws 3: each cell times its disc factors, in source-then-receiver order.
__global__ void k_taper_discs(float *f, long long n, const unsigned *start, const float *fac)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const unsigned a = start[i], b = start[i + 1];
        if (a == b) continue;
        float v = f[i];
        for (unsigned q = a; q < b; ++q) v *= fac[q];
        f[i] = v;
    }
}

int grid_blocks(long long n)
{
    return (int)std::min<long long>(std::max<long long>((n + 255) / 256, 1), 65535);
}

// The disc factors of the last sws-3 geometry, kept on its device.
struct DiscCache {
    int device = -1, nx = 0, ny = 0;
    float dh = 0.f, R = 0.f;
    std::vector<float> cx, cy;
    unsigned *start = nullptr;
    float *fac = nullptr;
};
std::mutex g_dc_mu;
DiscCache g_dc;
constexpr size_t kDiscMaxFactors = (size_t)1 << 28;

// Tables of the sws-3 taper for this geometry on `device` (current), built
// and uploaded unless the cache holds them.  Call with g_dc_mu held.
int disc_tables(int device, int nx, int ny, const rz_grad_taper &t)
{
    std::vector<float> cx(t.nsrc + t.ntr), cy(t.nsrc + t.ntr);
    for (int n = 0; n < t.nsrc; ++n) {
        cx[n] = t.src_x_m[n];
        cy[n] = t.src_z_m[n];
    }
    for (int n = 0; n < t.ntr; ++n) {   // recpos * DH, a float product (taper_grad.c:239)
        cx[t.nsrc + n] = (float)(t.rec_ix[n] + 1) * t.dh;
        cy[t.nsrc + n] = (float)(t.rec_iy[n] + 1) * t.dh;
    }
    if (g_dc.start && g_dc.device == device && g_dc.nx == nx && g_dc.ny == ny && g_dc.dh == t.dh &&
        g_dc.R == t.srtradius && g_dc.cx == cx && g_dc.cy == cy)
        return RZ_OK;
    std::vector<unsigned> start;
    std::vector<float> fac;
    const std::string why = disc_csr(nx, ny, t.dh, t.srtradius, (int)cx.size(), cx.data(), cy.data(),
                                     kDiscMaxFactors, start, fac);
    if (!why.empty()) return fail_global(RZ_ERR_INVALID, "taper_grad: %s", why.c_str());
    if (g_dc.start) {
        (void)hipSetDevice(g_dc.device);
        (void)hipDeviceSynchronize();
        (void)hipFree(g_dc.start);
        (void)hipFree(g_dc.fac);
        g_dc = DiscCache{};
        (void)hipSetDevice(device);
    }
    unsigned *sd = nullptr;
    float *fd = nullptr;
    if (hipMalloc(&sd, start.size() * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&fd, std::max<size_t>(fac.size(), 1) * sizeof(float)) != hipSuccess) {
        if (sd) (void)hipFree(sd);
        return fail_global(RZ_ERR_NOMEM, "taper_grad: device allocation failed");
    }
    if (hipMemcpy(sd, start.data(), start.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
        (!fac.empty() && hipMemcpy(fd, fac.data(), fac.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)) {
        (void)hipFree(sd);
        (void)hipFree(fd);
        return fail_global(RZ_ERR_HIP, "taper_grad: upload failed");
    }
    g_dc.device = device; g_dc.nx = nx; g_dc.ny = ny; g_dc.dh = t.dh; g_dc.R = t.srtradius;
    g_dc.cx = std::move(cx); g_dc.cy = std::move(cy);
    g_dc.start = sd; g_dc.fac = fd;
    return RZ_OK;
}

}  // namespace
}  // namespace rz

extern "C" int rz_taper_grad(int device, void *hip_stream, float *field, int nx, int ny, const rz_grad_taper *t,
                             float *coeff, int on_device)
{
    using namespace rz;
    if (!field || !t) return fail_global(RZ_ERR_INVALID, "taper_grad: null pointer");
    if (nx < 1 || ny < 1 || (long long)nx * ny > (1LL << 31))
        return fail_global(RZ_ERR_INVALID, "taper_grad: NX, NY must be >= 1 with NX*NY <= 2^31");
    const int sws = t->sws;
    if (sws < 1 || sws > 6) return fail_global(RZ_ERR_INVALID, "taper_grad: sws %d (1..6)", sws);
    if (sws == 3) {
        if (!(t->srtradius > 0.f) || !std::isfinite(t->srtradius) || !(t->dh > 0.f))
            return fail_global(RZ_ERR_INVALID, "taper_grad: SRTRADIUS and DH must be finite and > 0");
        if (t->nsrc < 0 || t->ntr < 0 || (t->nsrc > 0 && (!t->src_x_m || !t->src_z_m)) ||
            (t->ntr > 0 && (!t->rec_ix || !t->rec_iy)))
            return fail_global(RZ_ERR_INVALID, "taper_grad: source or receiver positions missing");
    }
    if (sws >= 4 && !t->file) return fail_global(RZ_ERR_INVALID, "taper_grad: no taper file contents");
    const size_t n = (size_t)nx * ny;
    // host factors (sws 1, 2, 4..6)
    std::vector<float> col, row, cell;
    if (sws == 1) col = grad_taper_columns(nx, t->gradt1, t->gradt2, t->gradt3, t->gradt4);
    if (sws == 2) row = grad_taper_rows(ny, t->gradt1, t->gradt2, t->dh, t->exp_taper_grad_hor);
    if (sws >= 4) {   // taper.bin is read x outer, y inner (taper_grad.c:275-282)
        cell.resize(n);
        for (int c = 0; c < nx; ++c)
            for (int k = 0; k < ny; ++k) cell[(size_t)k * nx + c] = t->file[(size_t)c * ny + k];
    }
    if (coeff && sws != 3)
        for (int k = 0; k < ny; ++k)
            for (int c = 0; c < nx; ++c)
                coeff[(size_t)k * nx + c] = sws == 1 ? col[c] : (sws == 2 ? row[k] : cell[(size_t)k * nx + c]);

    hipStream_t st = (hipStream_t)hip_stream;
    if (hipSetDevice(device) != hipSuccess) return fail_global(RZ_ERR_HIP, "taper_grad: no device %d", device);
    std::unique_lock<std::mutex> lock(g_dc_mu, std::defer_lock);
    if (sws == 3) {
        lock.lock();
        if (const int rc = disc_tables(device, nx, ny, *t)) return rc;
    }
    float *fd = nullptr, *cd = nullptr;
    auto cleanup = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        else (void)hipDeviceSynchronize();
        if (fd && !on_device) (void)hipFree(fd);
        if (cd) (void)hipFree(cd);
        return rc;
    };
    const size_t ncoef = sws == 1 ? (size_t)nx : (sws == 2 ? (size_t)ny : (sws == 3 ? 0 : n));
    if ((!on_device && hipMalloc(&fd, n * sizeof(float)) != hipSuccess) ||
        (ncoef && hipMalloc(&cd, ncoef * sizeof(float)) != hipSuccess))
        return cleanup(fail_global(RZ_ERR_NOMEM, "taper_grad: device allocation failed"));
    if (on_device) fd = field;
    hipError_t e = on_device ? hipSuccess
                             : hipMemcpyAsync(fd, field, n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && ncoef) {
        const float *src = sws == 1 ? col.data() : (sws == 2 ? row.data() : cell.data());
        e = hipMemcpyAsync(cd, src, ncoef * sizeof(float), hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) {
        if (sws == 3)
            hipLaunchKernelGGL(k_taper_discs, dim3(grid_blocks((long long)n)), dim3(256), 0, st, fd, (long long)n,
                               g_dc.start, g_dc.fac);
        else
            hipLaunchKernelGGL(k_taper_scale, dim3(grid_blocks((long long)n)), dim3(256), 0, st, fd, nx, ny,
                               sws == 1 ? cd : nullptr, sws == 2 ? cd : nullptr, sws >= 4 ? cd : nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess && !on_device) e = hipMemcpyAsync(field, fd, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = st ? hipStreamSynchronize(st) : hipDeviceSynchronize();
    if (e != hipSuccess) return cleanup(fail_global(RZ_ERR_HIP, "taper_grad: %s", hipGetErrorString(e)));
    return cleanup(RZ_OK);
}

extern "C" int rz_gauss_filt(int device, void *hip_stream, float *field, int nx, int ny, int *fsx, int *fsy,
                             int on_device)
{
    using namespace rz;
    if (const int rc = check_grid(field, nx, ny, fsx, fsy)) return rc;
    int hx = 0, hy = 0;
    if (!normalise_sizes(fsx, fsy, hx, hy)) return RZ_OK;
    if (hx > 4096 || hy > 4096) return fail_global(RZ_ERR_INVALID, "gauss_filt: filter size above 8193");
    const int wx = 2 * hx + 1, wy = 2 * hy + 1;
    std::vector<float> w((size_t)wx * wy);
    gauss_weights(*fsx, *fsy, hx, hy, w.data());

    hipStream_t st = (hipStream_t)hip_stream;
    if (hipSetDevice(device) != hipSuccess) return fail_global(RZ_ERR_HIP, "gauss_filt: no device %d", device);
    const size_t bytes = (size_t)nx * ny * sizeof(float);
    float *wd = nullptr, *src = nullptr, *dst = nullptr;
    auto cleanup = [&](int rc) {
        if (st) (void)hipStreamSynchronize(st);
        else (void)hipDeviceSynchronize();
        if (wd) (void)hipFree(wd);
        if (src) (void)hipFree(src);
        if (dst && !on_device) (void)hipFree(dst);
        return rc;
    };
    if (hipMalloc(&wd, w.size() * sizeof(float)) != hipSuccess || hipMalloc(&src, bytes) != hipSuccess ||
        (!on_device && hipMalloc(&dst, bytes) != hipSuccess))
        return cleanup(fail_global(RZ_ERR_NOMEM, "gauss_filt: device allocation failed"));
    if (on_device) dst = field;
    hipError_t e = hipMemcpyAsync(wd, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(src, field, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_gauss_filt(src, dst, wd, nx, ny, hx, hy, st);
    if (e == hipSuccess && !on_device) e = hipMemcpyAsync(field, dst, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = st ? hipStreamSynchronize(st) : hipDeviceSynchronize();
    if (e != hipSuccess) return cleanup(fail_global(RZ_ERR_HIP, "gauss_filt: %s", hipGetErrorString(e)));
    return cleanup(RZ_OK);
}

extern "C" int rz_gauss_filt_noalloc(int device, void *hip_stream, float *field, int nx, int ny, int *fsx,
                                     int *fsy)
{
    using namespace rz;
    if (const int rc = check_grid(field, nx, ny, fsx, fsy)) return rc;
    if (device < 0 || device >= kNaMaxDevices)
        return fail_global(RZ_ERR_INVALID, "gauss_filt_noalloc: device %d out of range", device);
    // checked before the sizes are touched: a rejected call changes nothing
    const size_t ncell = (size_t)nx * ny;
    if (ncell > kNaFloats - kNaWeights)
        return fail_global(RZ_ERR_INVALID, "gauss_filt_noalloc: %d x %d grid above the %zu-cell scratch", nx, ny,
                           (size_t)(kNaFloats - kNaWeights));
    int tx = *fsx, ty = *fsy, hx = 0, hy = 0;
    if (!normalise_sizes(&tx, &ty, hx, hy)) {
        *fsx = tx;
        *fsy = ty;
        return RZ_OK;
    }
    const size_t nw = (size_t)(2 * hx + 1) * (2 * hy + 1);
    if (hx > 4096 || hy > 4096 || nw > kNaWeights)
        return fail_global(RZ_ERR_INVALID, "gauss_filt_noalloc: a %d x %d window exceeds %zu weights", std::abs(tx),
                           std::abs(ty), (size_t)kNaWeights);
    *fsx = tx;
    *fsy = ty;
    float w[kNaWeights];
    gauss_weights(tx, ty, hx, hy, w);

    hipStream_t st = (hipStream_t)hip_stream;
    if (hipSetDevice(device) != hipSuccess) return fail_global(RZ_ERR_HIP, "gauss_filt_noalloc: no device %d", device);
    float *sp = nullptr;
    if (const int rc = scratch_on(device, &sp)) return rc;
    float *src = sp + kNaWeights;
    hipError_t e = hipMemcpyAsync(sp, w, nw * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(src, field, ncell * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = launch_gauss_filt(src, field, sp, nx, ny, hx, hy, st);
    if (e == hipSuccess) e = st ? hipStreamSynchronize(st) : hipDeviceSynchronize();
    if (e != hipSuccess) return fail_global(RZ_ERR_HIP, "gauss_filt_noalloc: %s", hipGetErrorString(e));
    return RZ_OK;
}
