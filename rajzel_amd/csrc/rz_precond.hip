// Gaussian smoothing on the device (gauss_filt.c:10-139): the SPATFILTER = 4
// branch of precond.c:35-37 on the gradient, and fatt.c:179 on the search
// direction.
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

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <mutex>
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

}  // namespace
}  // namespace rz

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
