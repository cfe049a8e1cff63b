// The other gradient smoothings of precond.c:25-33 on the device:
//
//  - smooth_grad.c:10-99 (SPATFILTER = 3): each cell becomes the median of
//    the FILT_SIZE_GRAD x FILT_SIZE_GRAD window around it, the grid's edges
//    replicated (smooth_grad.c:49-72).  The median is median2d.c's: the
//    window read row by row (y offset outer, x offset inner), then k/2 + 1
//    rounds of "take the first smallest value, overwrite it with the
//    window's first largest" (minpos and vecmax with strict comparisons).
//    One thread per cell runs those same rounds, so equal values (+0 and
//    -0 among them) resolve to the same element and NaN compares as there.
//
//  - smooth2.c:9-201 (SPATFILTER = 2): on the gradient v[x][z] (the file
//    order of smooth2.c:58-63, with the zero row and column 0 of its
//    matrix(0, n2, 0, n1)), a tridiagonal solve along x for every z in
//    0..n1-1 over x = 0..n2-1, then one along z over z = 1..n1-1 for every x
//    in 0..n2-1 (smooth2.c:125-152).  With the window function 1 everywhere
//    (win = [0, n1, 0, n2], smooth2.c:49-53,100) every line of a pass has the
//    same matrix, so tripd's decomposition (util.c:666-670) is the same
//    sequence of float operations for each: it is done once on the host and
//    every line runs only tripd's two substitutions (util.c:673-676), one
//    thread per line, with the reference's float operations in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/rajzel_amd.h"
#include "rz_internal.h"

namespace rz {
namespace {

int smooth_fail(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_global_error(buf);
    return code;
}

constexpr int kMedThreads = 256;

// One cell per thread, grid-stride; `work` holds k floats per thread.
__global__ __launch_bounds__(kMedThreads) void k_median(const float *__restrict__ src, float *__restrict__ dst,
                                                        int nx, int ny, int hfs, float *__restrict__ work)
{
    const int fs = 2 * hfs + 1, k = fs * fs;
    const long long nthr = (long long)gridDim.x * blockDim.x, tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float *t = work + tid * k;
    for (long long cell = tid; cell < (long long)nx * ny; cell += nthr) {
        const int j = (int)(cell / nx), i = (int)(cell - (long long)j * nx);
        // the window, row (y offset) outer and column (x offset) inner, edges replicated
        int n = 0;
        for (int jj = -hfs; jj <= hfs; ++jj) {
            const long long row = (long long)min(max(j + jj, 0), ny - 1) * nx;
            for (int ii = -hfs; ii <= hfs; ++ii) t[n++] = src[row + min(max(i + ii, 0), nx - 1)];
        }
        // vecmax: the first largest (median2d.c:29, vecmax)
        float maxglob = t[0];
        for (int q = 1; q < k; ++q)
            if (t[q] > maxglob) maxglob = t[q];
        // rounds 1 .. (k+1)/2 of the selection (median2d.c:31-36, minpos)
        float med = 0.0f;
        for (int r = 0; r < (k + 1) / 2; ++r) {
            float mini = t[0];
            int minp = 0;
            for (int q = 1; q < k; ++q)
                if (t[q] < mini) {
                    mini = t[q];
                    minp = q;
                }
            med = t[minp];
            t[minp] = maxglob;
        }
        dst[cell] = med;
    }
}

// Pass 1 of smooth2 (smooth2.c:125-137): line iz of v (x-slow, z-fast, row
// stride ld) over x = 0 .. n-1 with the decomposed matrix (d, e).
__global__ __launch_bounds__(256) void k_tri_x(float *__restrict__ v, long long ld, int n, int nlines,
                                               const float *__restrict__ d, const float *__restrict__ e)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= nlines) return;
    float *b = v + iz;
    for (int k = 1; k < n; ++k) b[k * ld] = b[k * ld] - e[k - 1] * b[(k - 1) * ld];
    b[(long long)(n - 1) * ld] = b[(long long)(n - 1) * ld] / d[n - 1];
    for (int k = n - 1; k > 0; --k) b[(k - 1) * ld] = b[(k - 1) * ld] / d[k - 1] - e[k - 1] * b[k * ld];
}

// Pass 2 (smooth2.c:138-152): line ix over z = 1 .. n, its first
// right-hand side plus r1 * w * v[ix][0] (w = 1).
__global__ __launch_bounds__(256) void k_tri_z(float *__restrict__ v, long long ld, int n, int nlines,
                                               const float *__restrict__ d, const float *__restrict__ e, float r1)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    if (ix >= nlines) return;
    float *row = v + (long long)ix * ld;
    float *b = row + 1;
    b[0] = b[0] + r1 * 1.0f * row[0];
    for (int k = 1; k < n; ++k) b[k] = b[k] - e[k - 1] * b[k - 1];
    b[n - 1] = b[n - 1] / d[n - 1];
    for (int k = n - 1; k > 0; --k) b[k - 1] = b[k - 1] / d[k - 1] - e[k - 1] * b[k];
}

// tripd's decomposition (util.c:666-670) of a constant matrix, on the host
void tripd_decompose(std::vector<float> &d, std::vector<float> &e)
{
    const int n = (int)d.size();
    for (int k = 1; k < n; ++k) {
        const float temp = e[k - 1];
        e[k - 1] = temp / d[k - 1];
        d[k] -= temp * e[k - 1];
    }
}

}  // namespace
}  // namespace rz

using namespace rz;

extern "C" int rz_median_filt(int device, void *hip_stream, float *field, int nx, int ny, int *fsize, int on_device)
{
    if (!field || !fsize) return smooth_fail(RZ_ERR_INVALID, "median_filt: null pointer");
    if (nx < 1 || ny < 1) return smooth_fail(RZ_ERR_INVALID, "median_filt: NX, NY must be >= 1");
    // smooth_grad.c:28-35: 0 leaves the field; an even size moves away from 0
    if (*fsize == 0) return RZ_OK;
    if (!(*fsize % 2)) *fsize += *fsize > 0 ? 1 : -1;
    const int hfs = std::abs(*fsize) / 2;
    const long long k = (long long)(2 * hfs + 1) * (2 * hfs + 1);
    if (hfs > 64) return smooth_fail(RZ_ERR_INVALID, "median_filt: window above 129 x 129");
    const size_t ncell = (size_t)nx * ny, bytes = ncell * sizeof(float);
    if (hipSetDevice(device) != hipSuccess) return smooth_fail(RZ_ERR_INVALID, "median_filt: no device %d", device);
    // at most 256 MiB of per-thread windows
    const long long by_mem = std::max<long long>((256ll << 20) / (k * (long long)sizeof(float) * kMedThreads), 1);
    const int blocks = (int)std::min<long long>(
        std::min<long long>(std::max<long long>(((long long)ncell + kMedThreads - 1) / kMedThreads, 1), 2048), by_mem);
    float *src = nullptr, *dst = nullptr, *work = nullptr;
    hipStream_t st = (hipStream_t)hip_stream;
    auto done = [&](int rc) {
        (void)hipFree(src);
        if (!on_device) (void)hipFree(dst);
        (void)hipFree(work);
        return rc;
    };
    hipError_t e;
    if ((e = hipMalloc(&src, bytes)) != hipSuccess ||
        (e = hipMalloc(&work, (size_t)blocks * kMedThreads * k * sizeof(float))) != hipSuccess)
        return done(smooth_fail(RZ_ERR_NOMEM, "median_filt: device allocation failed"));
    if (on_device) {
        dst = field;
        e = hipMemcpyAsync(src, field, bytes, hipMemcpyDeviceToDevice, st);
    } else {
        if ((e = hipMalloc(&dst, bytes)) != hipSuccess)
            return done(smooth_fail(RZ_ERR_NOMEM, "median_filt: device allocation failed"));
        e = hipMemcpyAsync(src, field, bytes, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) {
        k_median<<<blocks, kMedThreads, 0, st>>>(src, dst, nx, ny, hfs, work);
        e = hipGetLastError();
    }
    if (e == hipSuccess && !on_device) e = hipMemcpyAsync(field, dst, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = st ? hipStreamSynchronize(st) : hipDeviceSynchronize();
    if (e != hipSuccess) return done(smooth_fail(RZ_ERR_HIP, "median_filt: %s", hipGetErrorString(e)));
    return done(RZ_OK);
}

extern "C" int rz_smooth2(int device, void *hip_stream, const float *in, float *out, int n2, int n1, float wd_damp,
                          float wd_damp1)
{
    if (!in || !out) return smooth_fail(RZ_ERR_INVALID, "smooth2: null pointer");
    if (n2 < 2 || n1 < 3) return smooth_fail(RZ_ERR_INVALID, "smooth2: needs NX >= 2 and NY >= 3");
    // smooth2.c:27-35: the damping parameters (float products with the double 0.25)
    float r1 = wd_damp, r2 = wd_damp1;
    r1 = r1 * r1 * 0.25;
    r2 = r2 * r2 * 0.25;
    const float w = 1.0f;
    // the matrices of both passes as smooth2.c:127-131,140-148 set them (w = 1), decomposed once
    std::vector<float> d1(n2), e1(n2 - 1), d2(n1 - 1), e2(n1 - 2);
    for (int ix = 0; ix < n2 - 1; ++ix) {
        d1[ix] = 1.0 + r2 * (w + w);
        e1[ix] = -r2 * w;
    }
    d1[0] -= r2 * w;
    d1[n2 - 1] = 1.0 + r2 * w;
    for (int iz = 0; iz < n1 - 2; ++iz) {
        d2[iz] = 1.0 + r1 * (w + w);
        e2[iz] = -r1 * w;
    }
    d2[n1 - 2] = 1.0 + r1 * w;
    tripd_decompose(d1, e1);
    tripd_decompose(d2, e2);

    if (hipSetDevice(device) != hipSuccess) return smooth_fail(RZ_ERR_INVALID, "smooth2: no device %d", device);
    hipStream_t st = (hipStream_t)hip_stream;
    // v[0..n2][0..n1], x-slow: the data at [1..n2][1..n1], zeros elsewhere
    const long long ld = n1 + 1;
    const size_t vb = (size_t)(n2 + 1) * ld * sizeof(float);
    std::vector<float> vh((size_t)(n2 + 1) * ld, 0.0f);
    for (int i = 1; i <= n2; ++i)
        std::copy(in + (size_t)(i - 1) * n1, in + (size_t)i * n1, vh.begin() + (size_t)i * ld + 1);
    float *v = nullptr, *fac = nullptr;
    auto done = [&](int rc) {
        (void)hipFree(v);
        (void)hipFree(fac);
        return rc;
    };
    const size_t nf = d1.size() + e1.size() + d2.size() + e2.size();
    hipError_t e;
    if ((e = hipMalloc(&v, vb)) != hipSuccess || (e = hipMalloc(&fac, nf * sizeof(float))) != hipSuccess)
        return done(smooth_fail(RZ_ERR_NOMEM, "smooth2: device allocation failed"));
    float *dd1 = fac, *de1 = dd1 + d1.size(), *dd2 = de1 + e1.size(), *de2 = dd2 + d2.size();
    if ((e = hipMemcpyAsync(v, vh.data(), vb, hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(dd1, d1.data(), d1.size() * sizeof(float), hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(de1, e1.data(), e1.size() * sizeof(float), hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(dd2, d2.data(), d2.size() * sizeof(float), hipMemcpyHostToDevice, st)) == hipSuccess &&
        (e = hipMemcpyAsync(de2, e2.data(), e2.size() * sizeof(float), hipMemcpyHostToDevice, st)) == hipSuccess) {
        // pass 1: for z = 0 .. n1-1 (its row z of v is element z of every column)
        k_tri_x<<<(n1 + 255) / 256, 256, 0, st>>>(v, ld, n2, n1, dd1, de1);
        // pass 2: for x = 0 .. n2-1 over z = 1 .. n1-1
        k_tri_z<<<(n2 + 255) / 256, 256, 0, st>>>(v, ld, n1 - 1, n2, dd2, de2, r1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(vh.data(), v, vb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = st ? hipStreamSynchronize(st) : hipDeviceSynchronize();
    if (e != hipSuccess) return done(smooth_fail(RZ_ERR_HIP, "smooth2: %s", hipGetErrorString(e)));
    for (int i = 1; i <= n2; ++i)
        std::copy(vh.begin() + (size_t)i * ld + 1, vh.begin() + (size_t)i * ld + 1 + n1, out + (size_t)(i - 1) * n1);
    return done(RZ_OK);
}
