// xc_splat.hip -- ImageFromValuesAndPoints (Geometry.cpp:661-685) with
// DistributePixel's bilinear weights (Maths.h:236-276), in the reference's
// order of additions.
//
// The reference adds, point by point, up to four weighted copies of each
// value into the zeroed raster, so a pixel's value is the in-order sum of the
// contributions that land on it.  Points on the integer grid put their value
// on one pixel (the other three weights are zero, and adding signed zeros to a
// non-zero value, or to +0, changes nothing), so device atomics give that sum
// exactly as long as no pixel gets more than two non-zero contributions.  For
// lists with off-grid points (rotated or tweaked), the four contributions of
// every point are emitted with their pixel index, stably radix-sorted by pixel
// (which keeps point order inside a pixel) and each pixel's run summed in
// order: the reference's value, bit for bit, and the same every run.
#include <math.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

#pragma clang fp contract(off)

// The four contributions of point i, in DistributePixel's order; out-of-raster ones get the key npix.
__global__ void k_splat_emit(const xc_point *__restrict__ pts, const double *__restrict__ val, int n, double xmin,
                             double ymin, int w, int h, uint32_t *__restrict__ key, double *__restrict__ cv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t npix = (uint32_t)w * (uint32_t)h;
    const double x = pts[i].x - xmin, y = pts[i].y - ymin, v = val[i];
    const int ix = (int)floor(x), iy = (int)floor(y);
    const double alpha = x - ix, beta = y - iy;
    const bool x0 = ix >= 0 && ix < w, x1 = ix + 1 >= 0 && ix + 1 < w;
    const bool y0 = iy >= 0 && iy < h, y1 = iy + 1 >= 0 && iy + 1 < h;
    const uint32_t p = (uint32_t)ix + (uint32_t)w * (uint32_t)iy;  // used only when in the raster
    key[4 * i + 0] = (x0 && y0) ? p : npix;
    cv[4 * i + 0] = (1.0 - alpha) * (1.0 - beta) * v;
    key[4 * i + 1] = (x0 && y1) ? p + (uint32_t)w : npix;
    cv[4 * i + 1] = (1.0 - alpha) * beta * v;
    key[4 * i + 2] = (x1 && y0) ? p + 1 : npix;
    cv[4 * i + 2] = alpha * (1.0 - beta) * v;
    key[4 * i + 3] = (x1 && y1) ? p + 1 + (uint32_t)w : npix;
    cv[4 * i + 3] = alpha * beta * v;
}

// Each run of equal keys (one pixel) summed in order by its first element's thread.
__global__ void k_splat_runs(const uint32_t *__restrict__ key, const double *__restrict__ cv, int m, uint32_t npix,
                             double *__restrict__ img)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t k = key[j];
    if (k >= npix || (j > 0 && key[j - 1] == k)) return;
    double s = 0.0;  // the zeroed raster pixel
    for (int t = j; t < m && key[t] == k; ++t) s += cv[t];
    img[k] = s;
}

#pragma clang fp contract(on)

static int key_bits(uint32_t npix)
{
    int b = 1;
    while (b < 32 && (npix >> b) != 0) ++b;
    return b;
}

// Copies n points and reports whether any coordinate is not an integer (blocks of 256
// points without branches, so the rounding test vectorises; coordinates beyond +-2^51
// count as off the grid).  dst may be null (check only).
bool copy_points_off_grid(xc_point *dst, const xc_point *p, int n)
{
    const double C = 6755399441055744.0;  // 1.5 * 2^52: (v + C) - C rounds v to an integer (|v| < 2^51)
    bool bad = false;
    for (int i0 = 0; i0 < n; i0 += 256) {
        const int i1 = n - i0 < 256 ? n : i0 + 256;
        for (int i = i0; i < i1; ++i) {
            const double x = p[i].x, y = p[i].y;
            if (dst) dst[i] = p[i];
            bad |= ((x + C) - C != x) | ((y + C) - C != y) | !(fabs(x) < 2251799813685248.0) |
                   !(fabs(y) < 2251799813685248.0);
        }
        if (bad && !dst) return true;
    }
    return bad;
}

bool points_off_grid(const xc_point *p, int n) { return copy_points_off_grid(nullptr, p, n); }

bool stage_points(xc_point *dp, const xc_point *p, double *da, const double *a, double *db, const double *b, int n)
{
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = n < (1 << 17) ? 1 : std::max(1, std::min(8, hw));
    std::vector<char> og(nt, 0);
    auto part = [&](int t) {
        const long long lo = (long long)n * t / nt, hi = (long long)n * (t + 1) / nt;
        const int c = (int)(hi - lo);
        og[t] = copy_points_off_grid(dp + lo, p + lo, c);
        if (da) memcpy(da + lo, a + lo, sizeof(double) * c);
        if (db) memcpy(db + lo, b + lo, sizeof(double) * c);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) {
        try {
            th.emplace_back(part, t);
        } catch (...) {  // no thread to be had: that part on this one
            part(t);
        }
    }
    part(0);
    for (auto &x : th) x.join();
    bool any = false;
    for (char c : og) any = any || c;
    return any;
}

size_t splat_scratch_bytes(int n, long long npix)
{
    if (n <= 0) return 0;
    const int m = 4 * n, bits = key_bits((uint32_t)npix);
    // the sort's temporary storage, asked once per (items, key bits)
    static std::mutex mu;
    static std::map<std::pair<int, int>, size_t> known;
    size_t cub = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = known.find({m, bits});
        if (it != known.end()) {
            cub = it->second;
        } else {
            (void)hipcub::DeviceRadixSort::SortPairs(nullptr, cub, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                     (const double *)nullptr, (double *)nullptr, m, 0, bits);
            known[{m, bits}] = cub;
        }
    }
    return 4 * 256 + (size_t)m * (2 * sizeof(uint32_t) + 2 * sizeof(double)) + cub;
}

int splat_points(xc_ctx *ctx, hipStream_t s, const xc_point *pts, const double *val, int n, double xmin, double ymin,
                 double *img, int w, int h, bool off_grid, void *scratch, size_t scratch_bytes)
{
    if (n <= 0) return XC_OK;
    if (!off_grid) {
        hipLaunchKernelGGL(k_splat_points, dim3((n + 255) / 256), dim3(256), 0, s, pts, val, n, xmin, ymin, img, w,
                           h);
        XC_HIP(ctx, hipGetLastError());
        return XC_OK;
    }
    const long long npix = (long long)w * h;
    if (npix >= 0xffffffffLL) return set_err(ctx, XC_E_SHAPE, "raster of %lld pixels too large to splat", npix);
    const int m = 4 * n;
    const size_t need = splat_scratch_bytes(n, npix);
    if (!scratch || scratch_bytes < need) return set_err(ctx, XC_E_NOMEM, "splat scratch %zu < %zu", scratch_bytes, need);
    auto align = [](char *p) { return (char *)(((uintptr_t)p + 255) & ~(uintptr_t)255); };
    char *q = align((char *)scratch);
    uint32_t *k0 = (uint32_t *)q;
    q = align(q + sizeof(uint32_t) * m);
    uint32_t *k1 = (uint32_t *)q;
    q = align(q + sizeof(uint32_t) * m);
    double *v0 = (double *)q;
    q = align(q + sizeof(double) * m);
    double *v1 = (double *)q;
    q = q + sizeof(double) * m;
    size_t cub = need - (size_t)(q - (char *)scratch);
    hipLaunchKernelGGL(k_splat_emit, dim3((n + 255) / 256), dim3(256), 0, s, pts, val, n, xmin, ymin, w, h, k0, v0);
    XC_HIP(ctx, hipGetLastError());
    XC_HIP(ctx, hipcub::DeviceRadixSort::SortPairs((void *)q, cub, k0, k1, v0, v1, m, 0, key_bits((uint32_t)npix), s));
    hipLaunchKernelGGL(k_splat_runs, dim3((m + 255) / 256), dim3(256), 0, s, k1, v1, m, (uint32_t)npix, img);
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

}  // namespace xc
