// xc_fmode.hip -- CorrImagesF's peak selection on the device, after the
// R-mode pipeline has written the centred R image and its candidate flags:
// MakeF (LoG+I filter of R, guard band, normalisation), OrderF (candidates
// by decreasing F, radix-sorted), FPeak (the first well-isolated candidate,
// evaluated a chunk of candidates at a time, one workgroup each, then walked
// in order by one workgroup that applies the core masks), ReturnR.
#include <limits.h>
#include <math.h>

#include <hipcub/hipcub.hpp>

#include "xc_fmode.hpp"
#include "xc_kernels.hpp"

namespace xc {

constexpr int F_BLOCKS = 1024;  // partial maxima of the F image
constexpr int F_CHUNK = 256;    // FPeak candidates evaluated per round

#pragma clang fp contract(off)
// MakeF (Correlation.cpp:2412-2492): F = R convolved with the 3x3 kernel
// {-1,-1,-1; -1,9,-1; -1,-1,-1} (its sum is 1, so prenormalisation leaves
// it as is).  The reference convolves through 2-D transforms; here the
// nine-point stencil is summed directly -- the same value up to rounding
// (the transforms' own rounding is of order 1e-15 relative).  Pixels
// within 2 of the border (the guard band) get F = 0 and lose their
// candidate flag; elsewhere every neighbour lies inside the image.  Writes
// the per-block maxima of F.
__global__ void __launch_bounds__(256) k_makef(const double *__restrict__ R, uint8_t *__restrict__ cand,
                                               double *__restrict__ F, int wR, int hR, double *__restrict__ bmax)
{
    __shared__ double sm[256];
    const long long nR = (long long)wR * hR;
    double m = -1.0E300;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nR; i += (long long)gridDim.x * 256) {
        const int y = (int)(i / wR), x = (int)(i - (long long)wR * y);
        double f = 0.0;
        if (x < 2 || x >= wR - 2 || y < 2 || y >= hR - 2) {
            cand[i] = 0;
        } else {
            const double *c = R + i;
            const double nb = c[-wR - 1] + c[-wR] + c[-wR + 1] + c[-1] + c[1] + c[wR - 1] + c[wR] + c[wR + 1];
            f = 9.0 * c[0] - nb;
        }
        F[i] = f;
        m = f > m ? f : m;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off && sm[threadIdx.x + off] > sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = sm[0];
}

// F / max(F, 1e-7) where F > 0, else 0 (MakeF's normalisation); sort keys
// (F for a candidate, -1 below every F otherwise) with the pixel index as
// value, and the candidate count (OrderF's list).
__global__ void __launch_bounds__(256) k_fnorm(double *__restrict__ F, const uint8_t *__restrict__ cand,
                                               const double *__restrict__ bmax, int nb, long long nR,
                                               double *__restrict__ keys, int *__restrict__ vals,
                                               int *__restrict__ ncand)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    double vmax = 1e-7;
    for (int b = 0; b < nb; ++b) vmax = bmax[b] > vmax ? bmax[b] : vmax;
    __syncthreads();
    int c = 0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nR; i += (long long)gridDim.x * 256) {
        const double f = F[i] > 0.0 ? F[i] / vmax : 0.0;
        F[i] = f;
        keys[i] = cand[i] ? f : -1.0;
        vals[i] = (int)i;
        c += cand[i] ? 1 : 0;
    }
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(ncand, cnt);
}

// FPeak's per-candidate tests (Correlation.cpp:2607-2760) for candidates
// j0 .. j0 + gridDim.x - 1 of `order`, one workgroup each.  rec = {status,
// core half-width}: status 0 = rejected before the core is masked (an axis
// never falls to half height, a micro-peak, or the outer band leaves the
// image), 1 = core masked then rejected (a high neighbour), 2 = accepted.
__global__ void __launch_bounds__(256) k_fpeak_eval(const double *__restrict__ F, const double *__restrict__ R,
                                                    const int *__restrict__ order, int j0, int wR, int hR,
                                                    double nbmaxht, int2 *__restrict__ rec)
{
    __shared__ int dmin[4];
    __shared__ int high;
    const int tid = threadIdx.x;
    if (tid < 4) dmin[tid] = INT_MAX;
    if (tid == 0) high = 0;
    __syncthreads();
    const int k = order[j0 + blockIdx.x];
    const int ry = k / wR, rx = k - wR * ry;
    const double fpk = F[k], half = 0.5 * fpk, tol = nbmaxht * fpk;
    // inner radius: per direction (up, down, left, right) the nearest pixel with F <= fpk / 2
    const int len[4] = {ry, hR - 1 - ry, rx, wR - 1 - rx};
    for (int dir = 0; dir < 4; ++dir)
        for (int d = 1 + tid; d <= len[dir]; d += 256) {
            const int x = dir == 2 ? rx - d : (dir == 3 ? rx + d : rx);
            const int y = dir == 0 ? ry - d : (dir == 1 ? ry + d : ry);
            if (F[x + (long long)wR * y] <= half) {
                atomicMin(&dmin[dir], d);
                break;
            }
        }
    __syncthreads();
    int st = 0, rm = 0;
    const bool closed = dmin[0] != INT_MAX && dmin[1] != INT_MAX && dmin[2] != INT_MAX && dmin[3] != INT_MAX;
    const int ri = min(min(dmin[0], dmin[1]), min(dmin[2], dmin[3]));
    const int ro = closed ? 3 * ri : 0;
    // uniform across the workgroup: every test below reads shared or per-candidate values only
    if (closed && !(ri <= 1 && R[k] < 0.1) &&
        !(ry - ro < 0 || ry + ro >= hR || rx - ro < 0 || rx + ro >= wR)) {
        rm = ri / 3 < 1 ? 1 : ri / 3;
        const int side = 2 * ro + 1;
        const long long tot = (long long)side * side;
        bool hi = false;
        for (long long p = tid; p < tot && !hi; p += 256) {
            const int yy = ry - ro + (int)(p / side), xx = rx - ro + (int)(p % side);
            if (yy >= ry - ri && yy <= ry + ri && xx >= rx - ri && xx <= rx + ri) continue;
            hi = F[xx + (long long)wR * yy] > tol;
        }
        if (hi) atomicOr(&high, 1);
        __syncthreads();
        st = high ? 1 : 2;
    }
    if (tid == 0) rec[blockIdx.x] = make_int2(st, rm);
}

// FPeak's walk over one evaluated chunk, in order: a candidate inside the
// core of an earlier one is skipped; one that reached the core step masks
// its core; the first accepted one ends the search (*found = its index).
__global__ void __launch_bounds__(256) k_fpeak_walk(const int *__restrict__ order, int j0, int cnt,
                                                    const int2 *__restrict__ rec, uint8_t *__restrict__ mask, int wR,
                                                    int *__restrict__ found)
{
    __shared__ int s_act, s_k, s_rm;
    const int tid = threadIdx.x;
    for (int j = 0; j < cnt; ++j) {
        if (tid == 0) {
            const int k = order[j0 + j];
            const int2 r = rec[j];
            s_k = k;
            s_rm = r.y;
            s_act = (r.x == 0 || mask[k]) ? 0 : r.x;
        }
        __syncthreads();
        const int act = s_act, k = s_k, rm = s_rm;
        if (act) {
            const int ry = k / wR, rx = k - wR * ry, side = 2 * rm + 1;
            for (int p = tid; p < side * side; p += 256)
                mask[(rx - rm + p % side) + (long long)wR * (ry - rm + p / side)] = 1;
        }
        __syncthreads();  // the core is painted before the next candidate's test, and s_* may be reused
        if (act == 2) {
            if (tid == 0) *found = k;
            return;
        }
    }
}

// ReturnR( dx, dy, rx, ry, R, R ) (Correlation.cpp:3016-3043) with ParabPeak
// (:2975-3010) on the R image, or no peak (dx = dy = 0, r = 0).
__global__ void k_fresult(const PairDev *__restrict__ pairs, const double *__restrict__ R, const int *__restrict__ found)
{
    const PairDev &P = pairs[0];
    xc_result res;
    res.r = 0.0;
    res.dx = 0.0;
    res.dy = 0.0;
    res.peak = 0.0;
    res.sub_x = 0.0;
    res.sub_y = 0.0;
    res.lag_x = 0;
    res.lag_y = 0;
    res.nx = P.Nx_ref;
    res.ny = P.Ny_ref;
    res._pad = 0;
    const int k = *found;
    if (k < 0) {
        res.status = XC_NO_PEAK;
    } else {
        const int wR = P.wR, hR = P.hR;
        const int iy = k / wR, ix = k - wR * iy;
        const double z1 = R[k];
        double sx = 0.0, sy = 0.0, z0, z2;
        if (ix - 1 >= 0 && ix + 1 < wR && (z0 = R[k - 1]) > 0.0 && (z2 = R[k + 1]) > 0.0) {
            const double d = 1 * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
            if (fabs(d) <= 8 * 1) sx = d;
        }
        if (iy - 1 >= 0 && iy + 1 < hR && (z0 = R[k - wR]) > 0.0 && (z2 = R[k + wR]) > 0.0) {
            const double d = 1 * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
            if (fabs(d) <= 8 * 1) sy = d;
        }
        res.status = XC_OK;
        res.r = z1;
        res.peak = z1;
        res.dx = (ix + sx) + (P.offx - P.nnegx);
        res.dy = (iy + sy) + (P.offy - P.nnegy);
        res.sub_x = sx;
        res.sub_y = sy;
        res.lag_x = ix - P.nnegx;
        res.lag_y = iy - P.nnegy;
    }
    *reinterpret_cast<xc_result *>(P.out) = res;
}
#pragma clang fp contract(on)

static size_t sort_temp_bytes(long long nR)
{
    size_t b = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, (const double *)nullptr, (double *)nullptr,
                                                       (const int *)nullptr, (int *)nullptr, (int)nR);
    return b;
}

static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

size_t fmode_bytes(long long nR)
{
    return al(8 * nR) * 4 + al(4 * nR) * 2 + al(nR) * 2 + al(8 * F_BLOCKS) + al(8 * F_CHUNK) + al(64) +
           al(sort_temp_bytes(nR) + 1);
}

void fmode_carve(char *base, long long nR, FWork &w)
{
    size_t o = 0;
    auto take = [&](size_t b) {
        char *p = base + o;
        o += al(b);
        return p;
    };
    w.R = (double *)take(8 * nR);
    w.F = (double *)take(8 * nR);
    w.kin = (double *)take(8 * nR);
    w.kout = (double *)take(8 * nR);
    w.vin = (int *)take(4 * nR);
    w.vout = (int *)take(4 * nR);
    w.cand = (uint8_t *)take(nR);
    w.mask = (uint8_t *)take(nR);
    w.bmax = (double *)take(8 * F_BLOCKS);
    w.rec = (int2 *)take(8 * F_CHUNK);
    w.ints = (int *)take(64);
    w.sort_bytes = sort_temp_bytes(nR);
    w.sort = take(w.sort_bytes + 1);
}

int fmode_select(xc_ctx *ctx, hipStream_t s, const PairDev *dP, int wR, int hR, double nbmaxht, const FWork &w)
{
    const long long nR = (long long)wR * hR;
    const int nb = (int)std::min<long long>(F_BLOCKS, (nR + 255) / 256);
    int *ncand = w.ints, *found = w.ints + 1;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, w.ints, 16LL, 0);
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, found, 1LL, -1);
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3(nb), dim3(256), 0, s, w.mask, nR, (uint8_t)0);
    hipLaunchKernelGGL(k_makef, dim3(nb), dim3(256), 0, s, w.R, w.cand, w.F, wR, hR, w.bmax);
    hipLaunchKernelGGL(k_fnorm, dim3(nb), dim3(256), 0, s, w.F, w.cand, w.bmax, nb, nR, w.kin, w.vin, ncand);
    XC_HIP(ctx, hipGetLastError());
    size_t tb = w.sort_bytes;
    XC_HIP(ctx, hipcub::DeviceRadixSort::SortPairsDescending(w.sort, tb, w.kin, w.kout, w.vin, w.vout, (int)nR, 0,
                                                             64, s));
    int hn = 0, hf = -1;
    XC_HIP(ctx, hipMemcpyAsync(&hn, ncand, sizeof(int), hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    for (int j0 = 0; j0 < hn && hf < 0; j0 += F_CHUNK) {
        const int cnt = std::min(F_CHUNK, hn - j0);
        hipLaunchKernelGGL(k_fpeak_eval, dim3(cnt), dim3(256), 0, s, w.F, w.R, w.vout, j0, wR, hR, nbmaxht, w.rec);
        hipLaunchKernelGGL(k_fpeak_walk, dim3(1), dim3(256), 0, s, w.vout, j0, cnt, w.rec, w.mask, wR, found);
        XC_HIP(ctx, hipGetLastError());
        XC_HIP(ctx, hipMemcpyAsync(&hf, found, sizeof(int), hipMemcpyDeviceToHost, s));
        XC_HIP(ctx, hipStreamSynchronize(s));
    }
    hipLaunchKernelGGL(k_fresult, dim3(1), dim3(1), 0, s, dP, w.R, found);
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

}  // namespace xc
