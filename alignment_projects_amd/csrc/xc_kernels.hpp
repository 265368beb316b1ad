// xc_kernels.hpp -- HIP kernels of the tile-pair FFT cross-correlation.
//
// Data flow for one correlation (every stage batched over a work list so the
// same kernels serve the plain one-pair FP64 form and the batched FP32 form):
//
//   pack / splat  -> compact A, B rasters (w x h, T)
//   sat_rows/cols -> summed-area tables S, S^2 (FP64 entries; FP32 on request) and int32 nz
//   row_r2c       -> row spectra [h][Nx/2+1] of A and B (zero pad implicit)
//   col_fused     -> per spectrum column: Ny-point forward FFT of A and B,
//                    P = F_B * conj(F_A) (whitened in S mode), Ny-point
//                    inverse FFT; only the hR valid lag rows are written
//   row_c2r_eval  -> per lag row: Nx-point c2r, NCC / R / S per lag with the
//                    legality rules, per-row best + neighbours
//   finalize      -> reduce rows, y-neighbours by a direct DFT of two lag
//                    rows, parabolic subpixel fit, result record
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>

#include <type_traits>

#include "../../include/xcorr.h"
#include "xc_fft.hpp"
#include "xc_types.hpp"

namespace xc {

// ------------------------------------------------------------------
// Small helpers
// ------------------------------------------------------------------

__host__ __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// "a better than b" in the peak order: larger value, then smaller order key.
__device__ __forceinline__ bool better(double va, long long ka, double vb, long long kb)
{
    return va > vb || (va == vb && ka < kb);
}

// Legality of an overlap of olw x olh (the LegalRgn descriptor).
__host__ __device__ __forceinline__ bool rgn_ok(const RuleDev &g, int olw, int olh)
{
    if (!g.rgn_enabled) return true;
    if (olw < g.min_w || olh < g.min_h) return false;
    if (!((long long)olw * olh > g.min_area)) return false;
    if (g.exact_w > 0 && olw != g.exact_w) return false;
    if (g.exact_h > 0 && olh != g.exact_h) return false;
    if ((g.or_w_ge > 0 || g.or_h_gt > 0) && !(olw >= g.or_w_ge || olh > g.or_h_gt)) return false;
    if (g.table) {
        if (olw > g.table_w || olh > g.table_h) return false;
        if (!g.table[(long long)(olh - 1) * g.table_w + (olw - 1)]) return false;
    }
    return true;
}

// IntegralTable (Correlation.cpp:160-201) for the box [L,R] x [B,T].
template <typename TT>
__device__ __forceinline__ TT box_sum(const TT *__restrict__ t, int w, int L, int R, int B, int T)
{
    TT r = gld(t, R + (long long)w * T);
    if (L > 0) r -= gld(t, L - 1 + (long long)w * T);
    if (B > 0) r -= gld(t, R + (long long)w * (B - 1));
    if (L > 0 && B > 0) r += gld(t, (L - 1) + (long long)w * (B - 1));
    return r;
}

// Interleaved (S, S^2) table entries, by table type TB: FP64 pairs (every
// form by default, as IntegrateImage keeps double tables), or FP32 pairs for
// tiles that ask for them -- each entry the FP64 sum rounded once; lookups
// widen it back to FP64 and all box arithmetic is FP64.
template <typename FT> struct TabE {
    typedef double2 type;
    static __host__ __device__ __forceinline__ type make(double a, double b) { return make_double2(a, b); }
};
template <> struct TabE<float> {
    typedef float2 type;
    static __host__ __device__ __forceinline__ type make(double a, double b) { return make_float2((float)a, (float)b); }
};
template <typename FT> using tab_t = typename TabE<FT>::type;

// Table entry i of an FP table (S, S^2) pair, stored through a global pointer.
template <typename FT> __device__ __forceinline__ void tab_st(void *t, long long i, double a, double b)
{
    typedef typename std::conditional<std::is_same<FT, float>::value, float, double>::type E;
    gst2(reinterpret_cast<E *>(t), 2 * i, (E)a, (E)b);
}

template <typename FT>
__device__ __forceinline__ double2 tab_at(const void *t, long long i)
{
    const cpx<FT> v = gld(reinterpret_cast<const cpx<FT> *>(t), i);
    return make_double2((double)v.x, (double)v.y);
}

// Exact integer tables (TabX, the TK_X form) of unmasked 8-bit images of at
// most 2^20 pixels: entry (x, y) is the uint64 (Sraw << 36) + Qraw, Sraw the
// sum of the raw pixel values (< 2^28) and Qraw that of their squares
// (< 2^36) over [0, x] x [0, y].  Neither field can carry into the other, so
// a box's four-corner combination in 64-bit (modular) arithmetic is the packed
// pair of the box's exact sums.
struct TabX {};
constexpr int TABX_QB = 36;
constexpr unsigned long long TABX_QMASK = (1ULL << TABX_QB) - 1;
constexpr long long TABX_MAX_PX = 1LL << 20;
template <> struct TabE<TabX> {
    typedef unsigned long long type;
};
template <typename TB> __host__ __device__ constexpr bool is_tabx() { return std::is_same<TB, TabX>::value; }

// A table corner as the box arithmetic carries it: the (S, S^2) pair widened
// to FP64 for the FP tables, the packed pair for TabX.
template <typename TB> struct CorT {
    typedef double2 type;
};
template <> struct CorT<TabX> {
    typedef unsigned long long type;
};
template <typename TB> using cor_t = typename CorT<TB>::type;

__host__ __device__ __forceinline__ double2 c_add(double2 a, double2 b)
{
    a.x += b.x;
    a.y += b.y;
    return a;
}
__host__ __device__ __forceinline__ double2 c_sub(double2 a, double2 b)
{
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
__host__ __device__ __forceinline__ unsigned long long c_add(unsigned long long a, unsigned long long b) { return a + b; }
__host__ __device__ __forceinline__ unsigned long long c_sub(unsigned long long a, unsigned long long b) { return a - b; }
template <typename C> __host__ __device__ __forceinline__ C c_zero();
template <> __host__ __device__ __forceinline__ double2 c_zero<double2>() { return make_double2(0.0, 0.0); }
template <> __host__ __device__ __forceinline__ unsigned long long c_zero<unsigned long long>() { return 0ULL; }

template <typename TB>
__device__ __forceinline__ cor_t<TB> tab_c(const void *t, long long i)
{
    if constexpr (is_tabx<TB>()) return gld(reinterpret_cast<const unsigned long long *>(t), i);
    else return tab_at<TB>(t, i);
}

// The same lookups on an (S, S^2) table: one load per corner, combined in
// box_sum's order.
template <typename TB>
__device__ __forceinline__ cor_t<TB> box_c(const void *__restrict__ t, int w, int L, int R, int B, int T)
{
    cor_t<TB> r = tab_c<TB>(t, R + (long long)w * T);
    if (L > 0) r = c_sub(r, tab_c<TB>(t, L - 1 + (long long)w * T));
    if (B > 0) r = c_sub(r, tab_c<TB>(t, R + (long long)w * (B - 1)));
    if (L > 0 && B > 0) r = c_add(r, tab_c<TB>(t, (L - 1) + (long long)w * (B - 1)));
    return r;
}

// An image's normalisation, which exact tables apply at lookup: mean, 1/sd
// and 1/sd^2 (FP64); for the FP32 evaluation also the image's exact raw sum
// T (the table's last entry), its pixel count N, (1/sd) / N and 1/sd^2 in FP32.
struct NmD {
    double m, i, i2;
    unsigned long long T;
    unsigned N;
    float iN, i2f;
};
__device__ __forceinline__ NmD nm_of(const ImgView &v)
{
    const double i = v.msd[2];
    return NmD{v.msd[0], i, i * i, 0ULL, 0u, 0.0f, 0.0f};
}
// nm_of plus the FP32 evaluation's terms, from the image's exact table (w x h).
__device__ __forceinline__ NmD nm_of_x(const ImgView &v, const void *tab, int w, int h)
{
    NmD r = nm_of(v);
    const unsigned N = (unsigned)(w * h);
    r.T = reinterpret_cast<const unsigned long long *>(tab)[N - 1] >> TABX_QB;
    r.N = N;
    r.iN = (float)(r.i / (double)N);
    r.i2f = (float)r.i2;
    return r;
}

#pragma clang fp contract(off)
// GetCorr's numerator and variance product (Correlation.cpp:716-754) from two
// exact box sums.  With v = (raw - mean) / sd over a box of n pixels,
//   s = (Sraw - n mean) / sd   and   n q - s^2 = (n Qraw - Sraw^2) / sd^2,
// the bracket formed in 64-bit integers (n Qraw < 2^56): the variance
// subtraction is exact and only its final scaling rounds.  nraw: n raw / Nxy.
// n <= 2^20 (the box lies inside a tile of at most 2^20 pixels), so every
// product is one 32 x 32 -> 64-bit multiply.
__device__ __forceinline__ unsigned long long tabx_var(unsigned n, unsigned long long a, unsigned *S)
{
    *S = (unsigned)(a >> TABX_QB);                                   // < 2^28
    const unsigned qlo = (unsigned)a, qhi = (unsigned)(a >> 32) & 15u;  // Qraw = qhi 2^32 + qlo
    const unsigned long long nq = (unsigned long long)n * qlo + ((unsigned long long)(n * qhi) << 32);
    return nq - (unsigned long long)*S * *S;
}

__device__ __forceinline__ void tabx_num_d(unsigned n, double nraw, unsigned long long a, unsigned long long b, NmD A,
                                           NmD B, double *num, double *d)
{
    unsigned Sa, Sb;
    const unsigned long long Da = tabx_var(n, a, &Sa), Db = tabx_var(n, b, &Sb);
    const double dn = (double)n;
    const double s1 = ((double)Sa - dn * A.m) * A.i;
    const double s2 = ((double)Sb - dn * B.m) * B.i;
    const double d1 = (double)Da * A.i2;
    const double d2 = (double)Db * B.i2;
    *num = nraw - s1 * s2;
    *d = d1 * d2;
}

// 64-bit integers to FP32 in two halves (each conversion rounds, the fma once more).
__device__ __forceinline__ float u64_to_f32(unsigned long long v)
{
    return fmaf((float)(unsigned)(v >> 32), 4294967296.0f, (float)(unsigned)v);
}
__device__ __forceinline__ float i64_to_f32(long long v)
{
    return fmaf((float)(int)(v >> 32), 4294967296.0f, (float)(unsigned)v);
}

// GetCorr's r from exact box sums in the batched form's FP32: every step that
// cancels is an exact integer -- the variance bracket n Qraw - Sraw^2 and the
// mean-removed sum Sraw N - n T (= N (Sraw - n mean)) -- and only well-
// conditioned products, the square root and the quotient are FP32.
__device__ __forceinline__ double tabx_r32(unsigned n, double raw, double invN, unsigned long long a,
                                           unsigned long long b, const NmD &A, const NmD &B)
{
    unsigned Sa, Sb;
    const unsigned long long Da = tabx_var(n, a, &Sa), Db = tabx_var(n, b, &Sb);
    const long long Ea = (long long)((unsigned long long)Sa * A.N) - (long long)((unsigned long long)n * A.T);
    const long long Eb = (long long)((unsigned long long)Sb * B.N) - (long long)((unsigned long long)n * B.T);
    const float s1 = i64_to_f32(Ea) * A.iN, s2 = i64_to_f32(Eb) * B.iN;
    const float d = (u64_to_f32(Da) * A.i2f) * (u64_to_f32(Db) * B.i2f);
    const float fn = (float)n;
    const float num = fn * (float)raw * (float)invN - s1 * s2;
    return d < fn * fn * 1.0E-9f ? 0.0 : (double)(num * rsqrtf(d));
}
#pragma clang fp contract(on)

// Row part of BoxesFromShifts (Geometry.cpp:150-176) for signed lag y.
struct RowBox {
    int a_b, a_t, b_b, b_t, olh;
};

__host__ __device__ __forceinline__ RowBox row_box(const PairDev &P, int y)
{
    RowBox rb;
    int B = imax(0, y), T = imin(P.h2 - 1, y + P.h1 - 1);
    T = imax(T, B);
    rb.b_b = B;
    rb.b_t = T;
    rb.a_b = B - y;
    rb.a_t = T - y;
    rb.olh = T - B + 1;
    return rb;
}

// Value of the correlation surface at signed lag (x, y) from its raw
// (unnormalised c2r) value, with the mode's semantics:
//   NCC: CLinCorr::CheckSize/CheckDensity/GetCorr (Correlation.cpp:665-754)
//   R  : RCalc::Valid + CalcR (Correlation.cpp:1794-1845)
//   S  : RCalc::Valid + CalcS (Correlation.cpp:1766)
// *legal receives the candidate flag (LegalRgn and LegalCnt), *odd the
// exit(44) condition of GetCorr.
#pragma clang fp contract(off)
template <typename FT>
__device__ __forceinline__ double eval_lag(const PairDev &P, const RowBox &rb, int x, double raw,
                                           bool *legal, bool *odd)
{
    int BL = imax(0, x), BR = imin(P.w2 - 1, x + P.w1 - 1);
    BR = imax(BR, BL);
    int AL = BL - x, AR = BR - x;
    int olw = BR - BL + 1;
    bool ok = rgn_ok(P.rule, olw, rb.olh);
    *odd = false;
    // Nx Ny is a power of two: multiplying by its reciprocal is the division, exactly
    const double invNxy = 1.0 / ((double)P.Nx * (double)P.Ny);
    if (P.mode == PM_NCC) {
        if (!ok) { *legal = false; return 0.0; }
        if (P.rule.cnt_enabled) {
            int c1 = box_sum(P.ZA, P.w1, AL, AR, rb.a_b, rb.a_t);
            int c2 = box_sum(P.ZB, P.w2, BL, BR, rb.b_b, rb.b_t);
            if (!((long long)c1 > P.rule.k1 && (long long)c2 > P.rule.k2)) { *legal = false; return 0.0; }
        }
    } else if (ok && P.rule.cnt_enabled) {  // a lag the region rule rejects needs no counts
        int c1 = box_sum(P.ZA, P.w1, AL, AR, rb.a_b, rb.a_t);
        int c2 = box_sum(P.ZB, P.w2, BL, BR, rb.b_b, rb.b_t);
        if (!((long long)c1 > P.rule.k1 && (long long)c2 > P.rule.k2)) ok = false;
    }
    *legal = ok;
    if (P.mode == PM_S) return raw * invNxy;
    double n = (double)(olw * rb.olh);
    const cor_t<FT> sqa = box_c<FT>(P.SQA, P.w1, AL, AR, rb.a_b, rb.a_t);
    const cor_t<FT> sqb = box_c<FT>(P.SQB, P.w2, BL, BR, rb.b_b, rb.b_t);
    double num, d;
    if constexpr (is_tabx<FT>()) {
        tabx_num_d((unsigned)(olw * rb.olh), n * raw * invNxy, sqa, sqb, nm_of(P.A), nm_of(P.B), &num, &d);
    } else {
        double s1 = sqa.x, q1 = sqa.y, s2 = sqb.x, q2 = sqb.y;
        num = n * raw * invNxy - s1 * s2;
        double d1 = n * q1 - s1 * s1;
        double d2 = n * q2 - s2 * s2;
        d = d1 * d2;
    }
    double r = (d < n * n * 1.0E-9 ? 0.0 : num / sqrt(d));
    if (P.mode == PM_NCC) {
        *odd = fabs(r) > 1.0001;
        return r;
    }
    return (r > -1.0 && r < 1.0 ? r : 0.0);
}
#pragma clang fp contract(on)

// Lag-row index of raster row iy, or -1 for the dead zone.
__device__ __forceinline__ int lag_row_of_raster(const PairDev &P, int iy)
{
    if (P.lag_order == LAG_ALL) return iy;
    int y;
    if (iy < P.nposy) y = iy;
    else if (iy >= P.Ny - P.nnegy) y = iy - P.Ny;
    else return -1;
    if (P.lag_order == LAG_RASTER) return y >= 0 ? y : P.nposy + (y + P.nnegy);
    return y + P.nnegy;
}

// Signed lag y of stored lag row yl.
constexpr int DEAD_ROW = -0x40000000;

__host__ __device__ __forceinline__ int lag_y_of_row(const PairDev &P, int yl)
{
    if (P.lag_order == LAG_ALL) {
        if (yl < P.nposy) return yl;
        if (yl >= P.Ny - P.nnegy) return yl - P.Ny;
        return DEAD_ROW;
    }
    if (P.lag_order == LAG_RASTER) return yl < P.nposy ? yl : yl - P.nposy - P.nnegy;
    return yl - P.nnegy;
}

// Stored lag row yl is computed (its signed lag lies in the pair's live window).
__host__ __device__ __forceinline__ bool row_live(const PairDev &P, int yl)
{
    const int y = lag_y_of_row(P, yl);
    return y >= P.ylo && y <= P.yhi;
}

// Signed lag x of raster column ix, with *dead for the dead zone.
__device__ __forceinline__ int lag_x_of_raster(const PairDev &P, int ix, bool *dead)
{
    *dead = false;
    if (ix >= P.Nx - P.nnegx) return ix - P.Nx;
    if (ix >= P.nposx) { *dead = true; return 0; }
    return ix;
}

// ------------------------------------------------------------------
// Fill / pack / splat
// ------------------------------------------------------------------

template <typename T>
__global__ void k_fill(T *__restrict__ p, long long n, T v)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long s = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

// Bilinear splat of a point list into a compact w x h raster
// (DistributePixel, Maths.h:236-276; ImageFromValuesAndPoints,
// Geometry.cpp:661-685).  Device-wide atomics: exact for point lists in
// which no pixel receives more than two non-zero contributions (integer
// point lists without triplicates).
#pragma clang fp contract(off)
static __global__ void k_splat_points(const xc_point *__restrict__ pts, const double *__restrict__ val, int n,
                               double xmin, double ymin, double *__restrict__ img, int w, int h)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = pts[i].x - xmin, y = pts[i].y - ymin, v = val[i];
    int ix = (int)floor(x), iy = (int)floor(y);
    double alpha = x - ix, beta = y - iy;
    if (ix >= 0 && ix < w) {
        if (iy >= 0 && iy < h) atomicAdd(&img[ix + (long long)w * iy], (1.0 - alpha) * (1.0 - beta) * v);
        if (iy + 1 >= 0 && iy + 1 < h) atomicAdd(&img[ix + (long long)w * (iy + 1)], (1.0 - alpha) * beta * v);
    }
    if (ix + 1 >= 0 && ix + 1 < w) {
        if (iy >= 0 && iy < h) atomicAdd(&img[ix + 1 + (long long)w * iy], alpha * (1.0 - beta) * v);
        if (iy + 1 >= 0 && iy + 1 < h) atomicAdd(&img[ix + 1 + (long long)w * (iy + 1)], alpha * beta * v);
    }
}
#pragma clang fp contract(on)

// ------------------------------------------------------------------
// Source pixels
// ------------------------------------------------------------------

__device__ __forceinline__ double load_raw(const ImgView &v, int x, int y)
{
    const long long i = (long long)y * v.stride + x;
    switch (v.dtype) {
    case XC_U8: return (double)reinterpret_cast<const uint8_t *>(v.data)[i];
    case XC_U16: return (double)reinterpret_cast<const uint16_t *>(v.data)[i];
    case XC_F32: return (double)reinterpret_cast<const float *>(v.data)[i];
    default: return reinterpret_cast<const double *>(v.data)[i];
    }
}

__device__ __forceinline__ bool present(const ImgView &v, int x, int y)
{
    return !v.mask || v.mask[(long long)y * v.mstride + x] != 0;
}

// Value of pixel (x, y) in transform precision: (raw - mean) * (1/sd)
// rounded to T, 0 outside the image or where the mask says absent.
template <typename T>
__device__ __forceinline__ T load_val(const ImgView &v, double mean, double inv, int x, int y)
{
    if (x >= v.w || y >= v.h || !present(v, x, y)) return (T)0;
    return (T)((load_raw(v, x, y) - mean) * inv);
}

// Row y of a typed source as complex pairs (x = 2p, 2p + 1 with
// p = t + TPF m) for registers m < ME, the values load_val gives.
template <typename T, typename InT, int ME, int TPF>
__device__ __forceinline__ void load_row_pairs(cpx<T> *v, const ImgView &V, double mean, double inv, int t, int y,
                                               bool valid)
{
    const InT *row = reinterpret_cast<const InT *>(V.data) + (long long)y * V.stride;
    const uint8_t *mrow = V.mask ? V.mask + (long long)y * V.mstride : nullptr;
    if constexpr (sizeof(T) == 4 && sizeof(InT) == 1) {
        // FP32 form, unmasked u8 row at an even address: one 16-bit load per pixel
        // pair and an FP32 fma for the normalisation (within an FP32 rounding of
        // the FP64 value)
        if (!mrow && (reinterpret_cast<uintptr_t>(row) & 1) == 0) {
            const uint16_t *r2 = reinterpret_cast<const uint16_t *>(row);
            const float sc = (float)inv, of = (float)(-mean * inv);
#pragma unroll
            for (int m = 0; m < ME; ++m) {
                const int p = t + TPF * m, x = 2 * p;
                T a = (T)0, b = (T)0;
                if (valid && x + 1 < V.w) {
                    const unsigned int u = gld(r2, p);
                    a = fmaf((float)(u & 255u), sc, of);
                    b = fmaf((float)(u >> 8), sc, of);
                } else if (valid && x < V.w) {
                    a = fmaf((float)gld(row, x), sc, of);
                }
                v[m].x = a;
                v[m].y = b;
            }
            return;
        }
    }
#pragma unroll
    for (int m = 0; m < ME; ++m) {
        const int x = 2 * (t + TPF * m);
        T a = (T)0, b = (T)0;
        if (valid && x < V.w && (!mrow || gld(mrow, x))) a = (T)(((double)gld(row, x) - mean) * inv);
        if (valid && x + 1 < V.w && (!mrow || gld(mrow, x + 1))) b = (T)(((double)gld(row, x + 1) - mean) * inv);
        v[m].x = a;
        v[m].y = b;
    }
}

// ------------------------------------------------------------------
// Image preparation: statistics (Stats/Normalize, Maths.cpp:111-212) and
// the summed-area tables of IntegrateImage (Correlation.cpp:49-105), built
// in one write pass: 64-column strips scanned top to bottom by one
// wavefront, with the left-of-strip contribution carried in from per-strip
// row totals.
// ------------------------------------------------------------------

template <typename InT>
__device__ __forceinline__ void stats_rows(const ImgView &V, int y0, int y1, int w, int tid, double &a, double &b,
                                           double &c)
{
    for (int y = y0; y < y1; ++y) {
        const InT *row = reinterpret_cast<const InT *>(V.data) + (long long)y * V.stride;
        const uint8_t *mrow = V.mask ? V.mask + (long long)y * V.mstride : nullptr;
        for (int xb = 0; xb < w; xb += 256 * 8) {
            double v[8];
            bool p[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {  // eight loads in flight per thread
                const int x = xb + tid + 256 * k;
                p[k] = x < w && (!mrow || mrow[x]);
                v[k] = x < w ? (double)row[x] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (p[k]) {
                    a += v[k];
                    b += v[k] * v[k];
                    c += 1.0;
                }
        }
    }
}

// Unmasked u8 rows with 16-byte aligned starts: the band's whole 16-byte
// pieces (row, piece) dealt out over all threads, 16 pixels a load; the
// columns past the last whole piece byte by byte.  Integer sums (exact, as
// the FP64 sums of integer pixels are).
__device__ __forceinline__ void stats_rows_u8x16(const ImgView &V, int y0, int y1, int w, int tid, double &a, double &b,
                                                 double &c)
{
    const uint8_t *px = reinterpret_cast<const uint8_t *>(V.data);
    const long long st = V.stride;
    const int w16 = w / 16, tw = w - 16 * w16, rows = y1 - y0;
    unsigned long long sa = 0, sq = 0, n = 0;
    if (w16 > 0)
        for (int i = tid; i < rows * w16; i += 256) {
            const int y = y0 + i / w16, x = 16 * (i % w16);
            const uint4 q = gld16(px + (long long)y * st + x);
            const unsigned int wd[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const unsigned int v = (wd[j] >> (8 * k)) & 0xffu;
                    sa += v;
                    sq += v * v;
                }
            n += 16;
        }
    if (tw > 0)
        for (int i = tid; i < rows * tw; i += 256) {
            const int y = y0 + i / tw, x = 16 * w16 + i % tw;
            const unsigned int v = gld(px, (long long)y * st + x);
            sa += v;
            sq += v * v;
            n += 1;
        }
    a += (double)sa;
    b += (double)sq;
    c += (double)n;
}

// Masked u8 rows whose pixel and mask rows start on 16-byte boundaries: the
// band's whole 16-byte pieces (row, piece) are dealt out over all threads,
// 16 pixels and their 16 mask bytes a pair of loads; the columns past the
// last whole piece byte by byte.  Integer sums (exact).
__device__ __forceinline__ void stats_rows_u8x16m(const ImgView &V, int y0, int y1, int w, int tid, double &a,
                                                  double &b, double &c)
{
    const uint8_t *px = reinterpret_cast<const uint8_t *>(V.data), *mk = V.mask;
    const long long st = V.stride, ms = V.mstride;
    const int w16 = w / 16, tw = w - 16 * w16, rows = y1 - y0;
    unsigned long long sa = 0, sq = 0, n = 0;
    if (w16 > 0)
        for (int i = tid; i < rows * w16; i += 256) {
            const int y = y0 + i / w16, x = 16 * (i % w16);
            const uint4 q = gld16(px + (long long)y * st + x), m = gld16(mk + (long long)y * ms + x);
            const unsigned int wd[4] = {q.x, q.y, q.z, q.w}, wm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if ((wm[j] >> (8 * k)) & 0xffu) {
                        const unsigned int v = (wd[j] >> (8 * k)) & 0xffu;
                        sa += v;
                        sq += v * v;
                        n += 1;
                    }
        }
    if (tw > 0)
        for (int i = tid; i < rows * tw; i += 256) {
            const int y = y0 + i / tw, x = 16 * w16 + i % tw;
            if (gld(mk, (long long)y * ms + x)) {
                const unsigned int v = gld(px, (long long)y * st + x);
                sa += v;
                sq += v * v;
                n += 1;
            }
        }
    a += (double)sa;
    b += (double)sq;
    c += (double)n;
}

// u8 rows stored back to back (stride w, and the mask's too): the band
// [i0, i1) as one run of bytes, eight pixels (and mask bytes) in flight per
// thread whatever the row width, summed in integers (exact, as the FP64 sums
// of integer pixels are).
__device__ __forceinline__ void stats_flat_u8(const ImgView &V, long long i0, long long i1, int tid, double &a,
                                              double &b, double &c)
{
    const uint8_t *px = reinterpret_cast<const uint8_t *>(V.data), *mk = V.mask;
    unsigned long long sa = 0, sq = 0, n = 0;
    for (long long ib = i0; ib < i1; ib += 256 * 8) {
        unsigned v[8];
        bool p[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long long i = ib + tid + 256 * k;
            p[k] = i < i1;
            v[k] = p[k] ? gld(px, i) : 0u;
            if (mk) p[k] = p[k] && gld(mk, i) != 0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (p[k]) {
                sa += v[k];
                sq += v[k] * v[k];
                n += 1;
            }
    }
    a += (double)sa;
    b += (double)sq;
    c += (double)n;
}

// u8 rows at any stride (and mask rows at theirs): threads stride along x, each
// load step takes eight rows at once, summed in integers (exact).
__device__ __forceinline__ void stats_rows8_u8(const ImgView &V, int y0, int y1, int w, int tid, double &a, double &b,
                                               double &c)
{
    const uint8_t *px = reinterpret_cast<const uint8_t *>(V.data), *mk = V.mask;
    const long long st = V.stride, ms = V.mstride;
    unsigned long long sa = 0, sq = 0, n = 0;
    for (int y = y0; y < y1; y += 8)
        for (int x = tid; x < w; x += 256) {
            unsigned v[8];
            bool p[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                p[k] = y + k < y1;
                v[k] = p[k] ? gld(px, (long long)(y + k) * st + x) : 0u;
                if (mk) p[k] = p[k] && gld(mk, (long long)(y + k) * ms + x) != 0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (p[k]) {
                    sa += v[k];
                    sq += v[k] * v[k];
                    n += 1;
                }
        }
    a += (double)sa;
    b += (double)sq;
    c += (double)n;
}

// Partial sums of present raw values: TP_PARTS workgroups per image, each
// over a band of rows; threads stride along x (coalesced).
static __global__ void __launch_bounds__(256) k_tp_stats(const TileDev *__restrict__ tiles)
{
    __shared__ double ss[256], sq[256], sc[256];
    const TileDev &D = tiles[blockIdx.x / TP_PARTS];
    const int part = blockIdx.x % TP_PARTS, tid = threadIdx.x;
    const int h = D.v.h, w = D.v.w;
    const int y0 = (int)((long long)h * part / TP_PARTS), y1 = (int)((long long)h * (part + 1) / TP_PARTS);
    double a = 0.0, b = 0.0, c = 0.0;
    const bool wide = D.v.dtype == XC_U8 && !D.v.mask && (reinterpret_cast<uintptr_t>(D.v.data) & 15) == 0 &&
                      (D.v.stride & 15) == 0;
    const bool wide_m = D.v.dtype == XC_U8 && D.v.mask &&
                        ((reinterpret_cast<uintptr_t>(D.v.data) | reinterpret_cast<uintptr_t>(D.v.mask) |
                          (uintptr_t)D.v.stride | (uintptr_t)D.v.mstride) & 15) == 0;
    const bool flat = D.v.dtype == XC_U8 && D.v.stride == w && (!D.v.mask || D.v.mstride == w);
    if (wide) stats_rows_u8x16(D.v, y0, y1, w, tid, a, b, c);
    else if (wide_m) stats_rows_u8x16m(D.v, y0, y1, w, tid, a, b, c);
    else if (flat) stats_flat_u8(D.v, (long long)y0 * w, (long long)y1 * w, tid, a, b, c);
    else if (D.v.dtype == XC_U8) stats_rows8_u8(D.v, y0, y1, w, tid, a, b, c);
    else switch (D.v.dtype) {  // uniform per workgroup
    case XC_U8: stats_rows<uint8_t>(D.v, y0, y1, w, tid, a, b, c); break;
    case XC_U16: stats_rows<uint16_t>(D.v, y0, y1, w, tid, a, b, c); break;
    case XC_F32: stats_rows<float>(D.v, y0, y1, w, tid, a, b, c); break;
    default: stats_rows<double>(D.v, y0, y1, w, tid, a, b, c); break;
    }
    ss[tid] = a;
    sq[tid] = b;
    sc[tid] = c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            ss[tid] += ss[tid + off];
            sq[tid] += sq[tid + off];
            sc[tid] += sc[tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        D.part[part * 3 + 0] = ss[0];
        D.part[part * 3 + 1] = sq[0];
        D.part[part * 3 + 2] = sc[0];
    }
}

// (mean, sd) per image: MeanStd::Stats with the n-1 deviation, or (0, 1).
static __global__ void k_tp_stats_fin(const TileDev *__restrict__ tiles, int ntiles)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles) return;
    const TileDev &D = tiles[i];
    double mean = 0.0, sd = 1.0;
    if (D.normalize) {
        double s = 0.0, q = 0.0, c = 0.0;
        for (int p = 0; p < TP_PARTS; ++p) {
            s += D.part[p * 3 + 0];
            q += D.part[p * 3 + 1];
            c += D.part[p * 3 + 2];
        }
        mean = s;
        sd = 0.0;
        if (c > 1.0) {
            mean = s / c;
            sd = sqrt(fmax(q - c * mean * mean, 0.0) / (c - 1.0));
        }
    }
    D.msd[0] = mean;
    D.msd[1] = sd;
    D.msd[2] = 1.0 / sd;
}

// Wavefront inclusive scans with DPP row shifts and row broadcasts (GFX9
// data-parallel primitives: no LDS round trip, unlike a lane permute).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xf, true);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL, ROWMASK>((int)(b & 0xffffffffLL));
    const int hi = dpp_i<CTRL, ROWMASK>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int wave_scan_i(int v)
{
    v += dpp_i<0x111, 0xf>(v);  // row_shr:1
    v += dpp_i<0x112, 0xf>(v);  // row_shr:2
    v += dpp_i<0x114, 0xf>(v);  // row_shr:4
    v += dpp_i<0x118, 0xf>(v);  // row_shr:8
    v += dpp_i<0x142, 0xa>(v);  // row_bcast:15 -> rows 1, 3
    v += dpp_i<0x143, 0xc>(v);  // row_bcast:31 -> rows 2, 3
    return v;
}

__device__ __forceinline__ double wave_scan_d(double v)
{
    v += dpp_d<0x111, 0xf>(v);
    v += dpp_d<0x112, 0xf>(v);
    v += dpp_d<0x114, 0xf>(v);
    v += dpp_d<0x118, 0xf>(v);
    v += dpp_d<0x142, 0xa>(v);
    v += dpp_d<0x143, 0xc>(v);
    return v;
}

constexpr int SAT_MAX_THREADS = 1024;

// Sum over lanes 0..15 of the wavefront, returned to every lane (DPP
// inclusive scan within the first 16-lane row, then a read of lane 15).
__device__ __forceinline__ double row16_sum_d(double v)
{
    v += dpp_d<0x111, 0xf>(v);
    v += dpp_d<0x112, 0xf>(v);
    v += dpp_d<0x114, 0xf>(v);
    v += dpp_d<0x118, 0xf>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), 15);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 15);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int row16_sum_i(int v)
{
    v += dpp_i<0x111, 0xf>(v);
    v += dpp_i<0x112, 0xf>(v);
    v += dpp_i<0x114, 0xf>(v);
    v += dpp_i<0x118, 0xf>(v);
    return __builtin_amdgcn_readlane(v, 15);
}

// Typed pixel read for the table sweep (no per-pixel dtype switch).
template <typename InT>
__device__ __forceinline__ double sweep_val(const ImgView &v, int x, int y)
{
    return (double)reinterpret_cast<const InT *>(v.data)[(long long)y * v.stride + x];
}

// Row bands of the table sweep: a lone image of h rows (the plain path's
// two) is swept as up to SAT_NB bands of >= 32 rows, each by its own
// workgroup; a band starts from the column sums of the bands above it
// (k_tp_colsum).  Batches of images fill the device one workgroup per image
// and sweep unbanded (nbands = 1), which moves fewer bytes.
constexpr int SAT_NB = 8;
__host__ __device__ inline int sat_bands(int h)
{
    const int nb = (h + 31) / 32;
    return nb < 1 ? 1 : (nb > SAT_NB ? SAT_NB : nb);
}
__host__ __device__ inline int tile_bands(const TileDev &D) { return D.nbands > 1 ? D.nbands : 1; }
__host__ __device__ inline int tile_band_rows(const TileDev &D) { return (D.v.h + tile_bands(D) - 1) / tile_bands(D); }

// Column-sum scratch of a tile: [SAT_NB][w] S, [SAT_NB][w] Q, then [SAT_NB][w] int Z.
__device__ __forceinline__ double *bandp_s(const TileDev &D, int b) { return D.bandp + (long long)b * D.v.w; }
__device__ __forceinline__ double *bandp_q(const TileDev &D, int b) { return D.bandp + (long long)(SAT_NB + b) * D.v.w; }
__device__ __forceinline__ int *bandp_z(const TileDev &D, int b)
{
    return reinterpret_cast<int *>(D.bandp + 2LL * SAT_NB * D.v.w) + (long long)b * D.v.w;
}

// The value the transforms see: (raw - mean) * (1/sd) rounded to T; absent pixels are 0.
template <typename T, typename InT, bool MASK>
__device__ __forceinline__ double sat_val(const TileDev &D, double raw, int x, int y, double mean, double inv)
{
    const bool pres = x < D.v.w && (!MASK || D.v.mask[(long long)y * D.v.mstride + x]);
    return pres ? (double)(T)((raw - mean) * inv) : 0.0;
}

// Column sums of rows [y0, y1) over the thread's CPT columns.
template <typename T, int CPT, typename InT, bool MASK>
__device__ __forceinline__ void band_colsum(const TileDev &D, int band, int y0, int y1)
{
    const int w = D.v.w, x0 = threadIdx.x * CPT;
    const double mean = D.v.msd[0], inv = D.v.msd[2];
    double cs[CPT], cq[CPT];
    int cz[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        cs[j] = 0.0;
        cq[j] = 0.0;
        cz[j] = 0;
    }
#pragma unroll 4
    for (int y = y0; y < y1; ++y) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int x = x0 + j;
            const double t = sat_val<T, InT, MASK>(D, x < w ? sweep_val<InT>(D.v, x, y) : 0.0, x, y, mean, inv);
            cs[j] += t;
            cq[j] += t * t;
            cz[j] += (t != 0.0);
        }
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const int x = x0 + j;
        if (x < w) {
            bandp_s(D, band)[x] = cs[j];
            bandp_q(D, band)[x] = cq[j];
            bandp_z(D, band)[x] = cz[j];
        }
    }
}

template <typename T, typename TB, int CPT, int NT, typename InT, bool MASK>
__device__ __forceinline__ void sat_sweep(const TileDev &D, int band, double *xs, double *xq, int *xz)
{
    constexpr int NWV = NT / 64;
    const int w = D.v.w, h = D.v.h, bh = tile_band_rows(D);
    const int y0 = band * bh, y1 = min(h, y0 + bh);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = tid * CPT;
    const double mean = D.v.msd[0], inv = D.v.msd[2];
    double cs[CPT], cq[CPT];
    int cz[CPT];
    constexpr int LA = 4;  // rows loaded ahead of the sweep
    double nxt[LA][CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        // carry: the column sums of the bands above, added in band order
        cs[j] = 0.0;
        cq[j] = 0.0;
        cz[j] = 0;
        const int x = x0 + j;
        if (x < w)
            for (int b = 0; b < band; ++b) {
                cs[j] += bandp_s(D, b)[x];
                cq[j] += bandp_q(D, b)[x];
                cz[j] += bandp_z(D, b)[x];
            }
#pragma unroll
        for (int a = 0; a < LA; ++a) nxt[a][j] = (x < w && y0 + a < y1) ? sweep_val<InT>(D.v, x, y0 + a) : 0.0;
    }
    const bool pair_st = (CPT % 2 == 0) && (w % 2 == 0);  // 16-byte aligned column pairs
    const bool wz = D.Z != nullptr;                        // count table wanted (uniform)
    static_assert(NWV <= 16, "the cross-wave prefix reads one wave total per lane of a 16-lane row");
    for (int y = y0; y < y1; ++y) {
        double ps[CPT], pq[CPT];
        int pz[CPT];
        double cur[CPT];
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            cur[j] = nxt[0][j];
#pragma unroll
            for (int a = 0; a + 1 < LA; ++a) nxt[a][j] = nxt[a + 1][j];
            const int x = x0 + j;
            nxt[LA - 1][j] = (x < w && y + LA < y1) ? sweep_val<InT>(D.v, x, y + LA) : 0.0;
        }
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int x = x0 + j;
            const double t = sat_val<T, InT, MASK>(D, cur[j], x, y, mean, inv);
            cs[j] += t;
            cq[j] += t * t;
            cz[j] += (t != 0.0);
            ps[j] = (j ? ps[j - 1] : 0.0) + cs[j];
            pq[j] = (j ? pq[j - 1] : 0.0) + cq[j];
            pz[j] = (j ? pz[j - 1] : 0) + cz[j];
        }
        const double ts = ps[CPT - 1], tq = pq[CPT - 1];
        const int tz = pz[CPT - 1];
        const double is = wave_scan_d(ts), iq = wave_scan_d(tq);
        const int iz = wz ? wave_scan_i(tz) : 0;
        const int par = (y & 1) * NWV;
        if (lane == 63) {
            xs[par + wv] = is;
            xq[par + wv] = iq;
            xz[par + wv] = iz;
        }
        __syncthreads();
        double bs = is - ts, bq = iq - tq;
        int bz = iz - tz;
        if constexpr (NWV > 1) {
            // totals of the waves to the left: lane k < wv holds wave k's, summed over the first 16 lanes
            const bool left = lane < wv;  // wv < NWV, so k stays inside the row's NWV slots
            const int k = par + (left ? lane : 0);
            bs += row16_sum_d(left ? xs[k] : 0.0);
            bq += row16_sum_d(left ? xq[k] : 0.0);
            if (wz) bz += row16_sum_i(left ? xz[k] : 0);
        }
        const long long row = (long long)y * w;
        if (pair_st) {
#pragma unroll
            for (int j = 0; j < CPT; j += 2) {
                if (x0 + j < w) {  // w even: x0 + j + 1 < w too
                    if (D.SQ) {
                        tab_st<TB>(D.SQ, row + x0 + j, bs + ps[j], bq + pq[j]);
                        tab_st<TB>(D.SQ, row + x0 + j + 1, bs + ps[j + 1], bq + pq[j + 1]);
                    }
                    if (D.Z) gst2(D.Z, row + x0 + j, bz + pz[j], bz + pz[j + 1]);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j) {
                if (x0 + j < w) {
                    if (D.SQ) tab_st<TB>(D.SQ, row + x0 + j, bs + ps[j], bq + pq[j]);
                    if (D.Z) gst(D.Z, row + x0 + j, bz + pz[j]);
                }
            }
        }
    }
}


// Exact integer tables (TabX) of an unmasked u8 tile.  Each thread owns CPT
// consecutive columns; a row's CPT pixels arrive in one load, issued LA rows
// ahead so that the store traffic of the rows between does not stall the
// sweep.  The rows' integer prefixes (v, v^2 [, non-zero]) are scanned across
// the row (DPP wavefront scans + a cross-wave exchange of RB rows at a time,
// one barrier per exchange) and
// each column's packed running sum (Sraw << 36) + Qraw gains the row's prefix;
// a row prefix fits 32 bits (w <= TP_MAX_W).  WZ: the count table too
// (pixels whose value rounded to T is non-zero, as the other sweeps count them).
template <int CPT> struct PixWord {
    typedef uint32_t type;
};
template <> struct PixWord<8> {
    typedef unsigned long long type;
};

template <int CPT>
__device__ __forceinline__ typename PixWord<CPT>::type load_pix(const uint8_t *row, int x0, int w, bool vec)
{
    typedef typename PixWord<CPT>::type W;
    if (vec) {
        if constexpr (CPT == 2) return gld(reinterpret_cast<const uint16_t *>(row + x0), 0);
        else return gld(reinterpret_cast<const W *>(row + x0), 0);
    }
    W v = 0;
#pragma unroll
    for (int j = 0; j < CPT; ++j)
        if (x0 + j < w) v |= (W)gld(row, x0 + j) << (8 * j);
    return v;
}

constexpr int SATX_RB = 4;           // rows per barrier of the exact sweep
constexpr int SATX_BLK = 32768;      // bytes of pixel rows an exact sweep stages in LDS at a time
constexpr int SATX_PER = 4;          // 16-byte pieces of a staged block per thread (at most)

// RB rows of the exact sweep from each thread's pixel words pix[r] (rows y0 + r,
// rows at or past h contribute nothing): integer prefixes across the row, one
// cross-wave exchange, packed running sums, stores.
template <typename T, int CPT, int NT, bool WZ, typename W>
__device__ __forceinline__ void satx_rows(const TileDev &D, int *xi, int y0, const W (&pix)[SATX_RB],
                                          unsigned long long (&acc)[CPT], int (&az)[CPT], double mean, double inv)
{
    constexpr int NWV = NT / 64;
    constexpr int NQ = WZ ? 3 : 2;  // S, Q [, non-zero]
    constexpr int RB = SATX_RB;
    const int w = D.v.w, h = D.v.h;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = tid * CPT;
    unsigned long long *sq = reinterpret_cast<unsigned long long *>(D.SQ);
    int pf[RB][NQ][CPT];  // per-thread row prefixes of RB rows
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int y = y0 + r;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int v = (y < h && x0 + j < w) ? (int)((pix[r] >> (8 * j)) & 255) : 0;
            pf[r][0][j] = (j ? pf[r][0][j - 1] : 0) + v;
            pf[r][1][j] = (j ? pf[r][1][j - 1] : 0) + v * v;
            if constexpr (WZ) {
                const bool nz = y < h && x0 + j < w && (T)(((double)v - mean) * inv) != (T)0;
                pf[r][2][j] = (j ? pf[r][2][j - 1] : 0) + (nz ? 1 : 0);
            }
        }
    }
    int base[RB][NQ];
    const int par = ((y0 / RB) & 1) * RB * NQ * NWV;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) {
            const int tot = pf[r][qi][CPT - 1];
            const int inc = wave_scan_i(tot);
            if (lane == 63) xi[par + (r * NQ + qi) * NWV + wv] = inc;
            base[r][qi] = inc - tot;
        }
    __syncthreads();
    if constexpr (NWV > 1) {
        const bool left = lane < wv;
        const int k = left ? lane : 0;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int qi = 0; qi < NQ; ++qi) base[r][qi] += row16_sum_i(left ? xi[par + (r * NQ + qi) * NWV + k] : 0);
    }
    const bool pair_st = (w % 2) == 0;  // 16-byte aligned column pairs
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int y = y0 + r;
        if (y >= h) break;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            acc[j] += ((unsigned long long)(unsigned)(base[r][0] + pf[r][0][j]) << TABX_QB) +
                      (unsigned long long)(unsigned)(base[r][1] + pf[r][1][j]);
            if constexpr (WZ) az[j] += base[r][2] + pf[r][2][j];
        }
        const long long row = (long long)y * w;
        if (pair_st) {
#pragma unroll
            for (int j = 0; j < CPT; j += 2)
                if (x0 + j < w) {  // w even: x0 + j + 1 < w too
                    gst2(sq, row + x0 + j, acc[j], acc[j + 1]);
                    if constexpr (WZ) gst2(D.Z, row + x0 + j, az[j], az[j + 1]);
                }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j)
                if (x0 + j < w) {
                    gst(sq, row + x0 + j, acc[j]);
                    if constexpr (WZ) gst(D.Z, row + x0 + j, az[j]);
                }
        }
    }
}

// Rows of the exact sweep staged through LDS: when rows are 16-byte aligned
// whole pieces (w, stride and the base address multiples of 16), the
// workgroup loads a block of BR rows with 16-byte loads, issued while the
// block before is swept, so that a wait for pixels waits only for the stores
// of the block before (not for those of the last few rows, as per-thread
// loads a few rows ahead would); else each thread loads its own pixels LA rows
// ahead.
template <typename T, int CPT, int NT, bool WZ>
__device__ __forceinline__ void sat_sweep_x(const TileDev &D, int *xi, uint4 *blk)
{
    typedef typename PixWord<CPT>::type W;
    constexpr int LA = 16;        // rows loaded ahead by the per-thread path
    constexpr int RB = SATX_RB;
    static_assert(NT / 64 <= 16, "the cross-wave prefix reads one wave total per lane of a 16-lane row");
    static_assert(CPT == 2 || CPT == 4 || CPT == 8, "one load per row");
    const int w = D.v.w, h = D.v.h;
    const int tid = threadIdx.x;
    const int x0 = tid * CPT;
    const uint8_t *src = reinterpret_cast<const uint8_t *>(D.v.data);
    const long long stride = D.v.stride;
    const double mean = D.v.msd[0], inv = D.v.msd[2];
    unsigned long long acc[CPT];
    int az[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        acc[j] = 0;
        az[j] = 0;
    }
    const bool staged = blk && ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)stride | (uintptr_t)w) & 15) == 0;
    if (staged) {
        const int cpr = w / 16;  // 16-byte pieces a row
        int br = SATX_BLK / w;   // rows a block: the LDS block, and at most SATX_PER pieces a thread
        br = br < (SATX_PER * NT) / cpr ? br : (SATX_PER * NT) / cpr;
        br = br > 32 ? 32 : br;
        br = br / RB * RB;  // whole barrier groups (w <= TP_MAX_W = 8192 keeps br >= 4)
        const int np = br * cpr;
        uint4 pf[SATX_PER];
        auto fetch = [&](int yb) {
#pragma unroll
            for (int k = 0; k < SATX_PER; ++k) {
                const int c = tid + NT * k, r = c / cpr;
                pf[k] = (c < np && yb + r < h)
                            ? gld16(src + (long long)(yb + r) * stride + 16 * (c - r * cpr))
                            : make_uint4(0, 0, 0, 0);
            }
        };
        fetch(0);
        for (int yb = 0; yb < h; yb += br) {
            __syncthreads();  // the block before is consumed
#pragma unroll
            for (int k = 0; k < SATX_PER; ++k) {
                const int c = tid + NT * k;
                if (c < np) blk[c] = pf[k];
            }
            __syncthreads();
            if (yb + br < h) fetch(yb + br);  // in flight while this block is swept
            const uint8_t *b = reinterpret_cast<const uint8_t *>(blk);
            for (int r0 = 0; r0 < br && yb + r0 < h; r0 += RB) {
                W pix[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    const uint8_t *p = b + (r0 + r) * w + x0;
                    if constexpr (CPT == 2) pix[r] = x0 < w ? *reinterpret_cast<const uint16_t *>(p) : 0;
                    else pix[r] = x0 < w ? *reinterpret_cast<const W *>(p) : 0;
                }
                satx_rows<T, CPT, NT, WZ, W>(D, xi, yb + r0, pix, acc, az, mean, inv);
            }
        }
        return;
    }
    const bool vec = x0 + CPT <= w && ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)stride) % CPT) == 0;
    W nxt[LA];
#pragma unroll
    for (int a = 0; a < LA; ++a) nxt[a] = (a < h && x0 < w) ? load_pix<CPT>(src + a * stride, x0, w, vec) : 0;
    for (int y0 = 0; y0 < h; y0 += RB) {
        W pix[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int y = y0 + r;
            pix[r] = nxt[0];
#pragma unroll
            for (int a = 0; a + 1 < LA; ++a) nxt[a] = nxt[a + 1];
            nxt[LA - 1] = (y + LA < h && x0 < w) ? load_pix<CPT>(src + (y + LA) * stride, x0, w, vec) : 0;
        }
        satx_rows<T, CPT, NT, WZ, W>(D, xi, y0, pix, acc, az, mean, inv);
    }
}

// FP64 (or FP32) (S, S^2) tables of u8 tiles: the rows are scanned in exact
// integer arithmetic (a row's prefix of v and v^2 fits in 32 bits for
// w <= TP_MAX_W), the raw sums accumulate down the columns in FP64 (exact
// integers), and each stored entry is normalised from them over the n present
// pixels of the box:
//   S = (Sraw - n mean) / sd,  Q = (Qraw - 2 mean Sraw + n mean^2) / sd^2
// (unmasked: n = (x+1)(y+1); masked: a third integer scan counts the present
// pixels).  This is the exact sum of the normalised values; the general sweep
// sums the values already rounded to T, which differs by the accumulated
// rounding of T (relative 1e-8 for FP32), far below the transforms' own
// error.  WZ: the count table (pixels whose value rounded to T is non-zero,
// as the general sweep counts them) is scanned and written too.  Integer DPP
// scans are one instruction per step, against three for an FP64 scan.  RB
// rows share one cross-wave exchange; pix / pm: the thread's pixel and mask
// bytes of each row.
template <typename T, typename TB, int CPT, int NT, bool MASK, bool WZ, typename W>
__device__ __forceinline__ void satu8_rows(const TileDev &D, int *xi, int y0, const W (&pix)[SATX_RB],
                                           const W (&pm)[SATX_RB], double (&as)[CPT], double (&aq)[CPT],
                                           double (&an)[CPT], int (&az)[CPT], double mean, double inv)
{
    constexpr int NWV = NT / 64;
    constexpr int NQ = 2 + (MASK ? 1 : 0) + (WZ ? 1 : 0);  // S, Q [, present] [, non-zero]
    constexpr int RB = SATX_RB;
    const int w = D.v.w, h = D.v.h;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = tid * CPT;
    const double inv2 = inv * inv, m2 = 2.0 * mean, msq = mean * mean;
    int pf[RB][NQ][CPT];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int y = y0 + r;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            bool pres = y < h && x0 + j < w;
            if constexpr (MASK) pres = pres && ((pm[r] >> (8 * j)) & 255) != 0;
            const int v = pres ? (int)((pix[r] >> (8 * j)) & 255) : 0;
            pf[r][0][j] = (j ? pf[r][0][j - 1] : 0) + v;
            pf[r][1][j] = (j ? pf[r][1][j - 1] : 0) + v * v;
            if constexpr (MASK) pf[r][2][j] = (j ? pf[r][2][j - 1] : 0) + (pres ? 1 : 0);
            if constexpr (WZ) {
                const bool nz = pres && (T)(((double)v - mean) * inv) != (T)0;
                pf[r][NQ - 1][j] = (j ? pf[r][NQ - 1][j - 1] : 0) + (nz ? 1 : 0);
            }
        }
    }
    int base[RB][NQ];
    const int par = ((y0 / RB) & 1) * RB * NQ * NWV;
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int qi = 0; qi < NQ; ++qi) {
            const int tot = pf[r][qi][CPT - 1];
            const int inc = wave_scan_i(tot);
            if (lane == 63) xi[par + (r * NQ + qi) * NWV + wv] = inc;
            base[r][qi] = inc - tot;
        }
    __syncthreads();
    if constexpr (NWV > 1) {
        const bool left = lane < wv;
        const int k = left ? lane : 0;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int qi = 0; qi < NQ; ++qi) base[r][qi] += row16_sum_i(left ? xi[par + (r * NQ + qi) * NWV + k] : 0);
    }
    const bool pair_st = (CPT % 2 == 0) && (w % 2 == 0);
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const int y = y0 + r;
        if (y >= h) break;
        const long long row = (long long)y * w;
        double sv[CPT], qv[CPT];
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            as[j] += (double)(base[r][0] + pf[r][0][j]);
            aq[j] += (double)(base[r][1] + pf[r][1][j]);
            double n;
            if constexpr (MASK) {
                an[j] += (double)(base[r][2] + pf[r][2][j]);
                n = an[j];
            } else {
                n = (double)(x0 + j + 1) * (double)(y + 1);
            }
            if constexpr (WZ) az[j] += base[r][NQ - 1] + pf[r][NQ - 1][j];
            sv[j] = (as[j] - n * mean) * inv;
            qv[j] = ((aq[j] - m2 * as[j]) + n * msq) * inv2;
        }
        if (pair_st) {
#pragma unroll
            for (int j = 0; j < CPT; j += 2)
                if (x0 + j < w) {
                    tab_st<TB>(D.SQ, row + x0 + j, sv[j], qv[j]);
                    tab_st<TB>(D.SQ, row + x0 + j + 1, sv[j + 1], qv[j + 1]);
                    if constexpr (WZ) gst2(D.Z, row + x0 + j, az[j], az[j + 1]);
                }
        } else {
#pragma unroll
            for (int j = 0; j < CPT; ++j)
                if (x0 + j < w) {
                    tab_st<TB>(D.SQ, row + x0 + j, sv[j], qv[j]);
                    if constexpr (WZ) gst(D.Z, row + x0 + j, az[j]);
                }
        }
    }
}

// The u8 sweep's pixel (and mask) rows: staged through LDS in blocks when
// rows are 16-byte aligned whole pieces (as sat_sweep_x stages them; the
// block holds the pixel rows in its first half and the mask rows in its
// second), else loaded by each thread LA rows ahead.
template <typename T, typename TB, int CPT, int NT, bool MASK, bool WZ>
__device__ __forceinline__ void sat_sweep_u8(const TileDev &D, int *xi, uint4 *blk)
{
    typedef typename PixWord<CPT>::type W;
    constexpr int LA = 16;
    constexpr int RB = SATX_RB;
    constexpr int PERH = SATX_PER / 2;  // 16-byte pieces a thread fetches per block, of pixels and of mask
    static_assert(NT / 64 <= 16, "the cross-wave prefix reads one wave total per lane of a 16-lane row");
    static_assert(CPT == 2 || CPT == 4 || CPT == 8, "one load per row");
    const int w = D.v.w, h = D.v.h;
    const int tid = threadIdx.x;
    const int x0 = tid * CPT;
    const uint8_t *src = reinterpret_cast<const uint8_t *>(D.v.data);
    const uint8_t *msk = D.v.mask;
    const long long stride = D.v.stride, mstride = D.v.mstride;
    const double mean = D.v.msd[0], inv = D.v.msd[2];
    double as[CPT], aq[CPT], an[CPT];
    int az[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        as[j] = 0.0;
        aq[j] = 0.0;
        an[j] = 0.0;
        az[j] = 0;
    }
    // rows starting at 16-byte boundaries with 16-byte multiple strides: a row is read as its cpr whole
    // 16-byte pieces (the last may run into the row's padding, which the stride holds and the sweep
    // ignores: columns >= w count as absent), staged with an LDS pitch of 16 cpr bytes
    uintptr_t al = reinterpret_cast<uintptr_t>(src) | (uintptr_t)stride;
    if constexpr (MASK) al |= reinterpret_cast<uintptr_t>(msk) | (uintptr_t)mstride;
    const int cpr = (w + 15) / 16;  // 16-byte pieces a row
    const int pitch = 16 * cpr;
    int br = (SATX_BLK / 2) / pitch;
    br = cpr > 0 && br > (PERH * NT) / cpr ? (PERH * NT) / cpr : br;
    br = br > 32 ? 32 : br;
    br = br / RB * RB;  // whole barrier groups; 0 for rows wider than the half block holds RB of
    if (blk && (al & 15) == 0 && br > 0) {
        const int np = br * cpr;
        uint8_t *bm = reinterpret_cast<uint8_t *>(blk) + SATX_BLK / 2;
        uint4 pf[PERH], pq[PERH];
        auto fetch = [&](int yb) {
#pragma unroll
            for (int k = 0; k < PERH; ++k) {
                const int c = tid + NT * k, r = c / cpr, q = c - r * cpr;
                const bool in = c < np && yb + r < h;
                pf[k] = in ? gld16(src + (long long)(yb + r) * stride + 16 * q)
                           : make_uint4(0, 0, 0, 0);
                if constexpr (MASK)
                    pq[k] = in ? gld16(msk + (long long)(yb + r) * mstride + 16 * q)
                               : make_uint4(0, 0, 0, 0);
            }
        };
        fetch(0);
        for (int yb = 0; yb < h; yb += br) {
            __syncthreads();  // the block before is consumed
#pragma unroll
            for (int k = 0; k < PERH; ++k) {
                const int c = tid + NT * k;
                if (c < np) {
                    blk[c] = pf[k];
                    if constexpr (MASK) reinterpret_cast<uint4 *>(bm)[c] = pq[k];
                }
            }
            __syncthreads();
            if (yb + br < h) fetch(yb + br);  // in flight while this block is swept
            const uint8_t *b = reinterpret_cast<const uint8_t *>(blk);
            for (int r0 = 0; r0 < br && yb + r0 < h; r0 += RB) {
                W pix[RB], pmw[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    const int o = (r0 + r) * pitch + x0;
                    if constexpr (CPT == 2) {
                        pix[r] = x0 < w ? *reinterpret_cast<const uint16_t *>(b + o) : 0;
                        pmw[r] = (MASK && x0 < w) ? *reinterpret_cast<const uint16_t *>(bm + o) : 0;
                    } else {
                        pix[r] = x0 < w ? *reinterpret_cast<const W *>(b + o) : 0;
                        pmw[r] = (MASK && x0 < w) ? *reinterpret_cast<const W *>(bm + o) : 0;
                    }
                }
                satu8_rows<T, TB, CPT, NT, MASK, WZ, W>(D, xi, yb + r0, pix, pmw, as, aq, an, az, mean, inv);
            }
        }
        return;
    }
    const bool vec = x0 + CPT <= w && ((reinterpret_cast<uintptr_t>(src) | (uintptr_t)stride) % CPT) == 0;
    const bool mvec =
        !MASK || (x0 + CPT <= w && ((reinterpret_cast<uintptr_t>(msk) | (uintptr_t)mstride) % CPT) == 0);
    W nxt[LA], nxm[LA];
#pragma unroll
    for (int a = 0; a < LA; ++a) {
        nxt[a] = (a < h && x0 < w) ? load_pix<CPT>(src + a * stride, x0, w, vec) : 0;
        if constexpr (MASK) nxm[a] = (a < h && x0 < w) ? load_pix<CPT>(msk + a * mstride, x0, w, mvec) : 0;
        else nxm[a] = 0;
    }
    for (int y0 = 0; y0 < h; y0 += RB) {
        W pix[RB], pmw[RB];
#pragma unroll
        for (int r = 0; r < RB; ++r) {
            const int y = y0 + r;
            pix[r] = nxt[0];
            pmw[r] = nxm[0];
#pragma unroll
            for (int a = 0; a + 1 < LA; ++a) {
                nxt[a] = nxt[a + 1];
                nxm[a] = nxm[a + 1];
            }
            nxt[LA - 1] = (y + LA < h && x0 < w) ? load_pix<CPT>(src + (y + LA) * stride, x0, w, vec) : 0;
            if constexpr (MASK)
                nxm[LA - 1] = (y + LA < h && x0 < w) ? load_pix<CPT>(msk + (y + LA) * mstride, x0, w, mvec) : 0;
        }
        satu8_rows<T, TB, CPT, NT, MASK, WZ, W>(D, xi, y0, pix, pmw, as, aq, an, az, mean, inv);
    }
}

#define XC_SAT_DISPATCH(CALL)                                                   \
    switch (D.v.dtype) {                                                        \
    case XC_U8:                                                                 \
        if (m) CALL(uint8_t, true); else CALL(uint8_t, false);                  \
        break;                                                                  \
    case XC_U16:                                                                \
        if (m) CALL(uint16_t, true); else CALL(uint16_t, false);                \
        break;                                                                  \
    case XC_F32:                                                                \
        if (m) CALL(float, true); else CALL(float, false);                      \
        break;                                                                  \
    default:                                                                    \
        if (m) CALL(double, true); else CALL(double, false);                    \
        break;                                                                  \
    }

// Column sums of every band but the last (the carries of the bands below).
template <typename T, int CPT, int NT>
__global__ void __launch_bounds__(NT) k_tp_colsum(const TileDev *__restrict__ tiles, const int *__restrict__ list, int gnb)
{
    const TileDev D = tiles[list[blockIdx.x / gnb]];
    const int band = blockIdx.x % gnb, nb = tile_bands(D), bh = tile_band_rows(D);
    if (band >= nb - 1) return;
    const bool m = D.v.mask != nullptr;
    const int y0 = band * bh, y1 = min(D.v.h, y0 + bh);
#define CS_CALL(InT, MK) band_colsum<T, CPT, InT, MK>(D, band, y0, y1)
    XC_SAT_DISPATCH(CS_CALL)
#undef CS_CALL
}

// Summed-area tables (IntegrateImage, Correlation.cpp:49-105): one
// workgroup per row band walks its rows top to bottom.  Each thread owns
// CPT consecutive columns and keeps their running column sums in registers
// (starting from the bands above, the next rows loaded ahead); a DPP
// wavefront scan plus a cross-wave exchange through LDS (double-buffered by
// row parity, one barrier per row) turns column sums into the 2-D prefix.
// Writes S, Q, Z once.
template <typename T, typename TB, int CPT, int NT>
__device__ __forceinline__ void sat_item(const TileDev &D, int band, double *xs, double *xq, int *xz, int *xi,
                                         uint4 *xblk)
{
    if (band >= tile_bands(D)) return;
    if constexpr (is_tabx<TB>()) {
        // the host gives exact tables to unmasked u8 tiles only, swept unbanded
        if (D.Z) sat_sweep_x<T, CPT, NT, true>(D, xi, xblk);
        else sat_sweep_x<T, CPT, NT, false>(D, xi, xblk);
        return;
    } else {
    const bool m = D.v.mask != nullptr;
    if (D.v.dtype == XC_U8 && D.SQ && tile_bands(D) == 1 && D.v.w <= TP_MAX_W) {
        if (m) {
            if (D.Z) sat_sweep_u8<T, TB, CPT, NT, true, true>(D, xi, xblk);
            else sat_sweep_u8<T, TB, CPT, NT, true, false>(D, xi, xblk);
        } else {
            if (D.Z) sat_sweep_u8<T, TB, CPT, NT, false, true>(D, xi, xblk);
            else sat_sweep_u8<T, TB, CPT, NT, false, false>(D, xi, xblk);
        }
        return;
    }
#define SW_CALL(InT, MK) sat_sweep<T, TB, CPT, NT, InT, MK>(D, band, xs, xq, xz)
    XC_SAT_DISPATCH(SW_CALL)
#undef SW_CALL
    }
}

// The sweep launch over nitems (tile, band) items: list[item / gnb] is the
// tile and item % gnb the band.  Workgroup b takes items b, b + gridDim.x,
// ..., so a launch may hold fewer workgroups than items (a sweep beside the
// row transforms then leaves them wave slots on every CU).
template <typename T, typename TB, int CPT, int NT>
__global__ void __launch_bounds__(NT)
    k_tp_sat(const TileDev *__restrict__ tiles, const int *__restrict__ list, int gnb, int stage, int nitems)
{
    __shared__ double xs[2 * NT / 64], xq[2 * NT / 64];
    __shared__ int xz[2 * NT / 64], xi[8 * SATX_RB * NT / 64];
    extern __shared__ uint4 xblk_dyn[];  // SATX_BLK bytes when stage: u8 sweeps' staged pixel (and mask) rows
    uint4 *xblk = stage ? xblk_dyn : nullptr;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        if (it != (int)blockIdx.x) __syncthreads();  // the item before has finished with the LDS
        const TileDev D = tiles[list[it / gnb]];  // by value: the table stores cannot alias it
        sat_item<T, TB, CPT, NT>(D, it % gnb, xs, xq, xz, xi, xblk);
    }
}

// Columns per thread of the sweep for width w: two up to 2048 columns (a
// wavefront's row stores are then contiguous 1 KB runs), then wider.  The
// launch uses the power of two >= w / cpt threads (at least 64): few
// shapes to compile, and images of similar widths share a launch.
__host__ __device__ inline int sat_cpt(int w) { return w <= 2048 ? 2 : w <= 4096 ? 4 : 8; }

__host__ __device__ inline int sat_threads(int w)
{
    const int c = sat_cpt(w);
    const int need = (w + c - 1) / c;
    int n = 64;
    while (n < need && n < SAT_MAX_THREADS) n *= 2;
    return n;
}

// ------------------------------------------------------------------
// Rotated splat, gather form: each output pixel sums the bilinear weights
// of the grid points that land next to it, in the grid's raster order --
// the order in which the reference's scatter (DistributePixel, Maths.h:
// 236-276) adds them -- so the result is bit-identical to the scatter and
// independent of scheduling.
// ------------------------------------------------------------------

// Each thread makes two neighbouring pixels of a row, (px, py) and (px+1, py),
// walking the grid points of both pixels' squares once; each pixel still sums
// its own contributions in the grid's raster order.
__host__ __device__ inline long long rot_splat_units(int w, int h) { return (long long)((w + 1) / 2) * h; }

// Bilinear weight of a point whose mapped cell is (ix, iy) with fractions
// (alpha, beta) on pixel (q, py), DistributePixel's four products; false when
// the point does not reach the pixel.
__device__ __forceinline__ bool rot_weight(int ix, int iy, double alpha, double beta, int q, int py, double *w)
{
    if (ix == q) {
        if (iy == py) *w = (1.0 - alpha) * (1.0 - beta);
        else if (iy + 1 == py) *w = (1.0 - alpha) * beta;
        else return false;
    } else if (ix + 1 == q) {
        if (iy == py) *w = alpha * (1.0 - beta);
        else if (iy + 1 == py) *w = alpha * beta;
        else return false;
    } else {
        return false;
    }
    return true;
}

#pragma clang fp contract(off)
// The normalised values (raw - mean) / sd of a rotated copy's source, the
// splat's per-point quotient evaluated once per run for every copy and
// every output pixel that reads the point (absent points too: the splat
// skips them).
static __global__ void __launch_bounds__(256) k_rot_norm(const RotNorm *__restrict__ srcs, const int2 *__restrict__ work)
{
    const int2 wi = work[blockIdx.x];
    const RotNorm N = srcs[wi.x];
    const long long i = (long long)wi.y * 256 + threadIdx.x;
    if (i >= (long long)N.src.w * N.src.h) return;
    const int y = (int)(i / N.src.w), x = (int)(i - (long long)y * N.src.w);
    const double mean = N.src.msd[0], sd = N.src.msd[1];
    gst(N.dst, i, (load_raw(N.src, x, y) - mean) / sd);
}

template <typename T>
__global__ void __launch_bounds__(256) k_rot_splat(const RotDev *__restrict__ rots, const int2 *__restrict__ work)
{
    const int2 wi = work[blockIdx.x];
    const RotDev R = rots[wi.x];
    const long long unit = (long long)wi.y * 256 + threadIdx.x;
    if (unit >= rot_splat_units(R.w, R.h)) return;
    const int wp = (R.w + 1) / 2;
    const int py = (int)(unit / wp), px = 2 * (int)(unit - (long long)py * wp);
    const bool two = px + 1 < R.w;
    // grid points whose mapped position lies in [px-1, px+1+two) x [py-1, py+1):
    // they lie in the inverse image of that rectangle (for a rotation, its
    // transpose), so in the integer points of its bounding box, taken with a
    // slack of 1e-7 px for the rounding of the corners (~1e-13 px here)
    double gx0 = 1e300, gx1 = -1e300, gy0 = 1e300, gy1 = -1e300;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double cx = (double)(px - 1 + (c & 1) * (two ? 3 : 2) + R.L), cy = (double)(py - 1 + 2 * (c >> 1) + R.B);
        const double gx = cx * R.u0 + cy * R.u1, gy = cx * R.u3 + cy * R.u4;
        gx0 = fmin(gx0, gx);
        gx1 = fmax(gx1, gx);
        gy0 = fmin(gy0, gy);
        gy1 = fmax(gy1, gy);
    }
    const int ix0 = imax(0, (int)ceil(gx0 - 1e-7)), ix1 = imin(R.src.w - 1, (int)floor(gx1 + 1e-7));
    const int iy0 = imax(0, (int)ceil(gy0 - 1e-7)), iy1 = imin(R.src.h - 1, (int)floor(gy1 + 1e-7));
    double acc0 = 0.0, acc1 = 0.0;
    for (int gy = iy0; gy <= iy1; ++gy) {
        for (int gx = ix0; gx <= ix1; ++gx) {
            if (!present(R.src, gx, gy)) continue;
            const double xr = (double)gx * R.t0 + (double)gy * R.t1;
            const double yr = (double)gx * R.t3 + (double)gy * R.t4;
            const double fx = xr - R.L, fy = yr - R.B;
            const int ix = (int)floor(fx), iy = (int)floor(fy);
            const double alpha = fx - ix, beta = fy - iy;
            double w0, w1;
            const bool h0 = rot_weight(ix, iy, alpha, beta, px, py, &w0);
            const bool h1 = two && rot_weight(ix, iy, alpha, beta, px + 1, py, &w1);
            if (!h0 && !h1) continue;
            const double v = gld(R.nsrc, (long long)gy * R.src.w + gx);
            if (h0) acc0 += w0 * v;
            if (h1) acc1 += w1 * v;
        }
    }
    T *out = reinterpret_cast<T *>(R.out) + (long long)py * R.w + px;
    out[0] = (T)acc0;
    if (two) out[1] = (T)acc1;
}
#pragma clang fp contract(on)

// ------------------------------------------------------------------
// Row r2c: one Nx-point real transform per image row, via an Nx/2-point
// complex transform and the split post-twiddle.
// ------------------------------------------------------------------

#define C2R_ROW(nh) ((nh) + 2 > lds_len(nh) ? (nh) + 2 : lds_len(nh))

// Workgroup-index remap so that consecutive logical work items run on the
// same XCD: the dispatcher deals workgroups to the 8 XCDs round robin, and
// neighbouring items share cache lines (column tiles) or table rows.
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

template <int NH> struct RowCfg {
    static constexpr int TPF = Plan<NH>::TPF;
    static constexpr int G = TPF >= 256 ? 1 : 256 / TPF;  // rows per workgroup
    static constexpr int THREADS = TPF * G;
};

// Row r2c shape: RowCfg<NH>'s thread groups, each transforming RPT rows at
// once (FP32 rows of >= 2048 points: two, sharing barriers and twiddles).
// ROWS rows make one work item; a launch may cover them with RPT rows per
// thread in ROWS / RPT thread groups (the autotuner's choice, same results:
// each row's transform is the same arithmetic in either shape).
template <typename T, int NH> struct R2cCfg {
    static constexpr int RPT = (sizeof(T) == 4 && Plan<NH>::TPF >= 256) ? 2 : 1;
    static constexpr int ROWS = RowCfg<NH>::G * RPT;  // rows per workgroup (per work item)
    // the other rows-per-thread shape the tuner may pick (0: none): FP32, >= 512-point rows
    static constexpr int ALT = (sizeof(T) == 4 && NH >= 256 && ROWS >= 2 && Plan<NH>::TPF * ROWS <= 1024)
                                   ? (RPT == 2 ? 1 : 2)
                                   : 0;
    static constexpr int threads(int rpt) { return Plan<NH>::TPF * (ROWS / rpt); }
};

template <typename T, int NH, int RPT = R2cCfg<T, NH>::RPT>
__global__ void __launch_bounds__((R2cCfg<T, NH>::threads(RPT)))
    k_row_r2c(const PairDev *__restrict__ pairs, const int4 *__restrict__ work, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    const int4 wi = work[blockIdx.x];  // {pair, image, y0, -}
    const PairDev &P = pairs[wi.x];
    const int img = wi.y, y0 = wi.z + g * RPT;
    const ImgView &V = img ? P.B : P.A;
    cpx<T> *dst = reinterpret_cast<cpx<T> *>(img ? P.specB : P.specA);
    const int H = NH + 1;
    const double mean = V.msd[0], inv = V.msd[2];

    cpx<T> v[RPT][E];
    LdsView<T> L[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) L[q] = LdsView<T>{smem + (g * RPT + q) * lds_len(NH), 1};
#define XC_LOAD_ROW(ME)                                                                                    \
    _Pragma("unroll") for (int q = 0; q < RPT; ++q) {                                                       \
        const int y = y0 + q;                                                                               \
        const bool valid = y < V.h;                                                                         \
        switch (V.dtype) { /* uniform per workgroup */                                                      \
        case XC_U8: load_row_pairs<T, uint8_t, ME, TPF>(v[q], V, mean, inv, t, y, valid); break;            \
        case XC_U16: load_row_pairs<T, uint16_t, ME, TPF>(v[q], V, mean, inv, t, y, valid); break;          \
        case XC_F32: load_row_pairs<T, float, ME, TPF>(v[q], V, mean, inv, t, y, valid); break;             \
        default: load_row_pairs<T, double, ME, TPF>(v[q], V, mean, inv, t, y, valid); break;                \
        }                                                                                                   \
    }
    if (E == 8 && V.w <= NH) {
        // x >= NH (registers m >= E/2) is padding for every row of this workgroup
        XC_LOAD_ROW(E / 2)
        fft_rows<-1, NH, T, RPT, true>(v, t, L, tw);
    } else {
        XC_LOAD_ROW(E)
        fft_rows<-1, NH, T, RPT, false>(v, t, L, tw);
    }
#undef XC_LOAD_ROW

    // the split's twiddles, loaded before the exchange so that they arrive during it
    cpx<T> ws[E];
#pragma unroll
    for (int m = 0; m < E; ++m) ws[m] = twiddle<-1>(tw, t + TPF * m, 2 * NH);
    lds_sync();
#pragma unroll
    for (int q = 0; q < RPT; ++q)
#pragma unroll
        for (int m = 0; m < E; ++m) L[q][t + TPF * m] = v[q][m];
    lds_sync();
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int y = y0 + q;
        if (y < V.h) {
            cpx<T> *row = dst + (long long)y * H;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int k = t + TPF * m;
                const cpx<T> zk = v[q][m];
                const cpx<T> zn = cconj(L[q][(NH - k) & (NH - 1)]);
                const cpx<T> a = cscale(zk + zn, (T)0.5);
                const cpx<T> b = rot90<-1>(cscale(zk - zn, (T)0.5));
                gst(row, k, a + cmul(ws[m], b));
            }
            if (t == 0) gst(row, NH, cpx<T>{v[q][0].x - v[q][0].y, (T)0});
        }
    }
}

// Row r2c of 8-bit rows (FP32 transforms), NP work items per workgroup in
// turn: the next item's pixels are loaded while this item's rows are
// transformed, and the split's twiddles are loaded once for all of them.  Each
// row is the same arithmetic as k_row_r2c's, so the spectra are bitwise the
// same: unmasked rows at even addresses (GEN false: two pixels a 16-bit load,
// normalised with an FP32 fma) or any rows (GEN true: pixel and mask bytes,
// normalised in FP64 and rounded, absent pixels 0).  The host takes this
// kernel only for rows that fit the upper-half padding (w <= NH).
template <int NH, int RPT, int NP, bool GEN>
__global__ void __launch_bounds__((R2cCfg<float, NH>::threads(RPT)))
    k_row_r2c_u8(const R2cU8 *__restrict__ items, int nitems, const cpx<float> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF, H = NH + 1;
    static_assert(E == 8, "8-bit row r2c: 8 elements a thread");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<float> *smem = reinterpret_cast<cpx<float> *>(smem_raw);
    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    LdsView<float> L[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) L[q] = LdsView<float>{smem + (g * RPT + q) * lds_len(NH), 1};
    cpx<float> ws[E];
#pragma unroll
    for (int m = 0; m < E; ++m) ws[m] = twiddle<-1>(tw, t + TPF * m, 2 * NH);
    const int i0 = blockIdx.x * NP;
    // pixel pairs x = 2p, 2p + 1 (p = t + TPF m, m < E / 2) of this thread's rows: the two pixels in bits
    // 0-15; GEN: bits 16 and 17 flag them present (inside the row and the mask), bit 18 marks an unmasked
    // row at an even address (k_row_r2c normalises those rows with the FP32 fma, the others in FP64)
    unsigned px[RPT][E / 2];
    auto fetch = [&](const R2cU8 &it) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = g * RPT + q;
            const uint8_t *row = it.src + (long long)r * it.stride;
            const uint8_t *mrow = it.mask ? it.mask + (long long)r * it.mstride : nullptr;
            const uint16_t *r2 = reinterpret_cast<const uint16_t *>(row);
#pragma unroll
            for (int m = 0; m < E / 2; ++m) {
                const int p = t + TPF * m, x = 2 * p;
                unsigned u = 0;
                if (r < it.nrows) {
                    if constexpr (GEN) {
                        const bool even = ((reinterpret_cast<uintptr_t>(row) | reinterpret_cast<uintptr_t>(mrow)) & 1) == 0;
                        if (even && x + 1 < it.w) {
                            // both pixels (and both mask bytes) in one 16-bit load each; a pixel whose mask
                            // byte is 0 is dropped as below
                            const unsigned pp = gld(r2, p);
                            const unsigned mm = mrow ? gld(reinterpret_cast<const uint16_t *>(mrow), p) : 0x0101u;
                            if (mm & 0x00ffu) u |= (pp & 0x00ffu) | 0x10000u;
                            if (mm & 0xff00u) u |= (pp & 0xff00u) | 0x20000u;
                        } else {
                            if (x < it.w && (!mrow || gld(mrow, x))) u |= gld(row, x) | 0x10000u;
                            if (x + 1 < it.w && (!mrow || gld(mrow, x + 1))) u |= ((unsigned)gld(row, x + 1) << 8) | 0x20000u;
                        }
                        if (!mrow && (reinterpret_cast<uintptr_t>(row) & 1) == 0) u |= 0x40000u;
                    } else {
                        if (x + 1 < it.w) u = gld(r2, p);
                        else if (x < it.w) u = gld(row, x);
                    }
                }
                px[q][m] = u;
            }
        }
    };
    fetch(items[i0]);
    for (int i = 0; i < NP; ++i) {
        const int ii = i0 + i;
        if (ii >= nitems) break;  // uniform over the workgroup
        const R2cU8 it = items[ii];
        const double mean = gld(it.msd, 0), inv = gld(it.msd, 2);
        const float sc = (float)inv, of = (float)(-mean * inv);
        cpx<float> v[RPT][E];
#pragma unroll
        for (int q = 0; q < RPT; ++q)
#pragma unroll
            for (int m = 0; m < E / 2; ++m) {
                const unsigned u = px[q][m];
                if constexpr (GEN) {
                    const float a = (float)(u & 255u), b = (float)((u >> 8) & 255u);
                    const bool f = (u & 0x40000u) != 0;
                    v[q][m].x = !(u & 0x10000u) ? 0.0f : f ? fmaf(a, sc, of) : (float)(((double)a - mean) * inv);
                    v[q][m].y = !(u & 0x20000u) ? 0.0f : f ? fmaf(b, sc, of) : (float)(((double)b - mean) * inv);
                } else {
                    const int x = 2 * (t + TPF * m);
                    const bool pr = g * RPT + q < it.nrows;
                    v[q][m].x = (pr && x < it.w) ? fmaf((float)(u & 255u), sc, of) : 0.0f;
                    v[q][m].y = (pr && x + 1 < it.w) ? fmaf((float)(u >> 8), sc, of) : 0.0f;
                }
            }
        if (i + 1 < NP && ii + 1 < nitems) fetch(items[ii + 1]);
        fft_rows<-1, NH, float, RPT, true>(v, t, L, tw);
        lds_sync();
#pragma unroll
        for (int q = 0; q < RPT; ++q)
#pragma unroll
            for (int m = 0; m < E; ++m) L[q][t + TPF * m] = v[q][m];
        lds_sync();
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = g * RPT + q;
            if (r < it.nrows) {
                cpx<float> *row = reinterpret_cast<cpx<float> *>(it.dst) + (long long)r * H;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const int k = t + TPF * m;
                    const cpx<float> zk = v[q][m];
                    const cpx<float> zn = cconj(L[q][(NH - k) & (NH - 1)]);
                    const cpx<float> a = cscale(zk + zn, 0.5f);
                    const cpx<float> b = rot90<-1>(cscale(zk - zn, 0.5f));
                    gst(row, k, a + cmul(ws[m], b));
                }
                if (t == 0) gst(row, NH, cpx<float>{v[q][0].x - v[q][0].y, 0.0f});
            }
        }
        // (the next item's first exchange waits at its own barrier before overwriting these rows)
    }
}

// ------------------------------------------------------------------
// Column pass, fused: forward Ny-point FFT of the A and B spectrum
// columns, cross power F_B conj(F_A) (Correlation.cpp:972-986; whitened
// for S mode, :2311-2317), inverse Ny-point FFT, write the hR lag rows.
// ------------------------------------------------------------------

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int pow2_floor(int n) { return n >= 2 ? 2 * pow2_floor(n / 2) : 1; }

// Column work items cover COL_STEP spectrum columns; a workgroup takes C of
// them (C divides COL_STEP), so one work list serves every C.
constexpr int COL_STEP = 16;

template <typename T, int NY> struct ColCfg {
    static constexpr int TPF = Plan<NY>::TPF;
    static constexpr int CW = sizeof(T) == 8 ? 4 : 8;  // columns per workgroup by default (64 B rows)
    static constexpr int CT = 1024 / TPF > 0 ? 1024 / TPF : 1;                       // <= 1024 threads
    // FP32: the A and B forward transforms run interleaved (sharing barriers and
    // twiddles), each column then holds two LDS regions
    static constexpr int AB = sizeof(T) == 4 ? 2 : 1;
    static constexpr int CL = 65536 / (AB * lds_len(NY) * (int)sizeof(cpx<T>)) > 0
                                  ? 65536 / (AB * lds_len(NY) * (int)sizeof(cpx<T>))
                                  : 1;  // <= 64 KB LDS
    static constexpr int C = pow2_floor(cmin(CW, cmin(CT, CL)));  // a power of two, so it divides COL_STEP
    static constexpr int THREADS = TPF * C;
    static constexpr long long lds_bytes(int CC) { return (long long)CC * AB * lds_len(NY) * (long long)sizeof(cpx<T>); }
    // a workgroup of CC columns fits: <= 1024 threads and <= 160 KB of LDS
    static constexpr bool fits(int CC)
    {
        return CC >= 1 && COL_STEP % CC == 0 && TPF * CC <= 1024 && lds_bytes(CC) <= 163840;
    }
};

// Spectrum element m of this thread's column: row t + TPF m, from the thread's
// element pointer e (row t) and the row step TPF H; 0 when !ok (padding row or
// column, whose load is skipped).
template <typename T>
__device__ __forceinline__ cpx<T> col_ld(const cpx<T> *e, bool ok, int m, long long step)
{
    return ok ? gld(e, m * step) : cpx<T>{(T)0, (T)0};
}

template <typename T, int NY, bool WHITEN, int C>
__global__ void __launch_bounds__(Plan<NY>::TPF * C)
    k_col_fused(const PairDev *__restrict__ pairs, const int2 *__restrict__ work, const cpx<T> *__restrict__ tw,
                int xcd)
{
    using PL = Plan<NY>;
    constexpr int E = PL::E, TPF = PL::TPF, SUB = COL_STEP / C;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    const int bid = xcd ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const int2 wi = work[bid / SUB];  // {pair, k0}
    const PairDev &P = pairs[wi.x];
    const int H = P.Nx / 2 + 1;
    const int k = wi.y + (bid % SUB) * C + c;
    const bool kv = k < H;
    const cpx<T> *SA = reinterpret_cast<const cpx<T> *>(P.specA);
    const cpx<T> *SB = reinterpret_cast<const cpx<T> *>(P.specB);
    cpx<T> *LG = reinterpret_cast<cpx<T> *>(P.lag);
    // row t of column k in each spectrum (a padding column reads nothing), then TPF rows per register
    const long long RS = (long long)TPF * H;
    const cpx<T> *EA = SA + (long long)t * H + k, *EB = SB + (long long)t * H + k;
    LdsView<T> L{smem + c, C};
    TwReg<NY, T> TW;
    tw_load<NY, T>(TW, t, tw);

    cpx<T> a[E], b[E];
    if constexpr (ColCfg<T, NY>::AB == 2) {
        // both forward transforms at once, B's in the second LDS region
        cpx<T> ab[2][E];
        const LdsView<T> L2[2] = {L, LdsView<T>{smem + C * lds_len(NY) + c, C}};
        if (E == 8 && P.A.h <= NY / 2 && P.B.h <= NY / 2) {
            // rows y >= NY/2 (registers m >= E/2) are padding of both images
#pragma unroll
            for (int m = 0; m < E / 2; ++m) {
                const int y = t + TPF * m;
                ab[0][m] = col_ld(EA, kv && y < P.A.h, m, RS);
                ab[1][m] = col_ld(EB, kv && y < P.B.h, m, RS);
            }
            fft_rows_tw<-1, NY, T, 2, true>(ab, t, L2, TW);
        } else {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const int y = t + TPF * m;
                ab[0][m] = col_ld(EA, kv && y < P.A.h, m, RS);
                ab[1][m] = col_ld(EB, kv && y < P.B.h, m, RS);
            }
            fft_rows_tw<-1, NY, T, 2, false>(ab, t, L2, TW);
        }
#pragma unroll
        for (int m = 0; m < E; ++m) {
            a[m] = ab[0][m];
            b[m] = ab[1][m];
        }
    } else {
    // rows y >= NY/2 (registers m >= E/2) are padding when the image is at most NY/2 high
    if (E == 8 && P.A.h <= NY / 2) {
#pragma unroll
        for (int m = 0; m < E / 2; ++m) {
            const int y = t + TPF * m;
            a[m] = col_ld(EA, kv && y < P.A.h, m, RS);
        }
        fft_upper_zero<-1, NY>(a, t, L, TW);
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = t + TPF * m;
            a[m] = col_ld(EA, kv && y < P.A.h, m, RS);
        }
        fft<-1, NY>(a, t, L, TW);
    }
    // (B's transform waits at its first exchange's barrier before it overwrites A's LDS rows)
    if (E == 8 && P.B.h <= NY / 2) {
#pragma unroll
        for (int m = 0; m < E / 2; ++m) {
            const int y = t + TPF * m;
            b[m] = col_ld(EB, kv && y < P.B.h, m, RS);
        }
        fft_upper_zero<-1, NY>(b, t, L, TW);
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int y = t + TPF * m;
            b[m] = col_ld(EB, kv && y < P.B.h, m, RS);
        }
        fft<-1, NY>(b, t, L, TW);
    }
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        cpx<T> p;
        p.x = b[m].x * a[m].x + b[m].y * a[m].y;
        p.y = b[m].y * a[m].x - b[m].x * a[m].y;
        if (P.filter) {
            // CorrPatchToImage's low-pass (Correlation.cpp:1255-1283): with i the index
            // into the [N][N/2+1] half spectrum, "x" = i / (N/2+1) is the row index.
            int fx = t + TPF * m;
            if (fx >= P.Ny / 2) fx -= P.Ny;
            const double rad2 = ((double)fx * fx + (double)k * k) / 10000.0;
            const T filt = rad2 > 10.0 ? (T)0 : (T)exp(-rad2);
            p.x = p.x * filt;
            p.y = p.y * filt;
        }
        if constexpr (WHITEN) {
            if constexpr (sizeof(T) == 8) {
                // the reference's arithmetic (Correlation.cpp:2311-2317): |P|, then P / sqrt(|P|)
                T mag = hypot(p.x, p.y);
                if (mag > (T)1e-10) {
                    T s = sqrt(mag);
                    p.x /= s;
                    p.y /= s;
                }
            } else {
                // FP32: P * |P|^(-1/2) from |P|^2 (|P| > 1e-10 is |P|^2 > 1e-20): one square root and
                // one reciprocal square root; |P|^2 beyond the FP32 range (un-normalised pixels) takes hypot
                const float m2 = p.x * p.x + p.y * p.y;
                if (m2 > 1e-20f) {
                    const float s = m2 < 1e38f ? __frsqrt_rn(__fsqrt_rn(m2)) : 1.0f / sqrtf(hypotf(p.x, p.y));
                    p.x *= s;
                    p.y *= s;
                }
            }
        }
        a[m] = p;
    }
    // (the inverse waits at its first exchange's barrier before it overwrites the forward's LDS rows)
    fft<+1, NY>(a, t, L, TW);
    if (kv) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int iy = t + TPF * m;
            const int yl = lag_row_of_raster(P, iy);
            if (yl >= 0 && row_live(P, yl)) gst(LG, (long long)yl * H + k, a[m]);
        }
    }
}

// ------------------------------------------------------------------
// Row c2r + evaluation: per stored lag row, Nx-point c2r (Nx/2-point
// complex inverse after the split pre-twiddle), mode value per lag, and
// the row's best candidate with its two in-row neighbours.
// ------------------------------------------------------------------

// Value of raster column ix of lag row y (signed; DEAD_ROW for a P2I dead
// row) from its raw c2r value, with the mode's scan semantics.
//   *cand : the lag competes for the peak
//   *win  : the lag belongs to the centred wR x hR image (R/S modes)
//   *odd  : GetCorr's exit(44) condition
template <typename FT>
__device__ __forceinline__ double lag_value(const PairDev &P, const RowBox &rb, int y, int ix, double raw,
                                            bool *cand, bool *win, bool *odd)
{
    *cand = false;
    *win = false;
    *odd = false;
    bool dead;
    const int x = lag_x_of_raster(P, ix, &dead);
    if (P.mode == PM_P2I) {
        // CorrPatchToImage scan (Correlation.cpp:1299-1327): raw surface, zones, disc
        if (!dead && y != DEAD_ROW) {
            const long long dxd = x - P.Ox, dyd = y - P.Oy;
            *cand = dxd * dxd + dyd * dyd <= (long long)P.radius * P.radius;
        }
        return raw;
    }
    if (dead) return 0.0;
    if (P.mode == PM_NCC) {
        // CorrPatches scan: disc about (Ox, Oy), then the rules (Correlation.cpp:1030-1061)
        const long long dxd = x - P.Ox, dyd = y - P.Oy;
        if (dxd * dxd + dyd * dyd > (long long)P.radius * P.radius) return 0.0;
        bool legal;
        const double v = eval_lag<FT>(P, rb, x, raw, &legal, odd);
        *cand = legal;
        return v;
    }
    bool legal, o;
    const double v = eval_lag<FT>(P, rb, x, raw, &legal, &o);
    *win = true;
    bool c = legal;
    if (P.Rx > 0 && P.Ry > 0) {
        // MaskA (Correlation.cpp:2011-2054) in centred coordinates
        const int ox = P.Ox - P.offx + P.nnegx, oy = P.Oy - P.offy + P.nnegy;
        const int cx = x + P.nnegx, cy = y + P.nnegy;
        if (cy < oy - P.Ry || cy > oy + P.Ry || cx < ox - P.Rx || cx > ox + P.Rx) c = false;
        else {
            const double rx2 = (double)P.Rx * P.Rx, asq = rx2 / ((double)P.Ry * P.Ry);
            const int xx = cx - ox, yy = cy - oy;
            if (xx * xx + asq * yy * yy > rx2) c = false;
        }
    }
    if (P.mincor > 0.0 && !(v > P.mincor)) c = false;
    *cand = c;
    return v;
}

// ---- CorrPatches rows (NCC mode): the legal lags of one row form an x
// interval, so a row is evaluated over that interval only.
//
// With olw(x) = min(w1, w2, w2 - x, w1 + x), every scalar LegalRgn term is
// "olw >= m" for a per-row m (min_w; olw > min_area / olh; the or-term when
// olh <= or_h_gt), and the disc is |x - Ox| <= s with s = floor(sqrt(r^2 -
// dy^2)).  exact_w, a tabulated rule and the count predicate stay per lag.
struct NccRow {
    int olh;
    int x0, x1;    // legal-interval bounds (empty when x0 > x1)
    double invN;   // 1 / (Nx Ny), a power of two: raw * invN == raw / Nxy exactly
    const void *sqAt, *sqAb, *sqBt, *sqBb;  // (S, S^2) table rows: band top, row below the band (or null)
    const int32_t *zAt, *zAb, *zBt, *zBb;
    NmD nA, nB;    // the images' normalisation (exact tables only)
};

__host__ __device__ __forceinline__ long long isqrt_floor(long long v)
{
    long long s = (long long)sqrt((double)v);
    while (s > 0 && s * s > v) --s;
    while ((s + 1) * (s + 1) <= v) ++s;
    return s;
}

template <typename FT>
__host__ __device__ __forceinline__ NccRow ncc_row(const PairDev &P, int y)
{
    NccRow q;
    const RowBox rb = row_box(P, y);
    const RuleDev &g = P.rule;
    q.olh = rb.olh;
    q.invN = 1.0 / ((double)P.Nx * (double)P.Ny);
    long long lo = -P.nnegx, hi = P.nposx - 1;
    const long long dy = (long long)y - P.Oy, r2 = (long long)P.radius * P.radius - dy * dy;
    bool ok = r2 >= 0;
    if (ok) {
        const long long s = isqrt_floor(r2);
        lo = lo > P.Ox - s ? lo : P.Ox - s;
        hi = hi < P.Ox + s ? hi : P.Ox + s;
    }
    long long m = 1;
    if (g.rgn_enabled) {
        ok = ok && rb.olh >= g.min_h && (g.exact_h <= 0 || rb.olh == g.exact_h);
        if (g.min_w > m) m = g.min_w;
        if (g.min_area >= 0 && g.min_area / rb.olh + 1 > m) m = g.min_area / rb.olh + 1;
        if ((g.or_w_ge > 0 || g.or_h_gt > 0) && !(rb.olh > g.or_h_gt) && g.or_w_ge > m) m = g.or_w_ge;
    }
    if (g.cnt_enabled) {
        // LegalCnt needs c1 > k1 and c2 > k2, and a box's present-pixel count is at most its area
        // olw olh: a lag with olw olh <= max(k1, k2) is no candidate, its R value 0
        const long long K = g.k1 > g.k2 ? g.k1 : g.k2;
        if (K >= 0 && K / rb.olh + 1 > m) m = K / rb.olh + 1;
    }
    // olw(x) >= m  <=>  min(w1, w2) >= m  &&  m - w1 <= x <= w2 - m
    ok = ok && m <= (P.w1 < P.w2 ? P.w1 : P.w2);
    if (m - P.w1 > lo) lo = m - P.w1;
    if (P.w2 - m < hi) hi = P.w2 - m;
    ok = ok && lo <= hi;
    q.x0 = ok ? (int)lo : 1;
    q.x1 = ok ? (int)hi : 0;
    const long long at = (long long)rb.a_t * P.w1, ab = (long long)(rb.a_b - 1) * P.w1;
    const long long bt = (long long)rb.b_t * P.w2, bb = (long long)(rb.b_b - 1) * P.w2;
    const tab_t<FT> *SA = reinterpret_cast<const tab_t<FT> *>(P.SQA), *SB = reinterpret_cast<const tab_t<FT> *>(P.SQB);
    q.sqAt = SA + at;
    q.sqBt = SB + bt;
    q.sqAb = rb.a_b > 0 ? SA + ab : nullptr;
    q.sqBb = rb.b_b > 0 ? SB + bb : nullptr;
    q.zAt = P.ZA ? P.ZA + at : nullptr;
    q.zBt = P.ZB ? P.ZB + bt : nullptr;
    q.zAb = (P.ZA && rb.a_b > 0) ? P.ZA + ab : nullptr;
    q.zBb = (P.ZB && rb.b_b > 0) ? P.ZB + bb : nullptr;
    q.nA = NmD{0.0, 0.0, 0.0, 0ULL, 0u, 0.0f, 0.0f};
    q.nB = q.nA;
#ifdef __HIP_DEVICE_COMPILE__
    if constexpr (is_tabx<FT>()) {
        q.nA = nm_of_x(P.A, P.SQA, P.w1, P.h1);
        q.nB = nm_of_x(P.B, P.SQB, P.w2, P.h2);
    }
#endif
    return q;
}

// IntegralTable (Correlation.cpp:160-201) with the box's two table rows
// given: t[R,T] - t[L-1,T] - t[R,B-1] + t[L-1,B-1], in that order.
template <typename TT>
__device__ __forceinline__ TT band_sum(const TT *__restrict__ top, const TT *__restrict__ bot, int L, int R)
{
    TT v = top[R];
    if (L > 0) v -= top[L - 1];
    if (bot) {
        v -= bot[R];
        if (L > 0) v += bot[L - 1];
    }
    return v;
}

// band_sum on an (S, S^2) table, in the same order.
template <typename TB>
__device__ __forceinline__ cor_t<TB> band_c(const void *__restrict__ top, const void *__restrict__ bot, int L, int R)
{
    cor_t<TB> v = tab_c<TB>(top, R);
    if (L > 0) v = c_sub(v, tab_c<TB>(top, L - 1));
    if (bot) {
        v = c_sub(v, tab_c<TB>(bot, R));
        if (L > 0) v = c_add(v, tab_c<TB>(bot, L - 1));
    }
    return v;
}

#pragma clang fp contract(off)
// GetCorr's value from the overlap size, the raw correlation sum and the two
// images' band sums (S, S^2): FP64, except that the FP32 (batched) form takes
// the final num / sqrt(d) in FP32.
template <typename T>
__device__ __forceinline__ double ncc_r(int olw, int olh, double invN, double raw, double s1, double q1, double s2,
                                        double q2)
{
    const double n = (double)(olw * olh);
    const double num = n * raw * invN - s1 * s2;
    const double d1 = n * q1 - s1 * s1;
    const double d2 = n * q2 - s2 * s2;
    const double d = d1 * d2;
    if constexpr (sizeof(T) == 8) return d < n * n * 1.0E-9 ? 0.0 : num / sqrt(d);
    else return d < n * n * 1.0E-9 ? 0.0 : (double)((float)num * rsqrtf((float)d));
}

// ncc_r from two boxes' table sums in the table's own form (exact tables:
// tabx_r32 in the FP32 form, tabx_num_d and ncc_r's final step in FP64).
template <typename T, typename TB>
__device__ __forceinline__ double ncc_rc(int olw, int olh, double invN, double raw, cor_t<TB> a, cor_t<TB> b,
                                         const NmD &nA, const NmD &nB)
{
    if constexpr (!is_tabx<TB>()) {
        return ncc_r<T>(olw, olh, invN, raw, a.x, a.y, b.x, b.y);
    } else if constexpr (sizeof(T) == 4) {
        return tabx_r32((unsigned)(olw * olh), raw, invN, a, b, nA, nB);
    } else {
        const double n = (double)(olw * olh);
        double num, d;
        tabx_num_d((unsigned)(olw * olh), n * raw * invN, a, b, nA, nB, &num, &d);
        if constexpr (sizeof(T) == 8) return d < n * n * 1.0E-9 ? 0.0 : num / sqrt(d);
        else return d < n * n * 1.0E-9 ? 0.0 : (double)((float)num * rsqrtf((float)d));
    }
}
#pragma clang fp contract(on)

// NCC value of lag x of a row (CLinCorr::CheckDensity/GetCorr, Correlation.cpp:697-754);
// the lag is inside the row's legal interval.  *cand: passes exact_w, the table and
// LegalCnt; *odd: |r| > 1.0001.
#pragma clang fp contract(off)
// FP64 (plain form): GetCorr's division and square root in FP64.  FP32
// (batched form): the sums and both subtractions stay FP64, the final
// num / sqrt(d) is taken in FP32 (the batched form's precision).
template <typename T, typename TB>
__device__ __forceinline__ double ncc_lag(const PairDev &P, const NccRow &q, int x, double raw, bool *cand, bool *odd)
{
    const int BL = x > 0 ? x : 0;
    const int BR = (P.w2 - 1 < x + P.w1 - 1) ? P.w2 - 1 : x + P.w1 - 1;
    const int AL = BL - x, AR = BR - x;
    const int olw = BR - BL + 1;
    const RuleDev &g = P.rule;
    const cor_t<TB> sqa = band_c<TB>(q.sqAt, q.sqAb, AL, AR);
    const cor_t<TB> sqb = band_c<TB>(q.sqBt, q.sqBb, BL, BR);
    bool c = true;
    if (g.exact_w > 0 && olw != g.exact_w) c = false;
    if (g.table && (olw > g.table_w || q.olh > g.table_h || !gld(g.table, (long long)(q.olh - 1) * g.table_w + (olw - 1))))
        c = false;
    if (g.cnt_enabled) {
        const int c1 = band_sum(q.zAt, q.zAb, AL, AR);
        const int c2 = band_sum(q.zBt, q.zBb, BL, BR);
        if (!((long long)c1 > g.k1 && (long long)c2 > g.k2)) c = false;
    }
    *cand = c;
    const double r = ncc_rc<T, TB>(olw, q.olh, q.invN, raw, sqa, sqb, q.nA, q.nB);
    *odd = c && fabs(r) > 1.0001;
    return c ? r : 0.0;
}

// The NCC scan of one lag row when no per-lag rule but the count rule
// applies (no exact width, no table): ncc_lag's arithmetic with the row's
// table rows held in registers, the edge cases of band_sum2 resolved per
// segment.  AB / BB: the band has a row below it in A / B's tables.  CNT:
// LegalCnt's threshold test on the count tables (a lag that fails it is no
// candidate).
// Table element i read as a global-memory load (the table rows arrive as
// generic pointers), widened to FP64.
template <typename FT> struct TabV;
template <> struct TabV<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct TabV<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename FT>
__device__ __forceinline__ double2 ld_sq(const void *p, int i)
{
    typedef typename TabV<FT>::type V;
    typedef __attribute__((address_space(1))) const V gv;
    const V v = ((gv *)p)[i];
    return make_double2((double)v.x, (double)v.y);
}

// Table corner i of a row, in the table's own form.
template <typename TB>
__device__ __forceinline__ cor_t<TB> ld_c(const void *p, int i)
{
    if constexpr (is_tabx<TB>()) {
        typedef __attribute__((address_space(1))) const unsigned long long gu;
        return ((gu *)p)[i];
    } else {
        return ld_sq<TB>(p, i);
    }
}

__device__ __forceinline__ int ld_z(const int32_t *p, int i)
{
    typedef __attribute__((address_space(1))) const int gi;
    return ((gi *)(const void *)p)[i];
}

// One x range of a row's legal interval.  NEG: x < 0 (A's band starts at column -x,
// B's at 0), else x >= 0 (A's at 0, B's at x).  BC: B's right edge is the constant
// w2 - 1 over the range (x >= w2 - w1), else A's is the constant w1 - 1; the
// constant corner's table entries are loaded once.  The sums are formed in the
// same order either way.
template <typename T, typename TB, int TPF, bool AB, bool BB, bool CNT, bool NEG, bool BC>
__device__ __forceinline__ void ncc_scan_range(const PairDev &P, const NccRow &q, const T *R, int t, int NX, int xlo,
                                               int xhi, double &bestv, long long &bestk, int &flags)
{
    typedef cor_t<TB> C;
    const void *At = q.sqAt, *Ab = q.sqAb, *Bt = q.sqBt, *Bb = q.sqBb;
    const int w1 = P.w1, w2 = P.w2, olh = q.olh;
    const double invN = q.invN;
    const long long k1 = P.rule.k1, k2 = P.rule.k2;
    const int KC = BC ? w2 - 1 : w1 - 1;  // the constant right edge
    const C ct = ld_c<TB>(BC ? Bt : At, KC);
    const C cb = (BC ? BB : AB) ? ld_c<TB>(BC ? Bb : Ab, KC) : c_zero<C>();
    for (int x = xlo + t; x <= xhi; x += TPF) {
        const int BR = BC ? w2 - 1 : x + w1 - 1, AR = BR - x;
        bool cand = true;
        C a, b;
        if constexpr (!NEG) {
            // x >= 0: A's band starts at column 0, B's at x; x = 0 has no left term:
            // subtract / add 0 instead (exact), so that every load is unconditional
            const int BL1 = x > 0 ? x - 1 : 0;
            const bool l = x > 0;
            if constexpr (CNT) {
                int c1 = ld_z(q.zAt, AR), c2 = ld_z(q.zBt, BR);
                const int bl = ld_z(q.zBt, BL1);
                if constexpr (AB) c1 -= ld_z(q.zAb, AR);
                c2 -= l ? bl : 0;
                if constexpr (BB) {
                    const int u = ld_z(q.zBb, BR), z = ld_z(q.zBb, BL1);
                    c2 -= u;
                    c2 += l ? z : 0;
                }
                cand = (long long)c1 > k1 && (long long)c2 > k2;
            }
            a = BC ? ld_c<TB>(At, AR) : ct;
            if constexpr (AB) a = c_sub(a, BC ? ld_c<TB>(Ab, AR) : cb);
            b = BC ? ct : ld_c<TB>(Bt, BR);
            const C bl = ld_c<TB>(Bt, BL1);
            b = c_sub(b, l ? bl : c_zero<C>());
            if constexpr (BB) {
                const C u = BC ? cb : ld_c<TB>(Bb, BR), z = ld_c<TB>(Bb, BL1);
                b = c_sub(b, u);
                b = c_add(b, l ? z : c_zero<C>());
            }
        } else {
            // x < 0: A's band starts at column -x, B's at 0
            const int AL1 = -x - 1;
            if constexpr (CNT) {
                int c1 = ld_z(q.zAt, AR) - ld_z(q.zAt, AL1), c2 = ld_z(q.zBt, BR);
                if constexpr (AB) {
                    c1 -= ld_z(q.zAb, AR);
                    c1 += ld_z(q.zAb, AL1);
                }
                if constexpr (BB) c2 -= ld_z(q.zBb, BR);
                cand = (long long)c1 > k1 && (long long)c2 > k2;
            }
            a = BC ? ld_c<TB>(At, AR) : ct;
            a = c_sub(a, ld_c<TB>(At, AL1));
            if constexpr (AB) {
                const C u = BC ? ld_c<TB>(Ab, AR) : cb, z = ld_c<TB>(Ab, AL1);
                a = c_sub(a, u);
                a = c_add(a, z);
            }
            b = BC ? ct : ld_c<TB>(Bt, BR);
            if constexpr (BB) b = c_sub(b, BC ? cb : ld_c<TB>(Bb, BR));
        }
        const int ix = NEG ? NX + x : x;
        const double r = ncc_rc<T, TB>(NEG ? BR + 1 : BR - x + 1, olh, invN, (double)R[ix], a, b, q.nA, q.nB);
        if (cand) {
            flags |= fabs(r) > 1.0001 ? 3 : 1;
            if (r > bestv) {
                bestv = r;
                bestk = ix;
            }
        }
    }
}

// A row's legal interval in raster order: x >= 0 first, then x < 0; each thread's
// keys ascend, so a strict > keeps the first of equal values.  Each sign splits at
// x = w2 - w1, where the constant right edge passes from A to B.
template <typename T, typename TB, int TPF, bool AB, bool BB, bool CNT>
__device__ __forceinline__ void ncc_scan_rows(const PairDev &P, const NccRow &q, const T *R, int t, int NX, double &bestv,
                                              long long &bestk, int &flags)
{
    const int sw = P.w2 - P.w1;
    const int p0 = q.x0 > 0 ? q.x0 : 0, p1 = q.x1;        // x >= 0
    const int n0 = q.x0, n1 = q.x1 < -1 ? q.x1 : -1;     // x < 0
    ncc_scan_range<T, TB, TPF, AB, BB, CNT, false, false>(P, q, R, t, NX, p0, p1 < sw - 1 ? p1 : sw - 1, bestv, bestk,
                                                      flags);
    ncc_scan_range<T, TB, TPF, AB, BB, CNT, false, true>(P, q, R, t, NX, p0 > sw ? p0 : sw, p1, bestv, bestk, flags);
    ncc_scan_range<T, TB, TPF, AB, BB, CNT, true, false>(P, q, R, t, NX, n0, n1 < sw ? n1 : sw, bestv, bestk, flags);
    ncc_scan_range<T, TB, TPF, AB, BB, CNT, true, true>(P, q, R, t, NX, n0 > sw + 1 ? n0 : sw + 1, n1, bestv, bestk,
                                                    flags);
}
#pragma clang fp contract(on)

template <typename T, typename TB, int TPF, bool CNT>
__device__ __forceinline__ void ncc_scan_fast(const PairDev &P, const NccRow &q, const T *R, int t, int NX, double &bestv,
                                              long long &bestk, int &flags)
{
    if (q.sqAb) {
        if (q.sqBb) ncc_scan_rows<T, TB, TPF, true, true, CNT>(P, q, R, t, NX, bestv, bestk, flags);
        else ncc_scan_rows<T, TB, TPF, true, false, CNT>(P, q, R, t, NX, bestv, bestk, flags);
    } else {
        if (q.sqBb) ncc_scan_rows<T, TB, TPF, false, true, CNT>(P, q, R, t, NX, bestv, bestk, flags);
        else ncc_scan_rows<T, TB, TPF, false, false, CNT>(P, q, R, t, NX, bestv, bestk, flags);
    }
}

// CorrPatches' R-image value at lag (x, y): the NCC of a candidate lag, 0
// for any other (outside the row's legal interval or failing a per-lag
// rule), as the scan leaves it.
template <typename T, typename TB>
__device__ __forceinline__ double ncc_value(const PairDev &P, int x, int y, double raw)
{
    const NccRow q = ncc_row<TB>(P, y);
    if (x < q.x0 || x > q.x1) return 0.0;
    bool cand, odd;
    return ncc_lag<T, TB>(P, q, x, raw, &cand, &odd);
}

#pragma clang fp contract(off)
// S mode's value of a lag, eval_lag's raw / (Nx Ny): a product of its own
// (never fused into the sums it feeds).
__device__ __forceinline__ double s_value(double raw, double invNxy) { return raw * invNxy; }
#pragma clang fp contract(on)

// One S-mode row (CorrImages / CThmScan sweeps; no exact width, no table
// rule): lag_value's results for each lag, in the same order, with the
// row's work hoisted.  Every non-dead lag's value feeds the sums; a lag is a
// candidate only inside the row's legal x interval -- the region rule's
// terms and the count-area bound, as for NCC rows, without a disc (the lag
// window is the linear range, so the overlap width never clamps) -- and
// there its count corners come from the count tables' band rows, then the
// ellipse (MaskA) and the mincor test.
// band_sum over a count table's band rows, read through global loads.
__device__ __forceinline__ int band_z(const int32_t *top, const int32_t *bot, int L, int R)
{
    int v = ld_z(top, R);
    if (L > 0) v -= ld_z(top, L - 1);
    if (bot) {
        v -= ld_z(bot, R);
        if (L > 0) v += ld_z(bot, L - 1);
    }
    return v;
}

template <typename T, int TPF>
__device__ __forceinline__ void s_scan(const PairDev &P, const RowBox &rb, int y, const T *R, int t, int NX,
                                       bool centred, double &bestv, long long &bestk, double &ssum, double &ssq,
                                       int &flags)
{
    const RuleDev &g = P.rule;
    long long m = 1;
    bool ok = true;
    if (g.rgn_enabled) {
        ok = rb.olh >= g.min_h && (g.exact_h <= 0 || rb.olh == g.exact_h);
        if (g.min_w > m) m = g.min_w;
        if (g.min_area >= 0 && g.min_area / rb.olh + 1 > m) m = g.min_area / rb.olh + 1;
        if ((g.or_w_ge > 0 || g.or_h_gt > 0) && !(rb.olh > g.or_h_gt) && g.or_w_ge > m) m = g.or_w_ge;
    }
    if (g.cnt_enabled) {
        const long long K = g.k1 > g.k2 ? g.k1 : g.k2;
        if (K >= 0 && K / rb.olh + 1 > m) m = K / rb.olh + 1;
    }
    ok = ok && rb.olh > 0 && m <= (P.w1 < P.w2 ? P.w1 : P.w2);
    const int x0 = ok ? (int)(m - P.w1) : 1, x1 = ok ? (int)(P.w2 - m) : 0;
    const int32_t *zAt = nullptr, *zAb = nullptr, *zBt = nullptr, *zBb = nullptr;
    if (g.cnt_enabled) {
        zAt = P.ZA + (long long)rb.a_t * P.w1;
        zBt = P.ZB + (long long)rb.b_t * P.w2;
        zAb = rb.a_b > 0 ? P.ZA + (long long)(rb.a_b - 1) * P.w1 : nullptr;
        zBb = rb.b_b > 0 ? P.ZB + (long long)(rb.b_b - 1) * P.w2 : nullptr;
    }
    const double invNxy = 1.0 / ((double)P.Nx * (double)P.Ny);
    for (int ix = t; ix < NX; ix += TPF) {
        bool dead;
        const int x = lag_x_of_raster(P, ix, &dead);
        if (dead) continue;
        const double val = s_value((double)R[ix], invNxy);
        ssum += val;
        ssq += val * val;
        bool cand = x >= x0 && x <= x1;
        if (cand && g.cnt_enabled) {
            const int BL = x > 0 ? x : 0, BR = (P.w2 - 1 < x + P.w1 - 1) ? P.w2 - 1 : x + P.w1 - 1;
            const int c1 = band_z(zAt, zAb, BL - x, BR - x), c2 = band_z(zBt, zBb, BL, BR);
            cand = (long long)c1 > g.k1 && (long long)c2 > g.k2;
        }
        if (cand && P.Rx > 0 && P.Ry > 0) {
            // MaskA (Correlation.cpp:2011-2054) in centred coordinates
            const int ox = P.Ox - P.offx + P.nnegx, oy = P.Oy - P.offy + P.nnegy;
            const int cx = x + P.nnegx, cy = y + P.nnegy;
            if (cy < oy - P.Ry || cy > oy + P.Ry || cx < ox - P.Rx || cx > ox + P.Rx) cand = false;
            else {
                const double rx2 = (double)P.Rx * P.Rx, asq = rx2 / ((double)P.Ry * P.Ry);
                const int xx = cx - ox, yy = cy - oy;
                if (xx * xx + asq * yy * yy > rx2) cand = false;
            }
        }
        if (cand && P.mincor > 0.0 && !(val > P.mincor)) cand = false;
        if (P.rimg) {
            // the centred R image and its candidate flags (CorrImagesF)
            const long long ir = (long long)(x + P.nnegx) + (long long)P.wR * (y + P.nnegy);
            gst(P.rimg, ir, val);
            gst(P.cimg, ir, (uint8_t)(cand ? 1 : 0));
        }
        if (cand) {
            flags |= 1;
            const long long key = centred ? (long long)(x + P.nnegx) : (long long)ix;
            if (better(val, key, bestv, bestk)) {
                bestv = val;
                bestk = key;
            }
        }
    }
}

// CNT: the variant for launches with count-rule NCC pairs (scans the count
// tables in the fast NCC path); without it such rows take the general path.
template <typename T, int NH, bool CNT, typename TB>
__global__ void __launch_bounds__(RowCfg<NH>::THREADS)
    k_row_c2r_eval(const PairDev *__restrict__ pairs, const int2 *__restrict__ work, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF, G = RowCfg<NH>::G;
    constexpr int H = NH + 1, NX = 2 * NH;
    constexpr int NW = TPF >= 64 ? TPF / 64 : 1;  // wavefronts per row
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    constexpr int RS = C2R_ROW(NH);  // LDS elements per row
    double *red_v = reinterpret_cast<double *>(smem + G * RS);
    long long *red_k = reinterpret_cast<long long *>(red_v + G * NW);
    double *red_s = reinterpret_cast<double *>(red_k + G * NW);
    double *red_q = red_s + G * NW;
    int *red_f = reinterpret_cast<int *>(red_q + G * NW);

    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF, lane = tid & 63;
    const int2 wi = work[xcd_remap(blockIdx.x, gridDim.x)];  // {pair, yl0}
    const PairDev &P = pairs[wi.x];
    const int yl = wi.y + g;
    const bool valid = yl < P.hR && row_live(P, yl);
    const cpx<T> *LG = reinterpret_cast<const cpx<T> *>(P.lag) + (long long)yl * H;
    cpx<T> *X = smem + g * RS;

    // the split's two inputs X[p] and X[NH - p] straight from the row (the mirrored
    // read is the same cache lines, read by another wavefront of the workgroup),
    // every load in flight before the first use
    cpx<T> xk[E], xn[E], ws[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        xk[m] = valid ? gld(LG, p) : cpx<T>{(T)0, (T)0};
        xn[m] = valid ? gld(LG, NH - p) : cpx<T>{(T)0, (T)0};
        ws[m] = twiddle<+1>(tw, p, 2 * NH);
    }
    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const cpx<T> xc = cconj(xn[m]);
        const cpx<T> s = xk[m] + xc, d = xk[m] - xc;
        v[m] = s + rot90<+1>(cmul(ws[m], d));
    }
    LdsView<T> L{X, 1};
    fft<+1, NH>(v, t, L, tw);

    // raw row back to LDS in natural order, then evaluate with unit-stride lanes
    T *R = reinterpret_cast<T *>(X);
    lds_sync();
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        R[2 * p] = v[m].x;
        R[2 * p + 1] = v[m].y;
    }
    lds_sync();

    const int y = lag_y_of_row(P, valid ? yl : 0);
    const RowBox rb = row_box(P, y == DEAD_ROW ? 0 : y);
    const bool centred = P.lag_order == LAG_CENTRED;
    double bestv = -1.0E300;
    long long bestk = 0x7fffffffffffffffLL;
    double ssum = 0.0, ssq = 0.0;
    int flags = 0;
    if (valid && P.mode == PM_NCC) {
        // raster order: x >= 0 first (ix = x), then x < 0 (ix = Nx + x); each thread's
        // keys ascend, so a strict > keeps the first of equal values
        const NccRow q = ncc_row<TB>(P, y);
        const int segs[2][2] = {{q.x0 > 0 ? q.x0 : 0, q.x1}, {q.x0, q.x1 < -1 ? q.x1 : -1}};
        const bool fast = P.rule.exact_w <= 0 && !P.rule.table;
        if (fast && !P.rule.cnt_enabled) ncc_scan_fast<T, TB, TPF, false>(P, q, R, t, NX, bestv, bestk, flags);
        else if (CNT && fast && q.zAt && q.zBt) ncc_scan_fast<T, TB, TPF, CNT>(P, q, R, t, NX, bestv, bestk, flags);
        else for (int sgi = 0; sgi < 2; ++sgi) {
            for (int x = segs[sgi][0] + t; x <= segs[sgi][1]; x += TPF) {
                const int ix = x >= 0 ? x : NX + x;
                bool cand, odd;
                const double val = ncc_lag<T, TB>(P, q, x, (double)R[ix], &cand, &odd);
                if (cand) {
                    flags |= odd ? 3 : 1;
                    if (val > bestv) {
                        bestv = val;
                        bestk = ix;
                    }
                }
            }
        }
    } else if (valid && P.mode == PM_S && P.rule.exact_w <= 0 && !P.rule.table) {
        s_scan<T, TPF>(P, rb, y, R, t, NX, centred, bestv, bestk, ssum, ssq, flags);
    } else if (valid) {
        for (int ix = t; ix < NX; ix += TPF) {
            bool cand, win, odd;
            const double val = lag_value<TB>(P, rb, y, ix, (double)R[ix], &cand, &win, &odd);
            if (odd) flags |= 2;
            if (win) {
                ssum += val;
                ssq += val * val;
                if (P.rimg) {
                    // the centred R image and its candidate flags (CorrImagesF)
                    bool dead;
                    const long long ir = (long long)(lag_x_of_raster(P, ix, &dead) + P.nnegx) +
                                         (long long)P.wR * (y + P.nnegy);
                    gst(P.rimg, ir, val);
                    gst(P.cimg, ir, (uint8_t)(cand ? 1 : 0));
                }
            }
            if (cand) {
                flags |= 1;
                bool dead;
                const long long key = centred ? (long long)(lag_x_of_raster(P, ix, &dead) + P.nnegx) : (long long)ix;
                if (better(val, key, bestv, bestk)) {
                    bestv = val;
                    bestk = key;
                }
            }
        }
    }
    // ---- reduction over the row's TPF threads: shuffles, then LDS across wavefronts ----
    constexpr int SPAN = TPF >= 64 ? 64 : TPF;
#pragma unroll
    for (int off = SPAN / 2; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bestv, off);
        const long long ok = __shfl_xor(bestk, off);
        ssum += __shfl_xor(ssum, off);
        ssq += __shfl_xor(ssq, off);
        flags |= __shfl_xor(flags, off);
        if (better(ov, ok, bestv, bestk)) {
            bestv = ov;
            bestk = ok;
        }
    }
    if constexpr (NW > 1) {
        if (lane == 0) {
            const int wv = t / 64;
            red_v[g * NW + wv] = bestv;
            red_k[g * NW + wv] = bestk;
            red_s[g * NW + wv] = ssum;
            red_q[g * NW + wv] = ssq;
            red_f[g * NW + wv] = flags;
        }
        __syncthreads();
        if (t == 0) {
            for (int k = 1; k < NW; ++k) {
                const int i = g * NW + k;
                if (better(red_v[i], red_k[i], bestv, bestk)) {
                    bestv = red_v[i];
                    bestk = red_k[i];
                }
                ssum += red_s[i];
                ssq += red_q[i];
                flags |= red_f[i];
            }
        }
    }
    if (t == 0 && valid) {
        RowRec r;
        r.best = bestv;
        r.key = (int)((flags & 1) ? bestk : -1);
        r.left = 0.0;
        r.right = 0.0;
        if (flags & 1) {
            // the two in-row neighbours of the best, recomputed from the raw row
            for (int s = 0; s < 2; ++s) {
                const long long kk = bestk + (s ? 1 : -1);
                int ix = -1;
                if (!centred) ix = (int)kk;
                else {
                    const int x = (int)kk - P.nnegx;
                    if (x >= -P.nnegx && x < P.nposx) ix = x >= 0 ? x : P.Nx + x;
                }
                double val = 0.0;
                if (ix >= 0 && ix < NX) {
                    if (P.mode == PM_NCC) {
                        bool dead;
                        const int x = lag_x_of_raster(P, ix, &dead);
                        if (!dead) val = ncc_value<T, TB>(P, x, y, (double)R[ix]);
                    } else {
                        bool cand, win, odd;
                        val = lag_value<TB>(P, rb, y, ix, (double)R[ix], &cand, &win, &odd);
                    }
                }
                if (s) r.right = val;
                else r.left = val;
            }
        }
        r.sum = ssum;
        r.sum2 = ssq;
        r.flags = flags;
        r.pad[0] = r.pad[1] = 0;
        // field by field through a global pointer (a generic store would count as an LDS operation)
        static_assert(offsetof(RowRec, key) == 5 * sizeof(double) && sizeof(RowRec) == 5 * sizeof(double) + 16,
                      "RowRec: five doubles, then key, flags and two pad words");
        double *rd = reinterpret_cast<double *>(P.rows + yl);
        int32_t *ri = reinterpret_cast<int32_t *>(rd + 5);
        gst(rd, 0, r.best);
        gst(rd, 1, r.left);
        gst(rd, 2, r.right);
        gst(rd, 3, r.sum);
        gst(rd, 4, r.sum2);
        gst(ri, 0, r.key);
        gst(ri, 1, r.flags);
        gst(ri, 2, 0);
        gst(ri, 3, 0);
    }
}

// Shared-memory bytes of k_row_c2r_eval.
template <typename T, int NH> constexpr size_t c2r_smem()
{
    constexpr int G = RowCfg<NH>::G, TPF = Plan<NH>::TPF;
    constexpr int NW = TPF >= 64 ? TPF / 64 : 1;
    return sizeof(cpx<T>) * G * C2R_ROW(NH) + G * NW * (sizeof(double) * 3 + sizeof(long long) + sizeof(int)) + 16;
}

// ------------------------------------------------------------------
// Finalize: one workgroup per pair.
// ------------------------------------------------------------------

// Raw c2r value at raster column ix of stored lag row yl, by a direct
// Hermitian DFT of the row (FP64 accumulation); block-reduced.
template <typename T>
__device__ double point_c2r(const PairDev &P, int yl, int ix, const cpx<T> *__restrict__ tw, double *scratch)
{
    const int NH = P.Nx / 2, H = NH + 1;
    const cpx<T> *row = reinterpret_cast<const cpx<T> *>(P.lag) + (long long)yl * H;
    double acc = 0.0;
    for (int kk = threadIdx.x; kk < H; kk += blockDim.x) {
        const cpx<T> X = row[kk];
        const long long e = ((long long)kk * ix) & (P.Nx - 1);
        const cpx<T> w = tw[e * (XC_TW_N / P.Nx)];  // e^{-2 pi i e / Nx}
        // Re(X e^{+2 pi i e/Nx}) = X.x cos + X.y * (-sin)... with w = cos - i sin
        double re = (double)X.x * (double)w.x + (double)X.y * (double)w.y;
        acc += (kk == 0 || kk == NH) ? re : 2.0 * re;
    }
    scratch[threadIdx.x] = acc;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) scratch[threadIdx.x] += scratch[threadIdx.x + off];
        __syncthreads();
    }
    double r = scratch[0];
    __syncthreads();
    return r;
}

template <typename T, typename TB>
__global__ void __launch_bounds__(256) k_finalize(const PairDev *__restrict__ pairs, const cpx<T> *__restrict__ tw)
{
    __shared__ double sv[256], ss[256], sq[256];
    __shared__ long long sk[256];
    __shared__ int sf[256];
    const PairDev &P = pairs[blockIdx.x];
    const int tid = threadIdx.x;
    const int wstride = P.lag_order == LAG_CENTRED ? P.wR : P.Nx;
    double bv = -1.0E300, s1 = 0.0, s2 = 0.0;
    long long bk = 0x7fffffffffffffffLL;
    int fl = 0;
    for (int yl = tid; yl < P.hR; yl += 256) {
        if (!row_live(P, yl)) continue;  // no legal lag: all 0, no candidate
        const RowRec r = P.rows[yl];
        fl |= r.flags;
        s1 += r.sum;
        s2 += r.sum2;
        if (r.flags & 1) {
            const long long key = (long long)yl * wstride + r.key;
            if (better(r.best, key, bv, bk)) {
                bv = r.best;
                bk = key;
            }
        }
    }
    sv[tid] = bv;
    sk[tid] = bk;
    ss[tid] = s1;
    sq[tid] = s2;
    sf[tid] = fl;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) {
            if (better(sv[tid + off], sk[tid + off], sv[tid], sk[tid])) {
                sv[tid] = sv[tid + off];
                sk[tid] = sk[tid + off];
            }
            ss[tid] += ss[tid + off];
            sq[tid] += sq[tid + off];
            sf[tid] |= sf[tid + off];
        }
        __syncthreads();
    }
    const int flags = sf[0];
    const double best = sv[0];
    const long long key = sk[0];
    const double sumS = ss[0], sumS2 = sq[0];
    __syncthreads();

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
    if (P.mode == PM_NCC && (flags & 2)) {
        res.status = XC_E_ODD_R;
    } else if (!(flags & 1)) {
        res.status = XC_NO_PEAK;
    } else {
        res.status = XC_OK;
        const int yl = (int)(key / wstride);
        const int kin = (int)(key - (long long)yl * wstride);
        const RowRec rr = P.rows[yl];
        const int y = lag_y_of_row(P, yl);
        int x;
        if (P.lag_order != LAG_CENTRED) {
            bool dead;
            x = lag_x_of_raster(P, kin, &dead);
        } else {
            x = kin - P.nnegx;
        }
        double sx = 0.0, sy = 0.0;
        const double z1 = best;
        if (P.mode == PM_P2I) {
            // biggest /= norm; exit(44) above 1.0001; ParabPeakFFT on the raw surface
            const double r = best / P.norm;
            if (x - 1 >= 0 && x + 1 < P.Nx_ref && rr.left != 0.0 && rr.right != 0.0) {
                const double d = 1.0 * (rr.left - rr.right) / (2 * (rr.left + rr.right - z1 - z1));
                if (fabs(d) <= 8) sx = d;
            }
            if (y - 1 >= 0 && y + 1 < P.Ny_ref) {
                const double z0 = point_c2r<T>(P, y - 1, x, tw, sv);
                const double z2 = point_c2r<T>(P, y + 1, x, tw, sv);
                if (z0 != 0.0 && z2 != 0.0) {
                    const double d = 1.0 * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
                    if (fabs(d) <= 8) sy = d;
                }
            }
            res.status = r > 1.0001 ? XC_E_ODD_R : XC_OK;
            res.r = r;
            res.peak = best;
            res.dx = x + sx;
            res.dy = y + sy;
        } else if (P.mode == PM_NCC) {
            // ParabPeakFFT (Correlation.cpp:482-517) on the scanned R image
            if (x - 1 >= 0 && x + 1 < P.Nx_ref && rr.left != 0.0 && rr.right != 0.0) {
                const double z0 = rr.left, z2 = rr.right;
                const double d = 1.0 * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
                if (fabs(d) <= 8) sx = d;
            }
            if (y - 1 >= 0 && y + 1 < P.Ny_ref) {
                double zz[2];
                for (int s = 0; s < 2; ++s) {
                    const int iy = y + (s ? 1 : -1);  // raster row (non-negative here)
                    const int ylr = lag_row_of_raster(P, iy);
                    double val = 0.0;  // rows outside the live window hold 0
                    if (ylr >= 0 && row_live(P, ylr)) {
                        const int yy = lag_y_of_row(P, ylr);
                        const double raw = point_c2r<T>(P, ylr, x, tw, sv);
                        val = ncc_value<T, TB>(P, x, yy, raw);  // 0 outside the disc and the rules
                    }
                    zz[s] = val;
                }
                if (zz[0] != 0.0 && zz[1] != 0.0) {
                    const double d = 1.0 * (zz[0] - zz[1]) / (2 * (zz[0] + zz[1] - z1 - z1));
                    if (fabs(d) <= 8) sy = d;
                }
            }
            res.r = best;
            res.peak = best;
            res.dx = x + sx + P.offx;
            res.dy = y + sy + P.offy;
        } else {
            // ParabPeak (Correlation.cpp:2975-3010) on the centred R or S image
            const int rx = x + P.nnegx, ry = y + P.nnegy;
            if (rx - 1 >= 0 && rx + 1 < P.wR && rr.left > 0.0 && rr.right > 0.0) {
                const double z0 = rr.left, z2 = rr.right;
                const double d = 1.0 * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
                if (fabs(d) <= 8) sx = d;
            }
            if (ry - 1 >= 0 && ry + 1 < P.hR) {
                double zz[2];
                for (int s = 0; s < 2; ++s) {
                    const int ylr = ry + (s ? 1 : -1);  // centred storage: row index == ry
                    const int yy = ylr - P.nnegy;
                    const int ix = x >= 0 ? x : P.Nx + x;
                    const double raw = point_c2r<T>(P, ylr, ix, tw, sv);
                    bool legal, odd;
                    const RowBox rb = row_box(P, yy);
                    zz[s] = eval_lag<TB>(P, rb, x, raw, &legal, &odd);
                }
                if (zz[0] > 0.0 && zz[1] > 0.0) {
                    const double d = 1.0 * (zz[0] - zz[1]) / (2 * (zz[0] + zz[1] - z1 - z1));
                    if (fabs(d) <= 8) sy = d;
                }
            }
            res.peak = best;
            res.dx = rx + sx + P.offx - P.nnegx;
            res.dy = ry + sy + P.offy - P.nnegy;
            if (P.mode == PM_S) {
                // Stats (Maths.cpp:111-141, n-1) then 10 (r - mn) / (sd sqrt(n))
                const double n = (double)P.wR * (double)P.hR;
                const double mn = sumS / n;
                const double sd = sqrt(fmax(sumS2 - n * mn * mn, 0.0) / (n - 1.0));
                res.r = 10.0 * (best - mn) / (sd * sqrt(n));
            } else {
                res.r = best;
            }
        }
        res.sub_x = sx;
        res.sub_y = sy;
        res.lag_x = x;
        res.lag_y = y;
    }
    if (tid == 0) *reinterpret_cast<xc_result *>(P.out) = res;
}

}  // namespace xc
