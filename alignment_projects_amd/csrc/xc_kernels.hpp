// xc_kernels.hpp -- HIP kernels of the tile-pair FFT cross-correlation.
//
// Data flow for one correlation (every stage batched over a work list so the
// same kernels serve the plain one-pair FP64 form and the batched FP32 form):
//
//   pack / splat  -> compact A, B rasters (w x h, T)
//   sat_rows/cols -> FP64 summed-area tables S, S^2 and int32 nz
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

#include "../../include/xcorr.h"
#include "xc_fft.hpp"
#include "xc_types.hpp"

namespace xc {

// ------------------------------------------------------------------
// Small helpers
// ------------------------------------------------------------------

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// "a better than b" in the peak order: larger value, then smaller order key.
__device__ __forceinline__ bool better(double va, long long ka, double vb, long long kb)
{
    return va > vb || (va == vb && ka < kb);
}

// Legality of an overlap of olw x olh (the LegalRgn descriptor).
__device__ __forceinline__ bool rgn_ok(const RuleDev &g, int olw, int olh)
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
    TT r = t[R + (long long)w * T];
    if (L > 0) r -= t[L - 1 + (long long)w * T];
    if (B > 0) r -= t[R + (long long)w * (B - 1)];
    if (L > 0 && B > 0) r += t[(L - 1) + (long long)w * (B - 1)];
    return r;
}

// Row part of BoxesFromShifts (Geometry.cpp:150-176) for signed lag y.
struct RowBox {
    int a_b, a_t, b_b, b_t, olh;
};

__device__ __forceinline__ RowBox row_box(const PairDev &P, int y)
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
__device__ __forceinline__ double eval_lag(const PairDev &P, const RowBox &rb, int x, double raw,
                                           bool *legal, bool *odd)
{
    int BL = imax(0, x), BR = imin(P.w2 - 1, x + P.w1 - 1);
    BR = imax(BR, BL);
    int AL = BL - x, AR = BR - x;
    int olw = BR - BL + 1;
    bool ok = rgn_ok(P.rule, olw, rb.olh);
    *odd = false;
    double Nxy = (double)P.Nx * (double)P.Ny;
    if (P.mode == PM_NCC) {
        if (!ok) { *legal = false; return 0.0; }
        if (P.rule.cnt_enabled) {
            int c1 = box_sum(P.ZA, P.w1, AL, AR, rb.a_b, rb.a_t);
            int c2 = box_sum(P.ZB, P.w2, BL, BR, rb.b_b, rb.b_t);
            if (!((long long)c1 > P.rule.k1 && (long long)c2 > P.rule.k2)) { *legal = false; return 0.0; }
        }
    } else if (P.rule.cnt_enabled) {
        int c1 = box_sum(P.ZA, P.w1, AL, AR, rb.a_b, rb.a_t);
        int c2 = box_sum(P.ZB, P.w2, BL, BR, rb.b_b, rb.b_t);
        if (!((long long)c1 > P.rule.k1 && (long long)c2 > P.rule.k2)) ok = false;
    }
    *legal = ok;
    if (P.mode == PM_S) return raw / Nxy;
    double n = (double)(olw * rb.olh);
    double s1 = box_sum(P.SA, P.w1, AL, AR, rb.a_b, rb.a_t);
    double q1 = box_sum(P.QA, P.w1, AL, AR, rb.a_b, rb.a_t);
    double s2 = box_sum(P.SB, P.w2, BL, BR, rb.b_b, rb.b_t);
    double q2 = box_sum(P.QB, P.w2, BL, BR, rb.b_b, rb.b_t);
    double num = n * raw / Nxy - s1 * s2;
    double d1 = n * q1 - s1 * s1;
    double d2 = n * q2 - s2 * s2;
    double d = d1 * d2;
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

__device__ __forceinline__ int lag_y_of_row(const PairDev &P, int yl)
{
    if (P.lag_order == LAG_ALL) {
        if (yl < P.nposy) return yl;
        if (yl >= P.Ny - P.nnegy) return yl - P.Ny;
        return DEAD_ROW;
    }
    if (P.lag_order == LAG_RASTER) return yl < P.nposy ? yl : yl - P.nposy - P.nnegy;
    return yl - P.nnegy;
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

// Value of pixel (x, y) in transform precision: (raw - mean) / sd rounded to
// T, 0 outside the image or where the mask says absent.
template <typename T>
__device__ __forceinline__ T load_val(const ImgView &v, double mean, double sd, int x, int y)
{
    if (x >= v.w || y >= v.h || !present(v, x, y)) return (T)0;
    return (T)((load_raw(v, x, y) - mean) / sd);
}

// ------------------------------------------------------------------
// Image preparation: statistics (Stats/Normalize, Maths.cpp:111-212) and
// the summed-area tables of IntegrateImage (Correlation.cpp:49-105), built
// in one write pass: 64-column strips scanned top to bottom by one
// wavefront, with the left-of-strip contribution carried in from per-strip
// row totals.
// ------------------------------------------------------------------

// Partial sums of present raw values: TP_PARTS workgroups per image.
static __global__ void __launch_bounds__(256) k_tp_stats(const TileDev *__restrict__ tiles)
{
    __shared__ double ss[256], sq[256], sc[256];
    const TileDev &D = tiles[blockIdx.x / TP_PARTS];
    const int part = blockIdx.x % TP_PARTS, tid = threadIdx.x;
    const long long n = (long long)D.v.w * D.v.h;
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long i = (long long)part * 256 + tid; i < n; i += (long long)TP_PARTS * 256) {
        const int y = (int)(i / D.v.w), x = (int)(i - (long long)y * D.v.w);
        if (present(D.v, x, y)) {
            const double t = load_raw(D.v, x, y);
            a += t;
            b += t * t;
            c += 1.0;
        }
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
}

// Strip row totals of v, v^2, (v != 0): work item = (tile, strip).
template <typename T>
__global__ void __launch_bounds__(256) k_tp_rowtot(const TileDev *__restrict__ tiles, const int2 *__restrict__ work)
{
    const int2 wi = work[blockIdx.x];
    const TileDev &D = tiles[wi.x];
    const int s = wi.y, h = D.v.h, x0 = s * TP_STRIP, x1 = imin(D.v.w, x0 + TP_STRIP);
    const double mean = D.msd[0], sd = D.msd[1];
    for (int y = threadIdx.x; y < h; y += blockDim.x) {
        double a = 0.0, b = 0.0, c = 0.0;
        for (int x = x0; x < x1; ++x) {
            const double t = (double)load_val<T>(D.v, mean, sd, x, y);
            a += t;
            b += t * t;
            c += (t != 0.0);
        }
        const long long o = (long long)s * h + y, plane = (long long)D.nstrips * h;
        D.rowtot[o] = a;
        D.rowtot[plane + o] = b;
        D.rowtot[2 * plane + o] = c;
    }
}

// Carries: carry[s][y] = sum_{y' <= y} sum_{s' < s} rowtot[s'][y'].  One
// workgroup per image, one thread per strip (<= 256 strips).
static __global__ void __launch_bounds__(256) k_tp_carry(const TileDev *__restrict__ tiles)
{
    __shared__ double sa[256], sb[256], sc[256];
    const TileDev &D = tiles[blockIdx.x];
    const int s = threadIdx.x, ns = D.nstrips, h = D.v.h;
    const long long plane = (long long)ns * h;
    double aa = 0.0, ab = 0.0, ac = 0.0;
    for (int y = 0; y < h; ++y) {
        const long long o = (long long)s * h + y;
        sa[s] = s < ns ? D.rowtot[o] : 0.0;
        sb[s] = s < ns ? D.rowtot[plane + o] : 0.0;
        sc[s] = s < ns ? D.rowtot[2 * plane + o] : 0.0;
        __syncthreads();
        double ea = 0.0, eb = 0.0, ec = 0.0;
        for (int k = 0; k < s && k < ns; ++k) {
            ea += sa[k];
            eb += sb[k];
            ec += sc[k];
        }
        aa += ea;
        ab += eb;
        ac += ec;
        if (s < ns) {
            D.carry[o] = aa;
            D.carry[plane + o] = ab;
            D.carry[2 * plane + o] = ac;
        }
        __syncthreads();
    }
}

// Strip scan: one wavefront per (tile, strip); lane = column.
template <typename T>
__global__ void __launch_bounds__(64) k_tp_scan(const TileDev *__restrict__ tiles, const int2 *__restrict__ work)
{
    const int2 wi = work[blockIdx.x];
    const TileDev &D = tiles[wi.x];
    const int s = wi.y, lane = threadIdx.x, w = D.v.w, h = D.v.h;
    const int x = s * TP_STRIP + lane;
    const bool in = x < w;
    const double mean = D.msd[0], sd = D.msd[1];
    const long long plane = (long long)D.nstrips * h;
    double cs = 0.0, cq = 0.0;
    int cz = 0;
    for (int y = 0; y < h; ++y) {
        const double t = in ? (double)load_val<T>(D.v, mean, sd, x, y) : 0.0;
        cs += t;
        cq += t * t;
        cz += (t != 0.0);
        double ps = cs, pq = cq;
        int pz = cz;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double us = __shfl_up(ps, off), uq = __shfl_up(pq, off);
            const int uz = __shfl_up(pz, off);
            if (lane >= off) {
                ps += us;
                pq += uq;
                pz += uz;
            }
        }
        if (in) {
            const long long o = (long long)s * h + y, i = (long long)y * w + x;
            D.S[i] = D.carry[o] + ps;
            D.Q[i] = D.carry[plane + o] + pq;
            D.Z[i] = (int32_t)D.carry[2 * plane + o] + pz;
        }
    }
}

// ------------------------------------------------------------------
// Row r2c: one Nx-point real transform per image row, via an Nx/2-point
// complex transform and the split post-twiddle.
// ------------------------------------------------------------------

template <int NH> struct RowCfg {
    static constexpr int TPF = Plan<NH>::TPF;
    static constexpr int G = TPF >= 256 ? 1 : 256 / TPF;  // rows per workgroup
    static constexpr int THREADS = TPF * G;
};

template <typename T, int NH>
__global__ void __launch_bounds__(RowCfg<NH>::THREADS)
    k_row_r2c(const PairDev *__restrict__ pairs, const int4 *__restrict__ work, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    const int4 wi = work[blockIdx.x];  // {pair, image, y0, -}
    const PairDev &P = pairs[wi.x];
    const int img = wi.y, y = wi.z + g;
    const ImgView &V = img ? P.B : P.A;
    cpx<T> *dst = reinterpret_cast<cpx<T> *>(img ? P.specB : P.specA);
    const bool valid = y < V.h;
    const int H = NH + 1;
    const double mean = V.msd[0], sd = V.msd[1];

    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        v[m].x = valid ? load_val<T>(V, mean, sd, 2 * p, y) : (T)0;
        v[m].y = valid ? load_val<T>(V, mean, sd, 2 * p + 1, y) : (T)0;
    }
    LdsView<T> L{smem + g * NH, 1};
    fft<-1, NH>(v, t, L, tw);

    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) L[t + TPF * m] = v[m];
    __syncthreads();
    if (valid) {
        cpx<T> *row = dst + (long long)y * H;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int k = t + TPF * m;
            const cpx<T> zk = v[m];
            const cpx<T> zn = cconj(L[(NH - k) & (NH - 1)]);
            const cpx<T> a = cscale(zk + zn, (T)0.5);
            const cpx<T> b = rot90<-1>(cscale(zk - zn, (T)0.5));
            row[k] = a + cmul(twiddle<-1>(tw, k, 2 * NH), b);
        }
        if (t == 0) row[NH] = cpx<T>{v[0].x - v[0].y, (T)0};
    }
}

// ------------------------------------------------------------------
// Column pass, fused: forward Ny-point FFT of the A and B spectrum
// columns, cross power F_B conj(F_A) (Correlation.cpp:972-986; whitened
// for S mode, :2311-2317), inverse Ny-point FFT, write the hR lag rows.
// ------------------------------------------------------------------

constexpr int cmin(int a, int b) { return a < b ? a : b; }

template <typename T, int NY> struct ColCfg {
    static constexpr int TPF = Plan<NY>::TPF;
    static constexpr int CW = sizeof(T) == 8 ? 4 : 8;  // columns per workgroup wanted (64 B rows)
    static constexpr int CT = 1024 / TPF > 0 ? 1024 / TPF : 1;                       // <= 1024 threads
    static constexpr int CL = 65536 / (NY * (int)sizeof(cpx<T>)) > 0 ? 65536 / (NY * (int)sizeof(cpx<T>)) : 1;  // <= 64 KB LDS
    static constexpr int C = cmin(CW, cmin(CT, CL));
    static constexpr int THREADS = TPF * C;
};

template <typename T, int NY, bool WHITEN>
__global__ void __launch_bounds__((ColCfg<T, NY>::THREADS))
    k_col_fused(const PairDev *__restrict__ pairs, const int2 *__restrict__ work, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NY>;
    constexpr int E = PL::E, TPF = PL::TPF, C = ColCfg<T, NY>::C;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    const int2 wi = work[blockIdx.x];  // {pair, k0}
    const PairDev &P = pairs[wi.x];
    const int H = P.Nx / 2 + 1;
    const int k = wi.y + c;
    const bool kv = k < H;
    const cpx<T> *SA = reinterpret_cast<const cpx<T> *>(P.specA);
    const cpx<T> *SB = reinterpret_cast<const cpx<T> *>(P.specB);
    cpx<T> *LG = reinterpret_cast<cpx<T> *>(P.lag);
    LdsView<T> L{smem + c, C};

    cpx<T> a[E], b[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int y = t + TPF * m;
        a[m] = (kv && y < P.A.h) ? SA[(long long)y * H + k] : cpx<T>{(T)0, (T)0};
    }
    fft<-1, NY>(a, t, L, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int y = t + TPF * m;
        b[m] = (kv && y < P.B.h) ? SB[(long long)y * H + k] : cpx<T>{(T)0, (T)0};
    }
    __syncthreads();
    fft<-1, NY>(b, t, L, tw);
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
            T mag = hypot(p.x, p.y);
            if (mag > (T)1e-10) {
                T s = sqrt(mag);
                p.x /= s;
                p.y /= s;
            }
        }
        a[m] = p;
    }
    __syncthreads();
    fft<+1, NY>(a, t, L, tw);
    if (kv) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int iy = t + TPF * m;
            const int yl = lag_row_of_raster(P, iy);
            if (yl >= 0) LG[(long long)yl * H + k] = a[m];
        }
    }
}

// ------------------------------------------------------------------
// Row c2r + evaluation: per stored lag row, Nx-point c2r (Nx/2-point
// complex inverse after the split pre-twiddle), mode value per lag, and
// the row's best candidate with its two in-row neighbours.
// ------------------------------------------------------------------

struct RedItem {
    double v;
    long long k;
};

template <typename T, int NH>
__global__ void __launch_bounds__(RowCfg<NH>::THREADS)
    k_row_c2r_eval(const PairDev *__restrict__ pairs, const int2 *__restrict__ work, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF, G = RowCfg<NH>::G;
    constexpr int H = NH + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<T> *smem = reinterpret_cast<cpx<T> *>(smem_raw);
    // reduction scratch after the G spectra (H + 1 entries each, keeps 16 B alignment)
    double *red_v = reinterpret_cast<double *>(smem + G * (H + 1));
    long long *red_k = reinterpret_cast<long long *>(red_v + G * TPF);
    double *red_s = reinterpret_cast<double *>(red_k + G * TPF);
    double *red_q = red_s + G * TPF;
    double *nb = red_q + G * TPF;  // [G][2] neighbour values

    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    const int2 wi = work[blockIdx.x];  // {pair, yl0}
    const PairDev &P = pairs[wi.x];
    const int yl = wi.y + g;
    const bool valid = yl < P.hR;
    const cpx<T> *LG = reinterpret_cast<const cpx<T> *>(P.lag) + (long long)yl * H;
    cpx<T> *X = smem + g * (H + 1);

    for (int i = t; i < H; i += TPF) X[i] = valid ? LG[i] : cpx<T>{(T)0, (T)0};
    __syncthreads();

    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        const cpx<T> xk = X[p];
        const cpx<T> xn = cconj(X[NH - p]);
        const cpx<T> s = xk + xn, d = xk - xn;
        v[m] = s + rot90<+1>(cmul(twiddle<+1>(tw, p, 2 * NH), d));
    }
    LdsView<T> L{X, 1};
    fft<+1, NH>(v, t, L, tw);

    // ---- evaluation ----
    const int y = lag_y_of_row(P, valid ? yl : 0);
    const RowBox rb = row_box(P, y == DEAD_ROW ? 0 : y);
    const bool centred = P.lag_order == LAG_CENTRED;
    const int dyd = y - P.Oy;
    const long long rsqr = (long long)P.radius * P.radius;
    double vals[2 * E];
    double bestv = -1.0E300;
    long long bestk = 0x7fffffffffffffffLL;
    double ssum = 0.0, ssq = 0.0;
    int flags = 0;
    // oval (MaskA, Correlation.cpp:2011-2054) in centred coordinates
    const bool oval = centred && P.Rx > 0 && P.Ry > 0;
    const int ox = P.Ox - P.offx + P.nnegx, oy = P.Oy - P.offy + P.nnegy;
    const double rx2 = (double)P.Rx * P.Rx, asq = rx2 / ((double)P.Ry * P.Ry);
    const int cyy = y + P.nnegy;
#pragma unroll
    for (int m = 0; m < E; ++m) {
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const int ix = 2 * (t + TPF * m) + hf;
            const double raw = hf ? (double)v[m].y : (double)v[m].x;
            double val = 0.0;
            bool cand = false;
            bool dead;
            const int x = lag_x_of_raster(P, ix, &dead);
            if (P.mode == PM_P2I) {
                // CorrPatchToImage scan (Correlation.cpp:1299-1327): raw surface, zones, disc
                val = raw;
                if (valid && !dead && y != DEAD_ROW) {
                    const int dxd = x - P.Ox;
                    cand = (long long)dxd * dxd + (long long)dyd * dyd <= rsqr;
                }
            } else if (valid && !dead) {
                if (!centred) {
                    // CorrPatches scan: disc about (Ox, Oy), then the rules
                    const int dxd = x - P.Ox;
                    if ((long long)dxd * dxd + (long long)dyd * dyd <= rsqr) {
                        bool odd;
                        val = eval_lag(P, rb, x, raw, &cand, &odd);
                        if (odd) flags |= 2;
                    }
                } else {
                    bool odd;
                    val = eval_lag(P, rb, x, raw, &cand, &odd);
                    if (P.mode == PM_S) {
                        ssum += val;
                        ssq += val * val;
                    }
                    if (oval) {
                        const int cxx = x + P.nnegx;
                        if (cyy < oy - P.Ry || cyy > oy + P.Ry || cxx < ox - P.Rx || cxx > ox + P.Rx) cand = false;
                        else {
                            const int xx = cxx - ox, yy = cyy - oy;
                            if (xx * xx + asq * yy * yy > rx2) cand = false;
                        }
                    }
                    if (P.mincor > 0.0 && !(val > P.mincor)) cand = false;
                }
            }
            vals[2 * m + hf] = val;
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
    // ---- reduction over the row's TPF threads ----
    double *rv = red_v + g * TPF;
    long long *rk = red_k + g * TPF;
    double *rs = red_s + g * TPF;
    double *rq = red_q + g * TPF;
    int *rf = reinterpret_cast<int *>(nb + 2 * G) + g * TPF;
    __syncthreads();
    rv[t] = bestv;
    rk[t] = bestk;
    rs[t] = ssum;
    rq[t] = ssq;
    rf[t] = flags;
    __syncthreads();
    for (int off = TPF / 2; off > 0; off >>= 1) {
        if (t < off) {
            if (better(rv[t + off], rk[t + off], rv[t], rk[t])) {
                rv[t] = rv[t + off];
                rk[t] = rk[t + off];
            }
            rs[t] += rs[t + off];
            rq[t] += rq[t + off];
            rf[t] |= rf[t + off];
        }
        __syncthreads();
    }
    const long long bk = rk[0];
    // neighbours key-1, key+1 in the same row (raster ix for NCC, centred x otherwise)
    if (t == 0) {
        nb[2 * g] = 0.0;
        nb[2 * g + 1] = 0.0;
    }
    __syncthreads();
    if (rf[0] & 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const int ix = 2 * (t + TPF * m) + hf;
                bool dead;
                const int x = lag_x_of_raster(P, ix, &dead);
                long long key;
                if (!centred) key = ix;
                else key = dead ? -100 : (long long)(x + P.nnegx);
                if (key == bk - 1) nb[2 * g] = vals[2 * m + hf];
                if (key == bk + 1) nb[2 * g + 1] = vals[2 * m + hf];
            }
        }
    }
    __syncthreads();
    if (t == 0 && valid) {
        RowRec r;
        r.best = rv[0];
        r.key = (int)(rf[0] & 1 ? bk : -1);
        r.left = nb[2 * g];
        r.right = nb[2 * g + 1];
        r.sum = rs[0];
        r.sum2 = rq[0];
        r.flags = rf[0];
        r.pad[0] = r.pad[1] = 0;
        P.rows[yl] = r;
    }
}

// Shared-memory bytes of k_row_c2r_eval.
template <typename T, int NH> constexpr size_t c2r_smem()
{
    constexpr int G = RowCfg<NH>::G, TPF = Plan<NH>::TPF;
    return sizeof(cpx<T>) * G * (NH + 2) + G * TPF * (sizeof(double) * 3 + sizeof(long long) + sizeof(int)) +
           sizeof(double) * 2 * G + 16;
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

template <typename T>
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
                    double val = 0.0;
                    if (ylr >= 0) {
                        const int yy = lag_y_of_row(P, ylr);
                        const double raw = point_c2r<T>(P, ylr, x, tw, sv);
                        const int dxd = x - P.Ox, dyd = yy - P.Oy;
                        if ((long long)dxd * dxd + (long long)dyd * dyd <= (long long)P.radius * P.radius) {
                            bool legal, odd;
                            const RowBox rb = row_box(P, yy);
                            const double r = eval_lag(P, rb, x, raw, &legal, &odd);
                            val = legal ? r : 0.0;
                        }
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
                    zz[s] = eval_lag(P, rb, x, raw, &legal, &odd);
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
