// xc_seam.hip -- the FFT seam (FFT_2D / IFT_2D, Correlation_fft_fftw.cpp:10-92)
// and CorrPatchToImage (Correlation.cpp:1186-1355) in FP64 on the device.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "xc_exec.hpp"
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

// Row r2c of a dense real array in[rows][2*NH] into out[rows][NH+1].
template <typename T, int NH>
__global__ void __launch_bounds__(RowCfg<NH>::THREADS)
    k_seam_r2c(const T *__restrict__ in, int rows, cpx<T> *__restrict__ out, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF, G = RowCfg<NH>::G;
    __shared__ cpx<T> smem[G * lds_len(NH)];
    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    const int y = blockIdx.x * G + g;
    const bool valid = y < rows;
    const T *src = in + (long long)y * 2 * NH;
    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        v[m].x = valid ? src[2 * p] : (T)0;
        v[m].y = valid ? src[2 * p + 1] : (T)0;
    }
    LdsView<T> L{smem + g * lds_len(NH), 1};
    fft<-1, NH>(v, t, L, tw);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) L[t + TPF * m] = v[m];
    __syncthreads();
    if (valid) {
        cpx<T> *row = out + (long long)y * (NH + 1);
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

// In-place column transform of data[NY][H] (DIR -1 forward, +1 inverse).
template <typename T, int NY, int DIR>
__global__ void __launch_bounds__((ColCfg<T, NY>::THREADS))
    k_seam_col(cpx<T> *__restrict__ data, int H, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NY>;
    constexpr int E = PL::E, TPF = PL::TPF, C = ColCfg<T, NY>::C;
    __shared__ cpx<T> smem[C * lds_len(NY)];
    const int tid = threadIdx.x, c = tid % C, t = tid / C;
    const int k = blockIdx.x * C + c;
    const bool kv = k < H;
    LdsView<T> L{smem + c, C};
    cpx<T> a[E];
#pragma unroll
    for (int m = 0; m < E; ++m) a[m] = kv ? data[(long long)(t + TPF * m) * H + k] : cpx<T>{(T)0, (T)0};
    fft<DIR, NY>(a, t, L, tw);
    if (kv) {
#pragma unroll
        for (int m = 0; m < E; ++m) data[(long long)(t + TPF * m) * H + k] = a[m];
    }
}

// Row c2r of in[rows][NH+1] into out[rows][2*NH] (unnormalised).
template <typename T, int NH>
__global__ void __launch_bounds__(RowCfg<NH>::THREADS)
    k_seam_c2r(const cpx<T> *__restrict__ in, int rows, T *__restrict__ out, const cpx<T> *__restrict__ tw)
{
    using PL = Plan<NH>;
    constexpr int E = PL::E, TPF = PL::TPF, G = RowCfg<NH>::G;
    __shared__ cpx<T> smem[G * C2R_ROW(NH)];
    const int tid = threadIdx.x, g = tid / TPF, t = tid % TPF;
    const int y = blockIdx.x * G + g;
    const bool valid = y < rows;
    cpx<T> *X = smem + g * C2R_ROW(NH);
    for (int i = t; i < NH + 1; i += TPF) X[i] = valid ? in[(long long)y * (NH + 1) + i] : cpx<T>{(T)0, (T)0};
    __syncthreads();
    cpx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int p = t + TPF * m;
        const cpx<T> xk = X[p], xn = cconj(X[NH - p]);
        v[m] = (xk + xn) + rot90<+1>(cmul(twiddle<+1>(tw, p, 2 * NH), xk - xn));
    }
    LdsView<T> L{X, 1};
    fft<+1, NH>(v, t, L, tw);
    if (valid) {
        T *dst = out + (long long)y * 2 * NH;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int p = t + TPF * m;
            dst[2 * p] = v[m].x;
            dst[2 * p + 1] = v[m].y;
        }
    }
}

// CorrPatchToImage's raster of patch1: i1[x + N*y] = iv1[i] at the points'
// truncated coordinates, later points overwriting earlier ones.
__global__ void k_p2i_claim(const xc_point *__restrict__ pts, int n, int N, int *__restrict__ owner)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)pts[i].x, y = (int)pts[i].y;
    if (x >= 0 && x < N && y >= 0 && y < N) atomicMax(&owner[x + (long long)N * y], i);
}

__global__ void k_p2i_assign(const xc_point *__restrict__ pts, const double *__restrict__ v, int n, int N,
                             const int *__restrict__ owner, double *__restrict__ img)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)pts[i].x, y = (int)pts[i].y;
    if (x >= 0 && x < N && y >= 0 && y < N && owner[x + (long long)N * y] == i) img[x + (long long)N * y] = v[i];
}

#define XC_SEAM_CASES(F) \
    F(1) F(2) F(4) F(8) F(16) F(32) F(64) F(128) F(256) F(512) F(1024) F(2048)

static int seam_r2c(xc_ctx *ctx, int NH, const double *in, int rows, cpx<double> *out)
{
    switch (NH) {
#define F(n)                                                                                                \
    case n:                                                                                                 \
        hipLaunchKernelGGL((k_seam_r2c<double, n>), dim3((rows + RowCfg<n>::G - 1) / RowCfg<n>::G),          \
                           dim3(RowCfg<n>::THREADS), 0, ctx->stream, in, rows, out, ctx->tw64);             \
        return XC_OK;
        XC_SEAM_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "row length %d unsupported", 2 * NH);
}

static int seam_c2r(xc_ctx *ctx, int NH, const cpx<double> *in, int rows, double *out)
{
    switch (NH) {
#define F(n)                                                                                                \
    case n:                                                                                                 \
        hipLaunchKernelGGL((k_seam_c2r<double, n>), dim3((rows + RowCfg<n>::G - 1) / RowCfg<n>::G),          \
                           dim3(RowCfg<n>::THREADS), 0, ctx->stream, in, rows, out, ctx->tw64);             \
        return XC_OK;
        XC_SEAM_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "row length %d unsupported", 2 * NH);
}

template <int DIR> static int seam_col(xc_ctx *ctx, int NY, cpx<double> *data, int H)
{
    switch (NY) {
#define F(n)                                                                                                     \
    case n:                                                                                                      \
        hipLaunchKernelGGL((k_seam_col<double, n, DIR>), dim3((H + ColCfg<double, n>::C - 1) / ColCfg<double, n>::C), \
                           dim3(ColCfg<double, n>::THREADS), 0, ctx->stream, data, H, ctx->tw64);                \
        return XC_OK;
        XC_SEAM_CASES(F) F(4096)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "column length %d unsupported", NY);
}

static bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

}  // namespace xc

using namespace xc;

extern "C" int xc_fft2d_r2c_f64(xc_ctx *ctx, const double *in, int Nfast, int Nslow, double *out, int *M_out)
{
    if (!ctx || !in || !out) return XC_E_ARG;
    if (!pow2(Nfast) || Nfast < 2 || !pow2(Nslow) || Nfast / 2 > max_nh<double>() || Nslow > max_ny<double>())
        return set_err(ctx, XC_E_SHAPE, "FFT_2D: %d x %d (powers of two, Nfast >= 2, <= %d x %d)", Nfast, Nslow,
                       2 * max_nh<double>(), max_ny<double>());
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const int NH = Nfast / 2, H = NH + 1;
    const size_t nin = (size_t)Nfast * Nslow, M = (size_t)Nslow * H;
    Layout lay;
    size_t i_in = lay.add(sizeof(double) * nin), i_out = lay.add(sizeof(cpx<double>) * M);
    int st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    double *din = (double *)(base + lay.off[i_in]);
    cpx<double> *dout = (cpx<double> *)(base + lay.off[i_out]);
    XC_HIP(ctx, hipMemcpyAsync(din, in, sizeof(double) * nin, hipMemcpyHostToDevice, ctx->stream));
    if ((st = seam_r2c(ctx, NH, din, Nslow, dout))) return st;
    if ((st = seam_col<-1>(ctx, Nslow, dout, H))) return st;
    XC_HIP(ctx, hipGetLastError());
    XC_HIP(ctx, hipMemcpyAsync(out, dout, sizeof(cpx<double>) * M, hipMemcpyDeviceToHost, ctx->stream));
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (M_out) *M_out = (int)M;
    return XC_OK;
}

extern "C" int xc_ift2d_c2r_f64(xc_ctx *ctx, const double *in, int Nfast, int Nslow, double *out)
{
    if (!ctx || !in || !out) return XC_E_ARG;
    if (!pow2(Nfast) || Nfast < 2 || !pow2(Nslow) || Nfast / 2 > max_nh<double>() || Nslow > max_ny<double>())
        return set_err(ctx, XC_E_SHAPE, "IFT_2D: %d x %d unsupported", Nfast, Nslow);
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const int NH = Nfast / 2, H = NH + 1;
    const size_t nout = (size_t)Nfast * Nslow, M = (size_t)Nslow * H;
    Layout lay;
    size_t i_in = lay.add(sizeof(cpx<double>) * M), i_out = lay.add(sizeof(double) * nout);
    int st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    cpx<double> *din = (cpx<double> *)(base + lay.off[i_in]);
    double *dout = (double *)(base + lay.off[i_out]);
    XC_HIP(ctx, hipMemcpyAsync(din, in, sizeof(cpx<double>) * M, hipMemcpyHostToDevice, ctx->stream));
    if ((st = seam_col<+1>(ctx, Nslow, din, H))) return st;
    if ((st = seam_c2r(ctx, NH, din, Nslow, dout))) return st;
    XC_HIP(ctx, hipGetLastError());
    XC_HIP(ctx, hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XC_OK;
}

extern "C" int xc_corr_patch_to_image_f64(xc_ctx *ctx, const xc_point *ip1, const double *iv1, int n1,
                                          const double *i2, int n_i2, int Ox, int Oy, int radius, int bFilter,
                                          double *dx, double *dy, double *r, xc_result *res)
{
    if (!ctx || !ip1 || !iv1 || !i2 || n1 <= 0 || n_i2 <= 0) return set_err(ctx, XC_E_ARG, "null or empty input");
    const int N = (int)sqrt((double)n_i2);
    if (!pow2(N) || N < 2 || N / 2 > max_nh<double>() || N > max_ny<double>())
        return set_err(ctx, XC_E_SHAPE, "CorrPatchToImage: N = %d must be a power of two in [2, %d]", N,
                       2 * max_nh<double>());
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    IBox B1;
    bbox_from_points(B1, ip1, n1);
    const int w1 = B1.R - B1.L + 1, h1 = B1.T - B1.B + 1;
    // effective extent of the caller-embedded image (Correlation.cpp:1215-1233)
    int w2 = 0, h2 = 0;
    for (int i = 0; i < n_i2; ++i)
        if (i2[i] != 0.0) {
            const int y = i / N, x = i - N * y;
            h2 = std::max(h2, y);
            w2 = std::max(w2, x);
        }
    ++w2;
    ++h2;

    PairDev P;
    memset(&P, 0, sizeof P);
    setup_pair_geometry(P, w1, h1, w2, h2, PM_P2I);
    P.Nx = P.Nx_ref = N;
    P.Ny = P.Ny_ref = N;
    P.lag_order = LAG_ALL;
    P.hR = N;
    P.Ox = Ox;
    P.Oy = Oy;
    P.radius = radius;
    P.filter = bFilter ? 1 : 0;
    P.norm = (double)n1 * n_i2;
    const int NH = N / 2, H = NH + 1, G = row_group<double>(NH), C = col_group<double>(N), GR = r2c_rows<double>(NH);
    const int nr2c = 2 * ((N + GR - 1) / GR), ncol = (H + C - 1) / C, nc2r = (N + G - 1) / G;

    Layout lay;
    size_t i_pts = lay.add(sizeof(xc_point) * n1), i_vals = lay.add(sizeof(double) * n1);
    size_t i_i2 = lay.add(sizeof(double) * n_i2), i_i1 = lay.add(sizeof(double) * n_i2);
    size_t i_own = lay.add(sizeof(int) * n_i2), i_msd = lay.add(sizeof(double) * 3);
    size_t i_spA = lay.add(sizeof(cpx<double>) * N * H), i_spB = lay.add(sizeof(cpx<double>) * N * H);
    size_t i_lag = lay.add(sizeof(cpx<double>) * N * H), i_rows = lay.add(sizeof(RowRec) * N);
    size_t i_res = lay.add(sizeof(xc_result)), i_pair = lay.add(sizeof(PairDev));
    size_t i_wr = lay.add(sizeof(int4) * nr2c), i_wc = lay.add(sizeof(int2) * ncol), i_wc2 = lay.add(sizeof(int2) * nc2r);
    int st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };
    for (int k = 0; k < 2; ++k) {
        ImgView &V = k ? P.B : P.A;
        V.data = at(k ? i_i2 : i_i1);
        V.dtype = XC_F64;
        V.w = N;
        V.h = N;
        V.stride = N;
        V.mstride = N;
        V.msd = (const double *)at(i_msd);
    }
    P.specA = at(i_spA);
    P.specB = at(i_spB);
    P.lag = at(i_lag);
    P.rows = (RowRec *)at(i_rows);
    P.out = at(i_res);

    const size_t up_lo = lay.off[i_pair], up_hi = lay.total;
    st = ensure_pinned(ctx, (up_hi - up_lo) + 64 + sizeof(xc_result));
    if (st) return st;
    char *hs = (char *)ctx->pinned;
    auto hp = [&](size_t i) { return hs + (lay.off[i] - up_lo); };
    memcpy(hp(i_pair), &P, sizeof P);
    int4 *wr = (int4 *)hp(i_wr);
    int k = 0;
    for (int img = 0; img < 2; ++img)
        for (int y = 0; y < N; y += GR) wr[k++] = make_int4(0, img, y, 0);
    int2 *wc = (int2 *)hp(i_wc);
    for (int i = 0; i < ncol; ++i) wc[i] = make_int2(0, i * C);
    int2 *wc2 = (int2 *)hp(i_wc2);
    for (int i = 0; i < nc2r; ++i) wc2[i] = make_int2(0, i * G);
    xc_result *hres = (xc_result *)(((uintptr_t)(hs + (up_hi - up_lo)) + 63) & ~(uintptr_t)63);

    hipStream_t s = ctx->stream;
    const double msd[3] = {0.0, 1.0, 1.0};
    XC_HIP(ctx, hipMemcpyAsync(base + up_lo, hs, up_hi - up_lo, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_msd), msd, sizeof msd, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_pts), ip1, sizeof(xc_point) * n1, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_vals), iv1, sizeof(double) * n1, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_i2), i2, sizeof(double) * n_i2, hipMemcpyHostToDevice, s));
    const int fb = (int)std::min<long long>(1024, ((long long)n_i2 + 255) / 256);
    hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, s, (double *)at(i_i1), (long long)n_i2, 0.0);
    hipLaunchKernelGGL(k_fill<int>, dim3(fb), dim3(256), 0, s, (int *)at(i_own), (long long)n_i2, -1);
    hipLaunchKernelGGL(k_p2i_claim, dim3((n1 + 255) / 256), dim3(256), 0, s, (const xc_point *)at(i_pts), n1, N,
                       (int *)at(i_own));
    hipLaunchKernelGGL(k_p2i_assign, dim3((n1 + 255) / 256), dim3(256), 0, s, (const xc_point *)at(i_pts),
                       (const double *)at(i_vals), n1, N, (const int *)at(i_own), (double *)at(i_i1));
    XC_HIP(ctx, hipGetLastError());
    const PairDev *dP = (const PairDev *)at(i_pair);
    if ((st = launch_row_r2c<double>(ctx, NH, nr2c, dP, (const int4 *)at(i_wr)))) return st;
    if ((st = launch_col<double>(ctx, N, false, 0, 1, ncol, dP, (const int2 *)at(i_wc)))) return st;
    if ((st = launch_row_c2r<double>(ctx, NH, nc2r, dP, (const int2 *)at(i_wc2)))) return st;
    if ((st = launch_finalize<double>(ctx, 1, dP))) return st;
    XC_HIP(ctx, hipMemcpyAsync(hres, at(i_res), sizeof(xc_result), hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    xc_result R = *hres;
    if (R.status == XC_OK) {
        if (dx) *dx = R.dx;
        if (dy) *dy = R.dy;
        if (r) *r = R.r;
    }
    if (res) *res = R;
    return R.status;
}

// ---------------------------------------------------------------------------
// FourierMatch (Metrics.cpp:517-658): the smallest-footprint rotation and the
// image box on the host (TightestBBox's 90 outline rotations), the splats,
// the three transforms and the spectrum reductions on the device.
// ---------------------------------------------------------------------------

namespace xc {

constexpr int FM_NT = 256;
constexpr int FM_PARTS = 256;   // partial sums per reduction
constexpr int FM_ROWS = 32;     // spectrum rows per workgroup

__device__ __forceinline__ void fm_block_sum2(double &a, double &b, double *sa, double *sb)
{
    const int t = threadIdx.x;
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int off = FM_NT / 2; off > 0; off >>= 1) {
        if (t < off) {
            sa[t] += sa[t + off];
            sb[t] += sb[t + off];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

// MeanSqrDiff's sum, partial per workgroup.
__global__ void __launch_bounds__(FM_NT) k_fm_msd(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                  double *__restrict__ part)
{
    __shared__ double sa[FM_NT], sb[FM_NT];
    double s = 0.0, z = 0.0;
    for (int i = blockIdx.x * FM_NT + threadIdx.x; i < n; i += FM_PARTS * FM_NT) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    fm_block_sum2(s, z, sa, sb);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = s, part[2 * blockIdx.x + 1] = 0.0;
}

__global__ void k_fm_diff(const double *__restrict__ i1, const double *__restrict__ i2, long long n,
                          double *__restrict__ d)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        d[i] = i1[i] - i2[i];
}

// FFT_r2c_lookup (Metrics.cpp:420-435): negative kx as the conjugate at (Nx - kx, ky).
__device__ __forceinline__ cpx<double> fm_lookup(const cpx<double> *__restrict__ c, int Nx, int Ny, int x, int y)
{
    const int M = Nx / 2 + 1;
    if (x < 0) x += Nx;
    if (y < 0) y += Ny;
    if (x > Nx / 2) {
        cpx<double> v = c[(Nx - x) + (long long)M * y];
        v.y = -v.y;
        return v;
    }
    return c[x + (long long)M * y];
}

// The low-frequency (dot, total) over |kx| <= xlim, |ky| <= ylim, partial per workgroup.
__global__ void __launch_bounds__(FM_NT)
    k_fm_lowfreq(const cpx<double> *__restrict__ f1, const cpx<double> *__restrict__ f2, int Nx, int Ny, int xlim,
                 int ylim, double *__restrict__ part)
{
    __shared__ double sa[FM_NT], sb[FM_NT];
    const int ny = 2 * ylim + 1;
    const long long cnt = (long long)(2 * xlim + 1) * ny;
    double dot = 0.0, tot = 0.0;
    for (long long k = (long long)blockIdx.x * FM_NT + threadIdx.x; k < cnt; k += (long long)FM_PARTS * FM_NT) {
        const int x = (int)(k / ny) - xlim, y = (int)(k % ny) - ylim;
        const cpx<double> v1 = fm_lookup(f1, Nx, Ny, x, y), v2 = fm_lookup(f2, Nx, Ny, x, y);
        tot += sqrt((v1.x * v1.x + v1.y * v1.y) * (v2.x * v2.x + v2.y * v2.y));
        dot += v1.x * v2.x + v1.y * v2.y;
    }
    fm_block_sum2(dot, tot, sa, sb);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = dot, part[2 * blockIdx.x + 1] = tot;
}

__global__ void __launch_bounds__(FM_NT) k_fm_sum_parts(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double sa[FM_NT], sb[FM_NT];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += FM_NT) a += part[2 * i], b += part[2 * i + 1];
    fm_block_sum2(a, b, sa, sb);
    if (threadIdx.x == 0) out[0] = a, out[1] = b;
}

// The difference power by integer wavelength (Metrics.cpp:603-628): bins in
// LDS per workgroup of FM_ROWS rows (ky = 1..Ny-1), then added to ps.
__global__ void __launch_bounds__(FM_NT)
    k_fm_spectrum(const cpx<double> *__restrict__ fd, int Nx, int Ny, int maxf, double *__restrict__ ps)
{
    extern __shared__ double hist[];
    const int M = Nx / 2 + 1;
    for (int i = threadIdx.x; i < maxf; i += FM_NT) hist[i] = 0.0;
    __syncthreads();
    const int y0 = 1 + blockIdx.x * FM_ROWS, y1 = min(Ny, y0 + FM_ROWS);
    for (int y = y0; y < y1; ++y) {
        const double wavey = (double)Ny / (y > Ny / 2 ? Ny - y : y);
        for (int x = 1 + threadIdx.x; x < M; x += FM_NT) {
            const double wavex = (double)Nx / x;
            int iwave = (int)sqrt(wavex * wavex + wavey * wavey);
            if (iwave > maxf - 1) iwave = maxf - 1;
            const cpx<double> v = fd[x + (long long)M * y];
            atomicAdd(&hist[iwave], v.x * v.x + v.y * v.y);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < maxf; i += FM_NT)
        if (hist[i] != 0.0) atomicAdd(&ps[i], hist[i]);
}

namespace {

constexpr double FM_PI = 3.14159265358;  // the reference's truncated pi (GenDefs.h)

struct DBoxH {
    double L, R, B, T;
};

// BBoxFromPoints, DBox form (Geometry.cpp:109-137).
void dbox_from_points(DBoxH &B, const xc_point *p, int n)
{
    B.L = B.R = p[0].x;
    B.B = B.T = p[0].y;
    for (int i = 1; i < n; ++i) {
        double t;
        if ((t = p[i].x) < B.L) B.L = t;
        else if (t > B.R) B.R = t;
        if ((t = p[i].y) < B.B) B.B = t;
        else if (t > B.T) B.T = t;
    }
    B.L = floor(B.L);
    B.R = ceil(B.R);
    B.B = floor(B.B);
    B.T = ceil(B.T);
}

#pragma clang fp contract(off)
// NUSetRot then Apply_R_Part (TAffine.cpp:116-122, 407-414).
void rotate_points(std::vector<xc_point> &q, const xc_point *p, int n, double rad)
{
    const double c = cos(rad), s = sin(rad), t1 = -s;
    q.resize(n);
    for (int i = 0; i < n; ++i) {
        const double x = p[i].x * c + p[i].y * t1;
        const double y = p[i].x * s + p[i].y * c;
        q[i].x = x;
        q[i].y = y;
    }
}

// TightestBBox (Geometry.cpp:187-267): the region's row outline (or the points
// when that is no smaller) rotated by whole degrees -45..45; the strictly
// least box area wins, 0 degrees first.
int tightest_bbox_deg(const xc_point *pts, int np)
{
    DBoxH B;
    dbox_from_points(B, pts, np);
    const int ny = (int)(B.T - B.B) + 1;
    std::vector<double> minx(ny, B.R + 1), maxx(ny, B.L - 1);
    for (int i = 0; i < np; ++i) {
        const int iy = (int)floor(pts[i].y - B.B);
        minx[iy] = fmin(minx[iy], pts[i].x);
        maxx[iy] = fmax(maxx[iy], pts[i].x);
    }
    std::vector<xc_point> outline;
    for (int iy = 0; iy < ny; ++iy)
        if (minx[iy] <= maxx[iy]) {
            outline.push_back({minx[iy], iy + B.B});
            outline.push_back({maxx[iy], iy + B.B});
        }
    const xc_point *set = outline.data();
    int ns = (int)outline.size();
    if (ns >= np) {
        set = pts;
        ns = np;
    }
    double best_area = (B.R - B.L) * (B.T - B.B);
    int best = 0;
    std::vector<xc_point> P;
    for (int angle = -45; angle <= 45; ++angle) {
        if (!angle) continue;
        rotate_points(P, set, ns, angle * FM_PI / 180);
        DBoxH box;
        dbox_from_points(box, P.data(), ns);
        const double area = (box.R - box.L) * (box.T - box.B);
        if (area < best_area) {
            best_area = area;
            best = angle;
        }
    }
    return best;
}
#pragma clang fp contract(on)

}  // namespace
}  // namespace xc

extern "C" int xc_fourier_match_f64(xc_ctx *ctx, const xc_point *pts, const double *av, const double *bv, int n,
                                    int wvlen, double *dot, xc_fourier_qa *qa)
{
    if (!ctx || !dot) return XC_E_ARG;
    if (n < 0 || (n > 0 && (!pts || !av || !bv))) return set_err(ctx, XC_E_ARG, "FourierMatch: null or negative input");
    if (wvlen < 2) return set_err(ctx, XC_E_ARG, "FourierMatch: wvlen %d < 2", wvlen);
    xc_fourier_qa q;
    memset(&q, 0, sizeof q);
    *dot = 0.0;
    if (!n) {  // MeanSqrDiff's "No data?"
        if (qa) *qa = q;
        return XC_OK;
    }
    // SmallestFootprint and MakeMetricImagesFFT's box (host geometry)
    q.deg = tightest_bbox_deg(pts, n);
    std::vector<xc_point> rot;
    const xc_point *P = pts;
    if (q.deg) {
        rotate_points(rot, pts, n, q.deg * FM_PI / 180);
        P = rot.data();
    }
    IBox B;
    bbox_from_points(B, P, n);
    const int Nx = ceil_pow2(B.R - B.L + 1), Ny = ceil_pow2(B.T - B.B + 1);
    if (Nx < 2 || Nx / 2 > max_nh<double>() || Ny > max_ny<double>())
        return set_err(ctx, XC_E_SHAPE, "FourierMatch: metric image %d x %d (need 2 <= Nx <= %d, Ny <= %d)", Nx, Ny,
                       2 * max_nh<double>(), max_ny<double>());
    const int NH = Nx / 2, H = NH + 1;
    const long long N2 = (long long)Nx * Ny, MS = (long long)H * Ny;
    const int maxf = (int)(sqrt((double)(Nx * Nx + Ny * Ny)) / 2.0) + 1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    // the three lists staged in pinned memory (with the on-grid test of the splat), then uploaded
    const size_t sp = (sizeof(xc_point) * n + 255) / 256 * 256, sa = (sizeof(double) * n + 255) / 256 * 256;
    int st = ensure_pinned(ctx, sp + 2 * sa);
    if (st) return st;
    xc_point *hp = (xc_point *)ctx->pinned;
    double *ha = (double *)((char *)ctx->pinned + sp), *hb = (double *)((char *)ctx->pinned + sp + sa);
    const bool og = stage_points(hp, P, ha, av, hb, bv, n);
    Layout lay;
    const size_t i_p = lay.add(sizeof(xc_point) * n), i_a = lay.add(sizeof(double) * n), i_b = lay.add(sizeof(double) * n);
    const size_t i_i1 = lay.add(sizeof(double) * N2), i_i2 = lay.add(sizeof(double) * N2), i_d = lay.add(sizeof(double) * N2);
    const size_t i_f1 = lay.add(sizeof(cpx<double>) * MS), i_f2 = lay.add(sizeof(cpx<double>) * MS);
    const size_t i_fd = lay.add(sizeof(cpx<double>) * MS), i_part = lay.add(sizeof(double) * 4 * FM_PARTS);
    const size_t i_ps = lay.add(sizeof(double) * maxf), i_out = lay.add(sizeof(double) * 4);
    const size_t spl = og ? splat_scratch_bytes(n, N2) : 0;
    const size_t i_spl = lay.add(spl ? spl : 1);
    st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };
    double *i1 = (double *)at(i_i1), *i2 = (double *)at(i_i2), *df = (double *)at(i_d);
    cpx<double> *f1 = (cpx<double> *)at(i_f1), *f2 = (cpx<double> *)at(i_f2), *fd = (cpx<double> *)at(i_fd);
    double *part = (double *)at(i_part), *ps = (double *)at(i_ps), *out = (double *)at(i_out);
    hipStream_t s = ctx->stream;
    XC_HIP(ctx, hipMemcpyAsync(at(i_p), hp, sizeof(xc_point) * n, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_a), ha, sizeof(double) * n, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_b), hb, sizeof(double) * n, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemsetAsync(i1, 0, sizeof(double) * N2, s));
    XC_HIP(ctx, hipMemsetAsync(i2, 0, sizeof(double) * N2, s));
    XC_HIP(ctx, hipMemsetAsync(ps, 0, sizeof(double) * maxf, s));
    k_fm_msd<<<FM_PARTS, FM_NT, 0, s>>>((const double *)at(i_a), (const double *)at(i_b), n, part);
    k_fm_sum_parts<<<1, FM_NT, 0, s>>>(part, FM_PARTS, out + 2);
    if ((st = splat_points(ctx, s, (const xc_point *)at(i_p), (const double *)at(i_a), n, (double)B.L, (double)B.B, i1,
                           Nx, Ny, og, at(i_spl), spl)))
        return st;
    if ((st = splat_points(ctx, s, (const xc_point *)at(i_p), (const double *)at(i_b), n, (double)B.L, (double)B.B, i2,
                           Nx, Ny, og, at(i_spl), spl)))
        return st;
    k_fm_diff<<<(int)std::min<long long>((N2 + 255) / 256, 8192), 256, 0, s>>>(i1, i2, N2, df);
    XC_HIP(ctx, hipGetLastError());
    if ((st = seam_r2c(ctx, NH, i1, Ny, f1))) return st;
    if ((st = seam_col<-1>(ctx, Ny, f1, H))) return st;
    if ((st = seam_r2c(ctx, NH, i2, Ny, f2))) return st;
    if ((st = seam_col<-1>(ctx, Ny, f2, H))) return st;
    if ((st = seam_r2c(ctx, NH, df, Ny, fd))) return st;
    if ((st = seam_col<-1>(ctx, Ny, fd, H))) return st;
    k_fm_lowfreq<<<FM_PARTS, FM_NT, 0, s>>>(f1, f2, Nx, Ny, Nx / wvlen, Ny / wvlen, part);
    k_fm_sum_parts<<<1, FM_NT, 0, s>>>(part, FM_PARTS, out);
    if (Ny > 1)
        k_fm_spectrum<<<(Ny - 1 + FM_ROWS - 1) / FM_ROWS, FM_NT, sizeof(double) * maxf, s>>>(fd, Nx, Ny, maxf, ps);
    XC_HIP(ctx, hipGetLastError());
    std::vector<double> hps(maxf);
    double ho[4];
    XC_HIP(ctx, hipMemcpyAsync(ho, out, sizeof ho, hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipMemcpyAsync(hps.data(), ps, sizeof(double) * maxf, hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    // the half-power bin (Metrics.cpp:634-650)
    double tot = 0.0, cum = 0.0;
    int i;
    for (i = 0; i < maxf; ++i) tot += hps[i];
    for (i = 0; cum < tot * 0.50 && i < maxf; ++i) cum += hps[i];
    q.dot = ho[0] / ho[1];
    q.energy = ho[1];
    q.msd = ho[2] / n;
    q.tot_power = tot;
    q.half_frac = cum / tot;
    q.half_index = i;
    q.maxf = maxf;
    q.nx = Nx;
    q.ny = Ny;
    *dot = q.dot;
    if (qa) *qa = q;
    return XC_OK;
}
