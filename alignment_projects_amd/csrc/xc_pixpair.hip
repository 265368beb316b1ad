// xc_pixpair.hip -- PixPair::Load's raster preparation on the device
// (CPixPair.cpp:276-460): Legendre flattening (LegPolyFlatten for 8-bit
// rasters, Maths.cpp:1006-1086), the DoG filter (MakeDoGKernel,
// CPixPair.cpp:36-80, applied as Convolve does, Correlation.cpp:371-457),
// Normalize (Maths.cpp:192-212) and the <= 2048-pixel box downsample
// (PixPair::Downsample, CPixPair.cpp:90-113).
//
// Flattening.  The reference removes each Legendre product F_k = X_a(x) Y_b(y)
// in turn, projecting the current residual: c_k = <F_k, v_k> / <F_k, F_k>,
// v_{k+1} = v_k - c_k F_k.  Since v_k = v_0 - sum_{j<k} c_j F_j,
// <F_k, v_k> = M_k - sum_{j<k} c_j <F_k, F_j>, with the moments
// M_k = <F_k, v_0> and the Gram entries <F_k, F_j> = <X_a, X_c> <Y_b, Y_d>
// separable.  So one pass over the raster gathers every moment (per-row
// partials r_a(y) = sum_x X_a(x) v_0(x, y), then M_ab = sum_y Y_b(y) r_a(y)),
// the host solves the short triangular recurrence, and two more passes
// (statistics, then the normalised write) recompute the residual from the
// 8-bit raster: 1 + 1 + 9 bytes per pixel in all.
//
// DoG.  Convolve pre-normalises the kernel by its sum, and a difference of
// two unit-sum Gaussians sums to rounding noise (about 1e-17 .. 1e-15, either
// sign): the filtered picture's scale and sign are set by those bits.  The
// kernel and its sum are therefore built on the host in the reference's
// order with the C library's exp, and the device applies the exact table
// (divided by the sum) as a direct sum with Convolve's circular geometry:
// out(x, y) = sum_{i,j} K'(i, j) S((x - j) mod Nx, (y - i) mod Ny) over the
// FFTSize(ws, wk/2) x FFTSize(hs, hk/2) grid, S zero outside the picture.
// LDS tiles of 32 x 32 outputs; each thread keeps 4 outputs of a column and
// slides a 4-value window down the tile (one LDS read per 4 FMAs); the
// kernel taps are wave-uniform scalar loads.
#include <math.h>
#include <string.h>

#include <vector>

#include "xc_host.hpp"

namespace xc {

constexpr int PP_NT = 256;
constexpr int PP_MAXORD = 6;
constexpr int PP_MAXR2 = 18;  // D = 6 r2 + 1 <= 109: the 32 x 32 tile and its apron fit 160 KB of LDS
constexpr int PP_TILE = 32;
constexpr int PP_RPT = 4;     // outputs per thread along y

// Deterministic LDS tree sum of K values over a 256-thread workgroup; lane 0 gets the totals.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double (*sh)[PP_NT])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int off = PP_NT / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][t] += sh[k][t + off];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = sh[k][0];
}

// Per-row moments r_a(y) = sum_x X_a(x) v(x, y), a = 0..ord.
__global__ void __launch_bounds__(PP_NT)
    k_pp_rowmom(const uint8_t *__restrict__ ras, int w, int ord, const double *__restrict__ xp,
                double *__restrict__ rowmom)
{
    __shared__ double sh[PP_MAXORD + 1][PP_NT];
    const int y = blockIdx.x;
    const uint8_t *row = ras + (long long)y * w;
    double acc[PP_MAXORD + 1];
#pragma unroll
    for (int a = 0; a <= PP_MAXORD; ++a) acc[a] = 0.0;
    for (int x = threadIdx.x; x < w; x += PP_NT) {
        const double v = row[x];
#pragma unroll
        for (int a = 0; a <= PP_MAXORD; ++a)
            if (a <= ord) acc[a] += xp[(long long)a * w + x] * v;
    }
    block_sums<PP_MAXORD + 1>(acc, sh);
    if (threadIdx.x == 0)
        for (int a = 0; a <= ord; ++a) rowmom[(long long)y * (ord + 1) + a] = acc[a];
}

// M[a * (ord+1) + b] = sum_y Y_b(y) r_a(y): one workgroup per (a, b).
__global__ void __launch_bounds__(PP_NT)
    k_pp_colmom(const double *__restrict__ rowmom, int h, int ord, const double *__restrict__ yp,
                double *__restrict__ mom)
{
    __shared__ double sh[1][PP_NT];
    const int a = blockIdx.x / (ord + 1), b = blockIdx.x % (ord + 1);
    double acc[1] = {0.0};
    for (int y = threadIdx.x; y < h; y += PP_NT) acc[0] += yp[(long long)b * h + y] * rowmom[(long long)y * (ord + 1) + a];
    block_sums<1>(acc, sh);
    if (threadIdx.x == 0) mom[blockIdx.x] = acc[0];
}

// The flattened residual at (x, y): v - sum_a q_a(y) X_a(x), q_a(y) = sum_b C_ab Y_b(y).
__device__ __forceinline__ void pp_row_poly(const double *__restrict__ C, const double *__restrict__ yp, int h, int ord,
                                            int y, double (&q)[PP_MAXORD + 1])
{
#pragma unroll
    for (int a = 0; a <= PP_MAXORD; ++a) {
        double s = 0.0;
        if (a <= ord)
            for (int b = 0; b <= ord; ++b) s += C[a * (ord + 1) + b] * yp[(long long)b * h + y];
        q[a] = s;
    }
}

__device__ __forceinline__ double pp_resid(double v, const double *__restrict__ xp, int w, int ord, int x,
                                           const double (&q)[PP_MAXORD + 1])
{
#pragma unroll
    for (int a = 0; a <= PP_MAXORD; ++a)
        if (a <= ord) v -= q[a] * xp[(long long)a * w + x];
    return v;
}

// Per-row sum and sum of squares of the residual.
__global__ void __launch_bounds__(PP_NT)
    k_pp_flat_stats(const uint8_t *__restrict__ ras, int w, int h, int ord, const double *__restrict__ xp,
                    const double *__restrict__ yp, const double *__restrict__ C, double *__restrict__ rowstat)
{
    __shared__ double sh[2][PP_NT];
    const int y = blockIdx.x;
    double q[PP_MAXORD + 1];
    pp_row_poly(C, yp, h, ord, y, q);
    const uint8_t *row = ras + (long long)y * w;
    double acc[2] = {0.0, 0.0};
    for (int x = threadIdx.x; x < w; x += PP_NT) {
        const double r = pp_resid((double)row[x], xp, w, ord, x, q);
        acc[0] += r;
        acc[1] += r * r;
    }
    block_sums<2>(acc, sh);
    if (threadIdx.x == 0) {
        rowstat[2 * y] = acc[0];
        rowstat[2 * y + 1] = acc[1];
    }
}

// Sum n (sum, sum of squares) partials in index order groups; tot = (sum, sum2).
__global__ void __launch_bounds__(PP_NT) k_pp_sum_parts(const double *__restrict__ part, int n, double *__restrict__ tot)
{
    __shared__ double sh[2][PP_NT];
    double acc[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += PP_NT) {
        acc[0] += part[2 * i];
        acc[1] += part[2 * i + 1];
    }
    block_sums<2>(acc, sh);
    if (threadIdx.x == 0) {
        tot[0] = acc[0];
        tot[1] = acc[1];
    }
}

// MeanStd::Stats (Maths.cpp:111-124), the n-1 deviation.
__device__ __forceinline__ void pp_mean_sd(const double *__restrict__ tot, long long n, double &avg, double &sd)
{
    avg = tot[0];
    sd = 0.0;
    if (n > 1) {
        avg /= (double)n;
        sd = sqrt(fmax(tot[1] - (double)n * avg * avg, 0.0) / ((double)n - 1.0));
    }
}

__global__ void __launch_bounds__(PP_NT)
    k_pp_flat_write(const uint8_t *__restrict__ ras, int w, int h, int ord, const double *__restrict__ xp,
                    const double *__restrict__ yp, const double *__restrict__ C, const double *__restrict__ tot,
                    double *__restrict__ out)
{
    const int y = blockIdx.x;
    double q[PP_MAXORD + 1], avg, sd;
    pp_row_poly(C, yp, h, ord, y, q);
    pp_mean_sd(tot, (long long)w * h, avg, sd);
    const uint8_t *row = ras + (long long)y * w;
    double *orow = out + (long long)y * w;
    for (int x = threadIdx.x; x < w; x += PP_NT) orow[x] = (pp_resid((double)row[x], xp, w, ord, x, q) - avg) / sd;
}

// Convolve's circular convolution by the D x D table K' (already divided by its
// sum; KnT[kx * D + ky] = K'[ky * D + kx]), 32 x 32 outputs per workgroup of 256
// threads (32 columns x 8 row groups of 4); per-workgroup (sum, sum2) of the
// outputs into part.
__global__ void __launch_bounds__(PP_NT)
    k_pp_conv(const double *__restrict__ src, int ws, int hs, int Nx, int Ny, const double *__restrict__ KnT, int D,
              double *__restrict__ dst, double *__restrict__ part)
{
    extern __shared__ double T[];  // (32 + D - 1)^2
    const int L = PP_TILE + D - 1, r = D / 2;
    const int x0 = blockIdx.x * PP_TILE, y0 = blockIdx.y * PP_TILE;
    for (int i = threadIdx.x; i < L * L; i += PP_NT) {
        const int ly = i / L, lx = i - ly * L;
        int sx = (x0 - r + lx) % Nx, sy = (y0 - r + ly) % Ny;
        sx += sx < 0 ? Nx : 0;
        sy += sy < 0 ? Ny : 0;
        T[i] = (sx < ws && sy < hs) ? src[(long long)sy * ws + sx] : 0.0;
    }
    __syncthreads();
    const int tx = threadIdx.x % PP_TILE, tq = threadIdx.x / PP_TILE;  // column, row group
    const int base = tq * PP_RPT;
    double acc[PP_RPT] = {0.0, 0.0, 0.0, 0.0};
    for (int kx = 0; kx < D; ++kx) {
        // out(x0+tx, y0+base+q) = sum_{kx,ky} K'[ky*D+kx] T[base+q+D-1-ky][tx+D-1-kx]
        const double *col = T + (tx + D - 1 - kx);
        double w0 = col[(base + D - 1) * L], w1 = col[(base + D) * L], w2 = col[(base + D + 1) * L],
               w3 = col[(base + D + 2) * L];
        const double *kc = KnT + (long long)kx * D;  // taps of kernel column kx, by row
        for (int ky = 0; ky < D; ++ky) {
            const double k = kc[ky];
            acc[0] += k * w0;
            acc[1] += k * w1;
            acc[2] += k * w2;
            acc[3] += k * w3;
            w3 = w2;
            w2 = w1;
            w1 = w0;
            w0 = col[max(base + D - 2 - ky, 0) * L];  // the last (clamped) read goes unused
        }
    }
    double st[2] = {0.0, 0.0};
    const int x = x0 + tx;
#pragma unroll
    for (int q = 0; q < PP_RPT; ++q) {
        const int y = y0 + base + q;
        if (x < ws && y < hs) {
            dst[(long long)y * ws + x] = acc[q];
            st[0] += acc[q];
            st[1] += acc[q] * acc[q];
        }
    }
    __syncthreads();  // T is reused as the reduction buffer
    double(*sh)[PP_NT] = reinterpret_cast<double(*)[PP_NT]>(T);
    block_sums<2>(st, sh);
    if (threadIdx.x == 0) {
        const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = st[0];
        part[2 * b + 1] = st[1];
    }
}

__global__ void k_pp_normalize(double *__restrict__ v, long long n, const double *__restrict__ tot)
{
    double avg, sd;
    pp_mean_sd(tot, n, avg, sd);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        v[i] = (v[i] - avg) / sd;
}

// PixPair::Downsample: scl x scl box means, summed in (dy, dx) order.
__global__ void k_pp_downsample(const double *__restrict__ src, int wf, int scl, int ws, int hs, double *__restrict__ dst)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ws * hs) return;
    const int iy = (int)(i / ws), ix = (int)(i - (long long)iy * ws);
    double sum = 0.0;
    for (int dy = 0; dy < scl; ++dy)
        for (int dx = 0; dx < scl; ++dx) sum += src[ix * scl + dx + (long long)wf * (iy * scl + dy)];
    dst[i] = sum / (scl * scl);
}

namespace {

#pragma clang fp contract(off)

// LegPolyCreate (Maths.cpp:962-997): L[o * n + s], o = 0..ord.
void leg_poly_create(std::vector<double> &L, int ord, int n)
{
    L.assign((size_t)(ord + 1) * n, 0.0);
    const double dn = n, n_1 = dn - 1;
    for (int s = 0; s < n; ++s) {
        const double x = (2.0 * s - n_1) / dn;
        const double v[7] = {1.0,
                             x,
                             1.5 * x * x - 0.5,
                             2.5 * x * x * x - 1.5 * x,
                             0.125 * (35.0 * x * x * x * x - 30.0 * x * x + 3.0),
                             0.125 * (63.0 * x * x * x * x * x - 70.0 * x * x * x + 15.0 * x),
                             0.0625 * (231.0 * x * x * x * x * x * x - 315.0 * x * x * x * x + 105.0 * x * x - 5.0)};
        for (int o = 0; o <= ord; ++o) L[(size_t)o * n + s] = v[o];
    }
}

// MakeDoGKernel (CPixPair.cpp:36-80), then Convolve's pre-normalisation by the
// kernel's sum (Correlation.cpp:394-405), in the reference's order.
int dog_table(std::vector<double> &K, int r1, int r2)
{
    const int d = 3 * r2, D = 2 * d + 1;
    double sum1 = 0.0, sum2 = 0.0;
    for (int y = -d; y <= d; ++y)
        for (int x = -d; x <= d; ++x) {
            const double R = x * x + y * y;
            sum1 += exp(-(R / (2 * r1 * r1)));
            sum2 += exp(-(R / (2 * r2 * r2)));
        }
    K.assign((size_t)D * D, 0.0);
    for (int y = -d; y <= d; ++y)
        for (int x = -d; x <= d; ++x) {
            const double R = x * x + y * y;
            const double v1 = exp(-(R / (2 * r1 * r1)));
            const double v2 = exp(-(R / (2 * r2 * r2)));
            K[x + d + D * (y + d)] = v1 / sum1 - v2 / sum2;
        }
    double nrm = 0.0;
    for (int i = 0; i < D * D; ++i) nrm += K[i];
    if (!nrm) nrm = 1.0;
    for (int i = 0; i < D * D; ++i) K[i] /= nrm;
    return D;
}

// The sequential projection coefficients C[a][b] from the moments M[a][b].
void flatten_coefs(const std::vector<double> &xp, int w, const std::vector<double> &yp, int h, int ord,
                   const double *M, double *C)
{
    const int n1 = ord + 1;
    std::vector<double> gx((size_t)n1 * n1, 0.0), gy((size_t)n1 * n1, 0.0);
    for (int a = 0; a < n1; ++a)
        for (int c = 0; c < n1; ++c) {
            double s = 0.0;
            for (int x = 0; x < w; ++x) s += xp[(size_t)a * w + x] * xp[(size_t)c * w + x];
            gx[a * n1 + c] = s;
            double t = 0.0;
            for (int y = 0; y < h; ++y) t += yp[(size_t)a * h + y] * yp[(size_t)c * h + y];
            gy[a * n1 + c] = t;
        }
    std::vector<int> ka, kb;  // the reference's term order: xo outer, yo inner, (0,0) skipped
    for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n1; ++b)
            if (a + b) {
                ka.push_back(a);
                kb.push_back(b);
            }
    std::vector<double> c(ka.size());
    for (int i = 0; i < n1 * n1; ++i) C[i] = 0.0;
    for (size_t k = 0; k < ka.size(); ++k) {
        double p = M[ka[k] * n1 + kb[k]];
        for (size_t j = 0; j < k; ++j) p -= c[j] * gx[ka[k] * n1 + ka[j]] * gy[kb[k] * n1 + kb[j]];
        c[k] = p / (gx[ka[k] * n1 + ka[k]] * gy[kb[k] * n1 + kb[k]]);
        C[ka[k] * n1 + kb[k]] = c[k];
    }
}

int pp_run(xc_ctx *ctx, const uint8_t *d_ras, int w, int h, int order, int dog, int r1, int r2, double *d_vf,
           double *d_vfflt, double *d_vs, double *d_vsflt, int ws, int hs, int scl)
{
    hipStream_t s = ctx->stream;
    const int n1 = order + 1;
    std::vector<double> xp, yp, K;
    leg_poly_create(xp, order, w);
    leg_poly_create(yp, order, h);
    const int D = dog ? dog_table(K, r1, r2) : 0;
    std::vector<double> KT(K.size());  // by kernel column, for contiguous tap loads
    for (int y = 0; y < D; ++y)
        for (int x = 0; x < D; ++x) KT[(size_t)x * D + y] = K[(size_t)y * D + x];
    const int Nx = dog ? ceil_pow2(w + D / 2 - 1) : 0, Ny = dog ? ceil_pow2(h + D / 2 - 1) : 0;
    const int gx = (w + PP_TILE - 1) / PP_TILE, gy = (h + PP_TILE - 1) / PP_TILE;
    const long long nconv = dog ? (long long)gx * gy : 0;

    Layout lay;
    const size_t i_xp = lay.add(sizeof(double) * xp.size()), i_yp = lay.add(sizeof(double) * yp.size());
    const size_t i_rm = lay.add(sizeof(double) * (size_t)h * n1), i_mom = lay.add(sizeof(double) * n1 * n1);
    const size_t i_c = lay.add(sizeof(double) * n1 * n1), i_rs = lay.add(sizeof(double) * 2 * (size_t)h);
    const size_t i_tot = lay.add(sizeof(double) * 4), i_k = lay.add(sizeof(double) * (K.size() + 1));
    const size_t i_part = lay.add(sizeof(double) * 2 * (size_t)(nconv + 1));
    int st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    auto at = [&](size_t i) { return (double *)(base + lay.off[i]); };
    st = ensure_pinned(ctx, sizeof(double) * (2 * n1 * n1 + 8));
    if (st) return st;
    double *hm = (double *)ctx->pinned;

    XC_HIP(ctx, hipMemcpyAsync(at(i_xp), xp.data(), sizeof(double) * xp.size(), hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_yp), yp.data(), sizeof(double) * yp.size(), hipMemcpyHostToDevice, s));
    if (dog) XC_HIP(ctx, hipMemcpyAsync(at(i_k), KT.data(), sizeof(double) * KT.size(), hipMemcpyHostToDevice, s));
    // moments -> host -> coefficients
    k_pp_rowmom<<<h, PP_NT, 0, s>>>(d_ras, w, order, at(i_xp), at(i_rm));
    k_pp_colmom<<<n1 * n1, PP_NT, 0, s>>>(at(i_rm), h, order, at(i_yp), at(i_mom));
    XC_HIP(ctx, hipGetLastError());
    XC_HIP(ctx, hipMemcpyAsync(hm, at(i_mom), sizeof(double) * n1 * n1, hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    flatten_coefs(xp, w, yp, h, order, hm, hm + n1 * n1);
    XC_HIP(ctx, hipMemcpyAsync(at(i_c), hm + n1 * n1, sizeof(double) * n1 * n1, hipMemcpyHostToDevice, s));
    // residual statistics, then the normalised write
    k_pp_flat_stats<<<h, PP_NT, 0, s>>>(d_ras, w, h, order, at(i_xp), at(i_yp), at(i_c), at(i_rs));
    k_pp_sum_parts<<<1, PP_NT, 0, s>>>(at(i_rs), h, at(i_tot));
    k_pp_flat_write<<<h, PP_NT, 0, s>>>(d_ras, w, h, order, at(i_xp), at(i_yp), at(i_c), at(i_tot), d_vf);
    XC_HIP(ctx, hipGetLastError());
    const long long n = (long long)w * h;
    if (dog) {
        const int L = PP_TILE + D - 1;
        const size_t lds = sizeof(double) * (size_t)L * L;
        if (lds > ctx->lds_max) return set_err(ctx, XC_E_SHAPE, "DoG of width %d needs %zu bytes of LDS", D, lds);
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pp_conv),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        k_pp_conv<<<dim3(gx, gy), PP_NT, lds, s>>>(d_vf, w, h, Nx, Ny, at(i_k), D, d_vfflt, at(i_part));
        k_pp_sum_parts<<<1, PP_NT, 0, s>>>(at(i_part), (int)nconv, at(i_tot) + 2);
        const int nb = (int)std::min<long long>((n + PP_NT - 1) / PP_NT, 8192);
        k_pp_normalize<<<nb, PP_NT, 0, s>>>(d_vfflt, n, at(i_tot) + 2);
        XC_HIP(ctx, hipGetLastError());
    }
    if (scl > 1) {
        const long long m = (long long)ws * hs;
        const int nb = (int)((m + PP_NT - 1) / PP_NT);
        if (d_vs) k_pp_downsample<<<nb, PP_NT, 0, s>>>(d_vf, w, scl, ws, hs, d_vs);
        if (dog && d_vsflt) k_pp_downsample<<<nb, PP_NT, 0, s>>>(d_vfflt, w, scl, ws, hs, d_vsflt);
        XC_HIP(ctx, hipGetLastError());
    }
    return XC_OK;
}

// Argument checks and the downsample geometry (CPixPair.cpp:427-441).
int pp_check(xc_ctx *ctx, int w, int h, int order, int dog, int r1, int r2, int &ws, int &hs, int &scl)
{
    if (w < 1 || h < 1) return set_err(ctx, XC_E_ARG, "empty raster");
    if (order < 0 || order > PP_MAXORD) return set_err(ctx, XC_E_ARG, "flatten order %d outside 0..6", order);
    if (dog && (r1 < 1 || r2 < 1 || r2 > PP_MAXR2))
        return set_err(ctx, XC_E_ARG, "DoG radii (%d, %d): need 1 <= r1 and 1 <= r2 <= %d", r1, r2, PP_MAXR2);
    ws = w;
    hs = h;
    scl = 1;
    while (ws > 2048 || hs > 2048) {
        ws /= 2;
        hs /= 2;
        scl *= 2;
    }
    if (ws * scl != w || hs * scl != h)
        return set_err(ctx, XC_E_ARG, "dimensions %d x %d not a multiple of the scale %d", w, h, scl);
    return XC_OK;
}

}  // namespace
}  // namespace xc

using namespace xc;

extern "C" int xc_pixpair_prep_dev_f64(xc_ctx *ctx, const uint8_t *ras, int w, int h, int order, int dog, int r1,
                                       int r2, double *vf, double *vfflt, double *vs, double *vsflt, int *ws_out,
                                       int *hs_out, int *scl_out)
{
    if (!ctx) return XC_E_ARG;
    if (!ras || !vf || (dog && !vfflt)) return set_err(ctx, XC_E_ARG, "null raster or output");
    int ws, hs, scl;
    int st = pp_check(ctx, w, h, order, dog, r1, r2, ws, hs, scl);
    if (st) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    st = pp_run(ctx, ras, w, h, order, dog, r1, r2, vf, vfflt, vs, vsflt, ws, hs, scl);
    if (st) return st;
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ws_out) *ws_out = ws;
    if (hs_out) *hs_out = hs;
    if (scl_out) *scl_out = scl;
    return XC_OK;
}

extern "C" int xc_pixpair_prep_f64(xc_ctx *ctx, const uint8_t *ras, int w, int h, int order, int dog, int r1,
                                   int r2, double *vf, double *vfflt, double *vs, double *vsflt, int *ws_out,
                                   int *hs_out, int *scl_out)
{
    if (!ctx) return XC_E_ARG;
    if (!ras) return set_err(ctx, XC_E_ARG, "null raster");
    int ws, hs, scl;
    int st = pp_check(ctx, w, h, order, dog, r1, r2, ws, hs, scl);
    if (st) return st;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h, m = (size_t)ws * hs;
    const bool down = scl > 1;
    DevBuf buf;  // raster and outputs, released before return
    Layout lay;
    const size_t i_r = lay.add(n), i_f = lay.add(sizeof(double) * n), i_g = lay.add(dog ? sizeof(double) * n : 1);
    const size_t i_s = lay.add(down ? sizeof(double) * m : 1), i_sg = lay.add(down && dog ? sizeof(double) * m : 1);
    XC_HIP(ctx, hipMalloc(&buf.p, lay.total));
    char *base = (char *)buf.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };
    auto done = [&](int code) {
        (void)hipFree(buf.p);
        return code;
    };
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(at(i_r), ras, n, hipMemcpyHostToDevice, s) != hipSuccess)
        return done(set_err(ctx, XC_E_DEVICE, "raster upload failed"));
    st = pp_run(ctx, (const uint8_t *)at(i_r), w, h, order, dog, r1, r2, (double *)at(i_f),
                dog ? (double *)at(i_g) : nullptr, down ? (double *)at(i_s) : nullptr,
                down && dog ? (double *)at(i_sg) : nullptr, ws, hs, scl);
    if (st) return done(st);
    bool ok = true;
    if (vf) ok &= hipMemcpyAsync(vf, at(i_f), sizeof(double) * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (vfflt && dog) ok &= hipMemcpyAsync(vfflt, at(i_g), sizeof(double) * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (vs && down) ok &= hipMemcpyAsync(vs, at(i_s), sizeof(double) * m, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (vsflt && down && dog)
        ok &= hipMemcpyAsync(vsflt, at(i_sg), sizeof(double) * m, hipMemcpyDeviceToHost, s) == hipSuccess;
    ok &= hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return done(set_err(ctx, XC_E_DEVICE, "result download failed"));
    if (ws_out) *ws_out = ws;
    if (hs_out) *hs_out = hs;
    if (scl_out) *scl_out = scl;
    return done(XC_OK);
}

extern "C" int xc_leg_poly_flatten_f64(xc_ctx *ctx, const uint8_t *src, int w, int h, int order, double *vals)
{
    if (!ctx) return XC_E_ARG;
    if (!src || !vals || w < 1 || h < 1) return set_err(ctx, XC_E_ARG, "LegPolyFlatten: null or empty raster");
    if (order < 0 || order > PP_MAXORD) return set_err(ctx, XC_E_ARG, "flatten order %d outside 0..6", order);
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)w * h;
    DevBuf buf;
    Layout lay;
    const size_t i_r = lay.add(n), i_f = lay.add(sizeof(double) * n);
    XC_HIP(ctx, hipMalloc(&buf.p, lay.total));
    char *base = (char *)buf.p;
    auto done = [&](int code) {
        (void)hipFree(buf.p);
        return code;
    };
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(base + lay.off[i_r], src, n, hipMemcpyHostToDevice, s) != hipSuccess)
        return done(set_err(ctx, XC_E_DEVICE, "raster upload failed"));
    int st = pp_run(ctx, (const uint8_t *)(base + lay.off[i_r]), w, h, order, 0, 0, 0, (double *)(base + lay.off[i_f]),
                    nullptr, nullptr, nullptr, w, h, 1);
    if (st) return done(st);
    if (hipMemcpyAsync(vals, base + lay.off[i_f], sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(set_err(ctx, XC_E_DEVICE, "result download failed"));
    return done(XC_OK);
}
