/*
 * xc_oracle.c -- CPU oracle for the FFT tile-pair cross-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C and FP64, the
 * algorithm of billkarsh/Alignment_Projects' correlation path so that the
 * HIP library can be checked against it.  It is loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker or as the timed CPU baseline, never as the thing measured or
 * shipped.  libxcorr.so does not link it.
 *
 * Every function cites the reference file:line it follows.  Compiled with
 * -ffp-contract=off so that products and sums round exactly as the
 * reference's x86-64 -O3 build (no FMA) rounds them.
 *
 * FFT: the reference calls FFTW3 (or MKL DFTI) -- third-party, absent from
 * the reference tree.  The oracle carries its own radix-2 FP64 complex FFT
 * (textbook Cooley-Tukey, exact-angle twiddles), pinned in tests against a
 * direct DFT and numpy.fft.  For the CPU baseline, the FFTW3 API can be
 * resolved at run time from an installed library exporting it (MKL's FFTW3
 * wrappers) and called in the reference's own pattern: a plan per call with
 * FFTW_ESTIMATE, executed under one mutex (Correlation_fft_fftw.cpp:10-92).
 */
#include "../include/xcorr.h"

#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cd;

/* GenDefs.h:16 -- the reference's truncated pi, used for deg->rad. */
#define REF_PI 3.14159265358

/* ------------------------------------------------------------------ */
/* Complex FFT (oracle's own)                                          */
/* ------------------------------------------------------------------ */

static pthread_mutex_t tw_mutex = PTHREAD_MUTEX_INITIALIZER;
static cd *tw_cache[32];

/* e^{-2 pi i k / n}, k < n/2, from long-double exact-angle cos/sin. */
static const cd *twiddles(int n)
{
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    pthread_mutex_lock(&tw_mutex);
    if (!tw_cache[lg]) {
        int h = n / 2 > 0 ? n / 2 : 1;
        cd *t = (cd *)malloc(sizeof(cd) * h);
        for (int k = 0; k < h; ++k) {
            long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
            t[k].re = (double)cosl(a);
            t[k].im = (double)-sinl(a);
        }
        tw_cache[lg] = t;
    }
    const cd *r = tw_cache[lg];
    pthread_mutex_unlock(&tw_mutex);
    return r;
}

/* In-place unnormalised complex DFT of power-of-two length n.
 * inverse != 0 uses e^{+2 pi i ...}. */
static void fft_c(cd *a, int n, int inverse)
{
    if (n <= 1) return;
    const cd *tw = twiddles(n);
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cd t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; ++j) {
                cd w = tw[j * step];
                if (inverse) w.im = -w.im;
                cd u = a[i + j], v = a[i + j + half];
                cd t;
                t.re = v.re * w.re - v.im * w.im;
                t.im = v.re * w.im + v.im * w.re;
                a[i + j].re = u.re + t.re;  a[i + j].im = u.im + t.im;
                a[i + j + half].re = u.re - t.re;  a[i + j + half].im = u.im - t.im;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Optional FFTW3-API provider (CPU baseline only)                     */
/* ------------------------------------------------------------------ */

typedef void *(*pf_plan_r2c)(int, int, double *, double (*)[2], unsigned);
typedef void *(*pf_plan_c2r)(int, int, double (*)[2], double *, unsigned);
typedef void (*pf_exec)(void *);
typedef void (*pf_destroy)(void *);
#define XO_FFTW_ESTIMATE (1U << 6)

static struct {
    int          active;
    void        *lib;
    pf_plan_r2c  r2c;
    pf_plan_c2r  c2r;
    pf_exec      exec;
    pf_destroy   destroy;
    char         name[256];
} prov;

/* mutex_fft (Correlation.cpp:26-27): serialises every FFTW call. */
static pthread_mutex_t mutex_fft = PTHREAD_MUTEX_INITIALIZER;

/* Select the FFT used by the 2-D transforms: 0 = oracle radix-2,
 * 1 = FFTW3 API from `libpath` (dlopen).  Returns 0 on success. */
int xo_set_fft_provider(int which, const char *libpath)
{
    if (which == 0) { prov.active = 0; strcpy(prov.name, "oracle radix-2 FP64"); return 0; }
    if (!prov.lib) {
        prov.lib = dlopen(libpath, RTLD_NOW | RTLD_LOCAL);
        if (!prov.lib) return -1;
        prov.r2c = (pf_plan_r2c)dlsym(prov.lib, "fftw_plan_dft_r2c_2d");
        prov.c2r = (pf_plan_c2r)dlsym(prov.lib, "fftw_plan_dft_c2r_2d");
        prov.exec = (pf_exec)dlsym(prov.lib, "fftw_execute");
        prov.destroy = (pf_destroy)dlsym(prov.lib, "fftw_destroy_plan");
        if (!prov.r2c || !prov.c2r || !prov.exec || !prov.destroy) {
            dlclose(prov.lib); prov.lib = NULL; return -2;
        }
    }
    prov.active = 1;
    snprintf(prov.name, sizeof prov.name, "FFTW3 API via %s", libpath);
    return 0;
}

const char *xo_fft_provider_name(void)
{
    if (!prov.name[0]) strcpy(prov.name, "oracle radix-2 FP64");
    return prov.name;
}

/* ------------------------------------------------------------------ */
/* 2-D real transforms: FFT_2D / IFT_2D semantics                      */
/* ------------------------------------------------------------------ */

/* FFT_2D (Correlation_fft_fftw.cpp:10-56): in[Nslow][Nfast] real ->
 * out[Nslow][Nfast/2+1] complex, unnormalised. */
int xo_fft2d_r2c(const double *in, int Nfast, int Nslow, double *out_)
{
    int H = Nfast / 2 + 1;
    cd *out = (cd *)out_;
    if (prov.active) {
        pthread_mutex_lock(&mutex_fft);
        void *p = prov.r2c(Nslow, Nfast, (double *)in, (double (*)[2])out, XO_FFTW_ESTIMATE);
        prov.exec(p);
        prov.destroy(p);
        pthread_mutex_unlock(&mutex_fft);
        return Nslow * H;
    }
    cd *row = (cd *)malloc(sizeof(cd) * (Nfast > Nslow ? Nfast : Nslow));
    for (int y = 0; y < Nslow; ++y) {
        const double *s = in + (size_t)y * Nfast;
        int nz = 0;
        for (int x = 0; x < Nfast; ++x) { row[x].re = s[x]; row[x].im = 0.0; nz |= (s[x] != 0.0); }
        if (nz) fft_c(row, Nfast, 0);
        else memset(row, 0, sizeof(cd) * Nfast);
        memcpy(out + (size_t)y * H, row, sizeof(cd) * H);
    }
    for (int k = 0; k < H; ++k) {
        for (int y = 0; y < Nslow; ++y) row[y] = out[(size_t)y * H + k];
        fft_c(row, Nslow, 0);
        for (int y = 0; y < Nslow; ++y) out[(size_t)y * H + k] = row[y];
    }
    free(row);
    return Nslow * H;
}

/* IFT_2D (Correlation_fft_fftw.cpp:67-92): unnormalised c2r.  The input
 * is not modified (the reference copies it before FFTW's c2r). */
int xo_ift2d_c2r(const double *in_, int Nfast, int Nslow, double *out)
{
    int H = Nfast / 2 + 1;
    const cd *in = (const cd *)in_;
    if (prov.active) {
        size_t M = (size_t)Nslow * H;
        cd *copy = (cd *)malloc(sizeof(cd) * M);
        memcpy(copy, in, sizeof(cd) * M);
        pthread_mutex_lock(&mutex_fft);
        void *p = prov.c2r(Nslow, Nfast, (double (*)[2])copy, out, XO_FFTW_ESTIMATE);
        prov.exec(p);
        prov.destroy(p);
        pthread_mutex_unlock(&mutex_fft);
        free(copy);
        return 0;
    }
    cd *tmp = (cd *)malloc(sizeof(cd) * (size_t)Nslow * H);
    cd *row = (cd *)malloc(sizeof(cd) * (Nfast > Nslow ? Nfast : Nslow));
    for (int k = 0; k < H; ++k) {
        for (int y = 0; y < Nslow; ++y) row[y] = in[(size_t)y * H + k];
        fft_c(row, Nslow, 1);
        for (int y = 0; y < Nslow; ++y) tmp[(size_t)y * H + k] = row[y];
    }
    for (int y = 0; y < Nslow; ++y) {
        const cd *t = tmp + (size_t)y * H;
        for (int k = 0; k < H; ++k) row[k] = t[k];
        for (int k = H; k < Nfast; ++k) { row[k].re = t[Nfast - k].re; row[k].im = -t[Nfast - k].im; }
        fft_c(row, Nfast, 1);
        for (int x = 0; x < Nfast; ++x) out[(size_t)y * Nfast + x] = row[x].re;
    }
    free(row);
    free(tmp);
    return 0;
}

/* 1-D complex transform, exported for the DFT known-answer tests. */
void xo_fft1d(double *a, int n, int inverse) { fft_c((cd *)a, n, inverse); }

/* ------------------------------------------------------------------ */
/* Geometry and maths helpers                                          */
/* ------------------------------------------------------------------ */

typedef struct { int L, R, B, T; } ibox;

/* CeilPow2 (Maths.cpp:97-105). */
int xo_ceil_pow2(int n)
{
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

/* FFTSize (Correlation.cpp:344-347), FFTSizeSQR (:356-359). */
int xo_fft_size(int n1, int n2) { return xo_ceil_pow2(n1 + n2 - 1); }
int xo_fft_size_sqr(int w1, int h1, int w2, int h2)
{
    int a = w1 + w2, b = h1 + h2;
    return xo_ceil_pow2((a > b ? a : b) - 1);
}

/* BBoxFromPoints(IBox&) (Geometry.cpp:78-106), including its
 * "else if" update order. */
static void bbox_from_points(ibox *B, const xc_point *p, int n)
{
    double xlo, xhi, ylo, yhi;
    xlo = xhi = p[0].x;
    ylo = yhi = p[0].y;
    for (int i = 1; i < n; ++i) {
        double t;
        if ((t = p[i].x) < xlo) xlo = t;
        else if (t > xhi) xhi = t;
        if ((t = p[i].y) < ylo) ylo = t;
        else if (t > yhi) yhi = t;
    }
    B->L = (int)floor(xlo);
    B->R = (int)ceil(xhi);
    B->B = (int)floor(ylo);
    B->T = (int)ceil(yhi);
}

void xo_bbox_from_points(const xc_point *p, int n, int *LRBT)
{
    ibox b;
    bbox_from_points(&b, p, n);
    LRBT[0] = b.L; LRBT[1] = b.R; LRBT[2] = b.B; LRBT[3] = b.T;
}

/* BoxesFromShifts (Geometry.cpp:150-176). */
static void boxes_from_shifts(ibox *B1, ibox *B2, int w1, int h1, int w2, int h2, int x, int y)
{
    B2->L = x > 0 ? x : 0;
    B2->R = (w2 - 1) < (x + w1 - 1) ? (w2 - 1) : (x + w1 - 1);
    B2->B = y > 0 ? y : 0;
    B2->T = (h2 - 1) < (y + h1 - 1) ? (h2 - 1) : (y + h1 - 1);
    if (B2->R < B2->L) B2->R = B2->L;
    if (B2->T < B2->B) B2->T = B2->B;
    B1->L = B2->L - x;
    B1->R = B2->R - x;
    B1->B = B2->B - y;
    B1->T = B2->T - y;
}

void xo_boxes_from_shifts(int w1, int h1, int w2, int h2, int x, int y, int *b1, int *b2)
{
    ibox B1, B2;
    boxes_from_shifts(&B1, &B2, w1, h1, w2, h2, x, y);
    b1[0] = B1.L; b1[1] = B1.R; b1[2] = B1.B; b1[3] = B1.T;
    b2[0] = B2.L; b2[1] = B2.R; b2[2] = B2.B; b2[3] = B2.T;
}

/* DistributePixel (Maths.h:236-276): bilinear splat, in the reference's
 * order of the four += updates. */
static void distribute_pixel(double x, double y, double val, double *image, int w, int h)
{
    int    ix = (int)floor(x);
    int    iy = (int)floor(y);
    long   i = ix + (long)w * iy;
    double alpha = x - ix;
    double beta = y - iy;

    if (ix >= 0 && ix < w) {
        if (iy >= 0 && iy < h) image[i] += (1.0 - alpha) * (1.0 - beta) * val;
        ++iy;
        if (iy >= 0 && iy < h) image[i + w] += (1.0 - alpha) * beta * val;
        --iy;
    }
    ++ix;
    ++i;
    if (ix >= 0 && ix < w) {
        if (iy >= 0 && iy < h) image[i] += alpha * (1.0 - beta) * val;
        ++iy;
        if (iy >= 0 && iy < h) image[i + w] += alpha * beta * val;
    }
}

/* ImageFromValuesAndPoints (Geometry.cpp:661-685): zero-filled w*h raster,
 * each point splatted at (x - xmin, y - ymin). */
void xo_image_from_values_and_points(double *I, int w, int h, const double *v,
                                     const xc_point *p, int n, int xmin, int ymin)
{
    memset(I, 0, sizeof(double) * (size_t)w * h);
    for (int i = 0; i < n; ++i)
        distribute_pixel(p[i].x - xmin, p[i].y - ymin, v[i], I, w, h);
}

/* MeanStd::Stats (Maths.cpp:111-125, n-1 form) over v[0..n). */
void xo_stats(const double *v, int n, double *avg, double *std)
{
    double sum = 0.0, sum2 = 0.0;
    for (int i = 0; i < n; ++i) { sum += v[i]; sum2 += v[i] * v[i]; }
    *avg = sum;
    *std = 0.0;
    if (n > 1) {
        *avg /= n;
        *std = sqrt(fmax(sum2 - n * (*avg) * (*avg), 0.0) / (n - 1.0));
    }
}

/* Normalize (Maths.cpp:192-212): mean 0, sd 1; returns the sd. */
double xo_normalize(double *v, int n)
{
    double avg, std;
    xo_stats(v, n, &avg, &std);
    for (int i = 0; i < n; ++i) v[i] = (v[i] - avg) / std;
    return std;
}

/* IntegrateImage (Correlation.cpp:49-105): inclusive prefix sums of the
 * wS x hS region of I (row stride wI): S, S^2 and the non-zero count. */
void xo_integrate_image(double *S, double *S2, int *nz, int wS, int hS, const double *I, int wI)
{
    double t;
    S[0] = (t = I[0]);
    S2[0] = t * t;
    nz[0] = (t != 0.0);
    for (int x = 1; x < wS; ++x) {
        S[x] = S[x - 1] + (t = I[x]);
        S2[x] = S2[x - 1] + t * t;
        nz[x] = nz[x - 1] + (t != 0.0);
    }
    for (int y = 1; y < hS; ++y) {
        long j = (long)wS * y, k = (long)wI * y;
        S[j] = S[j - wS] + (t = I[k]);
        S2[j] = S2[j - wS] + t * t;
        nz[j] = nz[j - wS] + (t != 0.0);
    }
    for (int y = 1; y < hS; ++y) {
        long i = (long)wS * y, k = (long)wI * y;
        for (int x = 1; x < wS; ++x) {
            long j = i + x, m = k + x;
            S[j] = S[j - 1] + S[j - wS] - S[j - wS - 1] + (t = I[m]);
            S2[j] = S2[j - 1] + S2[j - wS] - S2[j - wS - 1] + t * t;
            nz[j] = nz[j - 1] + nz[j - wS] - nz[j - wS - 1] + (t != 0.0);
        }
    }
}

/* IntegralTable (Correlation.cpp:160-201). */
static double table_d(const double *t, int wS, const ibox *B)
{
    double r = t[B->R + (long)wS * B->T];
    if (B->L > 0) r -= t[B->L - 1 + (long)wS * B->T];
    if (B->B > 0) r -= t[B->R + (long)wS * (B->B - 1)];
    if (B->L > 0 && B->B > 0) r += t[(B->L - 1) + (long)wS * (B->B - 1)];
    return r;
}

static int table_i(const int *t, int wS, const ibox *B)
{
    int r = t[B->R + (long)wS * B->T];
    if (B->L > 0) r -= t[B->L - 1 + (long)wS * B->T];
    if (B->B > 0) r -= t[B->R + (long)wS * (B->B - 1)];
    if (B->L > 0 && B->B > 0) r += t[(B->L - 1) + (long)wS * (B->B - 1)];
    return r;
}

/* Legality rules: the descriptor form of the EvalType callbacks. */
static int rgn_ok(const xc_rgn_rule *g, int olw, int olh)
{
    if (!g || !g->enabled) return 1;
    if (olw < g->min_w || olh < g->min_h) return 0;
    if (!((int64_t)olw * olh > g->min_area)) return 0;
    if (g->exact_w > 0 && olw != g->exact_w) return 0;
    if (g->exact_h > 0 && olh != g->exact_h) return 0;
    if ((g->or_w_ge > 0 || g->or_h_gt > 0) && !(olw >= g->or_w_ge || olh > g->or_h_gt)) return 0;
    if (g->table) {
        if (olw > g->table_w || olh > g->table_h) return 0;
        if (!g->table[(long)(olh - 1) * g->table_w + (olw - 1)]) return 0;
    }
    return 1;
}

static int cnt_ok(const xc_cnt_rule *c, int c1, int c2)
{
    if (!c || !c->enabled) return 1;
    return (int64_t)c1 > c->k1 && (int64_t)c2 > c->k2;
}

int xo_rgn_ok(const xc_rgn_rule *g, int olw, int olh) { return rgn_ok(g, olw, olh); }

/* ------------------------------------------------------------------ */
/* Shared FFT correlation core                                         */
/* ------------------------------------------------------------------ */

/* rslt = IFT_2D(FFT_2D(i2) * conj(FFT_2D(i1))) (Correlation.cpp:965-989).
 * If whitened != NULL it also receives IFT_2D of the cross power divided
 * by sqrt(|.|) where |.| > 1e-10 (Correlation.cpp:2311-2319). */
static void cross_power_inverse(const double *i1, const double *i2, int Nx, int Ny,
                                double *rslt, double *whitened)
{
    int  H = Nx / 2 + 1;
    long M = (long)Ny * H;
    cd *f1 = (cd *)malloc(sizeof(cd) * M);
    cd *f2 = (cd *)malloc(sizeof(cd) * M);
    xo_fft2d_r2c(i2, Nx, Ny, (double *)f2);
    xo_fft2d_r2c(i1, Nx, Ny, (double *)f1);
    for (long i = 0; i < M; ++i) {
        /* fft1[i] = fft2[i] * conj(fft1[i]) */
        double a = f2[i].re, b = f2[i].im, c = f1[i].re, d = f1[i].im;
        f1[i].re = a * c + b * d;
        f1[i].im = b * c - a * d;
    }
    if (rslt) xo_ift2d_c2r((const double *)f1, Nx, Ny, rslt);
    if (whitened) {
        for (long i = 0; i < M; ++i) {
            double mag = hypot(f1[i].re, f1[i].im);
            if (mag > 1e-10) {
                double s = sqrt(mag);
                f1[i].re /= s;
                f1[i].im /= s;
            }
        }
        xo_ift2d_c2r((const double *)f1, Nx, Ny, whitened);
    }
    free(f1);
    free(f2);
}

/* LkFFT (Correlation.cpp:287-298). */
static double lk_fft(const double *fr, int nx, int ny, int x, int y)
{
    if (x < 0) x += nx;
    if (y < 0) y += ny;
    return fr[x + (long)nx * y];
}

/* ParabPeakFFT (Correlation.cpp:482-517): the refinement is applied only
 * for signed lags in [d, n-d) and non-zero neighbours. */
static void parab_peak_fft(double *xpk, double *ypk, int d, const double *I, int nx, int ny)
{
    int    ix = (int)*xpk, iy = (int)*ypk;
    double z1 = lk_fft(I, nx, ny, ix, iy), z0, z2;

    if (ix - d >= 0 && ix + d < nx &&
        (z0 = lk_fft(I, nx, ny, ix - d, iy)) != 0.0 &&
        (z2 = lk_fft(I, nx, ny, ix + d, iy)) != 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) *xpk += z0;
    }
    if (iy - d >= 0 && iy + d < ny &&
        (z0 = lk_fft(I, nx, ny, ix, iy - d)) != 0.0 &&
        (z2 = lk_fft(I, nx, ny, ix, iy + d)) != 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) *ypk += z0;
    }
}

void xo_parab_peak_fft(double *xpk, double *ypk, int d, const double *I, int nx, int ny)
{
    parab_peak_fft(xpk, ypk, d, I, nx, ny);
}

static void fill_result(xc_result *res, int status, double r, double dx, double dy, double peak,
                        int lx, int ly, double sx, double sy, int nx, int ny)
{
    if (!res) return;
    memset(res, 0, sizeof *res);
    res->status = status;
    res->r = r;
    res->dx = dx;
    res->dy = dy;
    res->peak = peak;
    res->lag_x = lx;
    res->lag_y = ly;
    res->sub_x = sx;
    res->sub_y = sy;
    res->nx = nx;
    res->ny = ny;
}

/* ------------------------------------------------------------------ */
/* CorrPatches                                                         */
/* ------------------------------------------------------------------ */

/* CorrPatches (Correlation.cpp:901-1170) with CLinCorr (:592-760).
 * rslt_out (optional, Nx*Ny doubles) receives the scanned R image: NCC at
 * legal lags, 0 elsewhere, in wrapped raster order. */
int xo_corr_patches(const xc_point *ip1, const double *iv1, int n1,
                    const xc_point *ip2, const double *iv2, int n2,
                    int Ox, int Oy, int radius,
                    const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                    double *dx, double *dy, double *r_out, xc_result *res,
                    double *rslt_out, int rslt_cap)
{
    if (n1 <= 0 || n2 <= 0) return XC_E_ARG;
    ibox B1, B2;
    bbox_from_points(&B1, ip1, n1);
    bbox_from_points(&B2, ip2, n2);
    int w1 = B1.R - B1.L + 1, h1 = B1.T - B1.B + 1;
    int w2 = B2.R - B2.L + 1, h2 = B2.T - B2.B + 1;
    int Nx = xo_fft_size(w1, w2), Ny = xo_fft_size(h1, h2);
    long Nxy = (long)Nx * Ny;

    double *i1 = (double *)malloc(sizeof(double) * Nxy);
    double *i2 = (double *)malloc(sizeof(double) * Nxy);
    double *rslt = (double *)malloc(sizeof(double) * Nxy);
    xo_image_from_values_and_points(i1, Nx, Ny, iv1, ip1, n1, B1.L, B1.B);
    xo_image_from_values_and_points(i2, Nx, Ny, iv2, ip2, n2, B2.L, B2.B);
    cross_power_inverse(i1, i2, Nx, Ny, rslt, NULL);

    /* CLinCorr::Initialize (:639-662) */
    double *s1 = (double *)malloc(sizeof(double) * (long)w1 * h1);
    double *q1 = (double *)malloc(sizeof(double) * (long)w1 * h1);
    int    *z1 = (int *)malloc(sizeof(int) * (long)w1 * h1);
    double *s2 = (double *)malloc(sizeof(double) * (long)w2 * h2);
    double *q2 = (double *)malloc(sizeof(double) * (long)w2 * h2);
    int    *z2 = (int *)malloc(sizeof(int) * (long)w2 * h2);
    xo_integrate_image(s1, q1, z1, w1, h1, i1, Nx);
    xo_integrate_image(s2, q2, z2, w2, h2, i2, Nx);

    double biggest = -1.0E30;
    int bigx = -1, bigy = -1;
    int nnegx = w1 - 1, nnegy = h1 - 1, nposx = w2, nposy = h2;
    int rsqr = radius * radius;
    int status = XC_OK;

    for (int iy = 0; iy < Ny && status == XC_OK; ++iy) {
        int dyy, y = iy;
        if (y >= Ny - nnegy) y -= Ny;
        else if (y >= nposy) {
            memset(&rslt[(long)Nx * iy], 0, Nx * sizeof(double));
            continue;
        }
        dyy = y - Oy;
        for (int ix = 0; ix < Nx; ++ix) {
            int ddx, x = ix;
            long i = ix + (long)Nx * iy;
            if (x >= Nx - nnegx) x -= Nx;
            else if (x >= nposx) { rslt[i] = 0.0; continue; }
            ddx = x - Ox;
            if (ddx * ddx + dyy * dyy > rsqr) { rslt[i] = 0.0; continue; }

            /* CheckSize (:665-694) */
            ibox OL1, OL2;
            boxes_from_shifts(&OL1, &OL2, w1, h1, w2, h2, x, y);
            int olw = OL1.R - OL1.L + 1, olh = OL1.T - OL1.B + 1;
            if (!rgn_ok(rgn, olw, olh)) { rslt[i] = 0.0; continue; }
            /* CheckDensity (:697-713) */
            int i1c = table_i(z1, w1, &OL1), i2c = table_i(z2, w2, &OL2);
            if (!cnt_ok(cnt, i1c, i2c)) { rslt[i] = 0.0; continue; }
            /* GetCorr (:716-754) */
            double n = (double)(olw * olh);
            double im1sum = table_d(s1, w1, &OL1), im1sum2 = table_d(q1, w1, &OL1);
            double im2sum = table_d(s2, w2, &OL2), im2sum2 = table_d(q2, w2, &OL2);
            double num = n * rslt[i] / Nxy - im1sum * im2sum;
            double d1 = n * im1sum2 - im1sum * im1sum;
            double d2 = n * im2sum2 - im2sum * im2sum;
            double d = d1 * d2;
            double r = (d < n * n * 1.0E-9 ? 0.0 : num / sqrt(d));
            if (fabs(r) > 1.0001) { status = XC_E_ODD_R; break; }
            rslt[i] = r;
            if (r > biggest) { biggest = r; bigx = x; bigy = y; }
        }
    }

    double rv = 0.0;
    if (status == XC_OK) {
        if (biggest < -2.0) {
            status = XC_NO_PEAK;
            fill_result(res, status, 0.0, dx ? *dx : 0.0, dy ? *dy : 0.0, 0.0, 0, 0, 0, 0, Nx, Ny);
        } else {
            double px = bigx, py = bigy;
            parab_peak_fft(&px, &py, 1, rslt, Nx, Ny);
            double ox = px + (B2.L - B1.L), oy = py + (B2.B - B1.B);
            if (dx) *dx = ox;
            if (dy) *dy = oy;
            rv = biggest;
            fill_result(res, status, biggest, ox, oy, biggest, bigx, bigy, px - bigx, py - bigy, Nx, Ny);
        }
    } else {
        fill_result(res, status, 0.0, 0, 0, 0, 0, 0, 0, 0, Nx, Ny);
    }
    if (r_out) *r_out = rv;
    if (rslt_out && rslt_cap >= Nxy) memcpy(rslt_out, rslt, sizeof(double) * Nxy);

    free(i1); free(i2); free(rslt);
    free(s1); free(q1); free(z1); free(s2); free(q2); free(z2);
    return status;
}

/* ------------------------------------------------------------------ */
/* CorrPatchToImage                                                    */
/* ------------------------------------------------------------------ */

/* CorrPatchToImage (Correlation.cpp:1186-1355). */
int xo_corr_patch_to_image(const xc_point *ip1, const double *iv1, int np1,
                           const double *i2, int Nsqr,
                           int Ox, int Oy, int radius, int bFilter,
                           double *dx, double *dy, double *r_out, xc_result *res)
{
    if (np1 <= 0 || Nsqr <= 0) return XC_E_ARG;
    int N = (int)sqrt((double)Nsqr);
    double norm = (double)np1 * Nsqr;
    ibox B1;
    bbox_from_points(&B1, ip1, np1);
    int w1 = B1.R - B1.L + 1, h1 = B1.T - B1.B + 1;

    int w2 = 0, h2 = 0;
    for (int i = 0; i < Nsqr; ++i) {
        if (i2[i] != 0.0) {
            int y = i / N, x = i - N * y;
            if (y > h2) h2 = y;
            if (x > w2) w2 = x;
        }
    }
    ++w2;
    ++h2;

    double *i1 = (double *)calloc((size_t)Nsqr, sizeof(double));
    for (int i = 0; i < np1; ++i) {
        int x = (int)ip1[i].x, y = (int)ip1[i].y;
        i1[x + N * y] = iv1[i];
    }

    int  H = N / 2 + 1;
    long M = (long)N * H;
    cd *f1 = (cd *)malloc(sizeof(cd) * M);
    cd *f2 = (cd *)malloc(sizeof(cd) * M);
    double *rslt = (double *)malloc(sizeof(double) * Nsqr);
    xo_fft2d_r2c(i2, N, N, (double *)f2);
    xo_fft2d_r2c(i1, N, N, (double *)f1);
    for (long i = 0; i < M; ++i) {
        double a = f2[i].re, b = f2[i].im, c = f1[i].re, d = f1[i].im;
        double pr = a * c + b * d, pi = b * c - a * d;
        if (bFilter) {
            /* :1255-1283 -- note the reference's x = i / yrow (row index). */
            int yrow = N / 2 + 1;
            int x = (int)(i / yrow);
            int y = (int)(i - (long)yrow * x);
            if (x >= N / 2) x -= N;
            double rad2 = ((double)x * x + (double)y * y) / 10000.0;
            double filt = rad2 > 10.0 ? 0.0 : exp(-rad2);
            pr = pr * filt;
            pi = pi * filt;
        }
        f1[i].re = pr;
        f1[i].im = pi;
    }
    xo_ift2d_c2r((const double *)f1, N, N, rslt);

    double biggest = -1.0E30;
    int bigx = -1, bigy = -1;
    int rsqr = radius * radius;
    for (int i = 0; i < Nsqr; ++i) {
        int y = i / N, x = i - N * y;
        if (y > N - h1) y -= N;
        else if (y >= h2) continue;
        if (x > N - w1) x -= N;
        else if (x >= w2) continue;
        int ddx = x - Ox, ddy = y - Oy;
        if (ddx * ddx + ddy * ddy > rsqr) continue;
        if (rslt[i] > biggest) { biggest = rslt[i]; bigx = x; bigy = y; }
    }
    biggest /= norm;

    int status = XC_OK;
    if (biggest > 1.0001) {
        status = XC_E_ODD_R;
        fill_result(res, status, biggest, 0, 0, biggest, bigx, bigy, 0, 0, N, N);
    } else {
        double px = bigx, py = bigy;
        parab_peak_fft(&px, &py, 1, rslt, N, N);
        if (dx) *dx = px;
        if (dy) *dy = py;
        if (r_out) *r_out = biggest;
        fill_result(res, status, biggest, px, py, biggest, bigx, bigy, px - bigx, py - bigy, N, N);
    }
    free(i1); free(f1); free(f2); free(rslt);
    return status;
}

/* ------------------------------------------------------------------ */
/* CorrImagesR / CorrImagesS                                           */
/* ------------------------------------------------------------------ */

/* ParabPeak (Correlation.cpp:2975-3010): neighbours must be > 0. */
static void parab_peak(double *xpk, double *ypk, int d, const double *I, int w, int h)
{
    int    ix = (int)*xpk, iy = (int)*ypk;
    double z1 = I[ix + (long)w * iy], z0, z2;
    if (ix - d >= 0 && ix + d < w &&
        (z0 = I[ix - d + (long)w * iy]) > 0.0 &&
        (z2 = I[ix + d + (long)w * iy]) > 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) *xpk += z0;
    }
    if (iy - d >= 0 && iy + d < h &&
        (z0 = I[ix + (long)w * (iy - d)]) > 0.0 &&
        (z2 = I[ix + (long)w * (iy + d)]) > 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) *ypk += z0;
    }
}

/* Convolve (Correlation.cpp:371-457): dst (ws x hs) = src convolved with
 * the wk x hk kernel K through 2-D transforms of size FFTSize(ws, wk/(1+sym))
 * x FFTSize(hs, hk/(1+sym)), the kernel centred on the origin with wrap,
 * divided by its sum when prenorm (a zero sum leaves it as is), product of
 * the two r2c transforms, unnormalised c2r, then / Nxy. */
int xo_convolve(double *dst, const double *src, int ws, int hs, const double *K, int wk, int hk, int sym, int prenorm)
{
    int Nx = xo_fft_size(ws, wk / (1 + sym)), Ny = xo_fft_size(hs, hk / (1 + sym));
    long Nxy = (long)Nx * Ny, M = (long)Ny * (Nx / 2 + 1);
    double *KK = (double *)calloc(Nxy, sizeof(double));
    double *SS = (double *)calloc(Nxy, sizeof(double));
    cd *kf = (cd *)malloc(sizeof(cd) * M), *sf = (cd *)malloc(sizeof(cd) * M);
    double nrm = 1.0;
    if (prenorm) {
        nrm = 0.0;
        for (int i = 0; i < wk * hk; ++i) nrm += K[i];
        if (!nrm) nrm = 1.0;
    }
    int w2 = wk / 2, h2 = hk / 2;
    for (int i = -h2; i <= h2; ++i) {
        int y = (i >= 0 ? i : Ny + i);
        for (int j = -w2; j <= w2; ++j) {
            int x = (j >= 0 ? j : Nx + j);
            KK[x + (long)Nx * y] = K[j + w2 + wk * (i + h2)] / nrm;
        }
    }
    xo_fft2d_r2c(KK, Nx, Ny, (double *)kf);
    for (int y = 0; y < hs; ++y) memcpy(SS + (long)Nx * y, src + (long)ws * y, sizeof(double) * ws);
    xo_fft2d_r2c(SS, Nx, Ny, (double *)sf);
    for (long i = 0; i < M; ++i) {  /* std::complex<double> operator*= */
        double re = sf[i].re * kf[i].re - sf[i].im * kf[i].im;
        double im = sf[i].re * kf[i].im + sf[i].im * kf[i].re;
        sf[i].re = re;
        sf[i].im = im;
    }
    xo_ift2d_c2r((const double *)sf, Nx, Ny, SS);
    for (int y = 0; y < hs; ++y)
        for (int x = 0; x < ws; ++x) dst[x + (long)ws * y] = SS[x + (long)Nx * y] / Nxy;
    free(KK); free(SS); free(kf); free(sf);
    return 0;
}

/* CCorImg::FPeak (Correlation.cpp:2580-2767): walk the candidates `order`
 * (highest F first); skip one inside the core of an earlier candidate;
 * inner radius ri = the least of the four axis distances to a pixel with
 * F <= fpk/2 (none in some direction: reject); reject ri <= 1 with
 * R < 0.1 and an outer band ro = 3 ri that leaves the image; mask the core
 * (half-width max(ri/3, 1)); accept when no pixel of the ro box outside
 * the ri box exceeds nbmaxht * fpk.  Returns 1 with (rx, ry), else 0. */
int xo_fpeak(const double *F, const double *R, int wR, int hR, const long *order, long n, double nbmaxht, int *rxo,
             int *ryo)
{
    uint8_t *mask = (uint8_t *)calloc((size_t)wR * hR, 1);
    int ok = 0;
    for (long i = 0; i < n && !ok; ++i) {
        long k = order[i];
        if (mask[k]) continue;
        double fpk = F[k], tol = nbmaxht * fpk;
        int ry = (int)(k / wR), rx = (int)(k - (long)wR * ry);
        int ri = 0x7fffffff, hit, y, x;
        hit = 0;
        for (y = ry - 1; y >= 0; --y) if (F[rx + (long)wR * y] <= 0.5 * fpk) { if (ry - y < ri) ri = ry - y; hit = 1; break; }
        if (!hit) continue;
        hit = 0;
        for (y = ry + 1; y < hR; ++y) if (F[rx + (long)wR * y] <= 0.5 * fpk) { if (y - ry < ri) ri = y - ry; hit = 1; break; }
        if (!hit) continue;
        hit = 0;
        for (x = rx - 1; x >= 0; --x) if (F[x + (long)wR * ry] <= 0.5 * fpk) { if (rx - x < ri) ri = rx - x; hit = 1; break; }
        if (!hit) continue;
        hit = 0;
        for (x = rx + 1; x < wR; ++x) if (F[x + (long)wR * ry] <= 0.5 * fpk) { if (x - rx < ri) ri = x - rx; hit = 1; break; }
        if (!hit) continue;
        if (ri <= 1 && R[k] < 0.1) continue;                /* Micropk */
        int ro = 3 * ri;
        if (ry - ro < 0 || ry + ro >= hR || rx - ro < 0 || rx + ro >= wR) continue;  /* Ro_edge */
        int rm = ri / 3;
        if (rm < 1) rm = 1;
        for (y = ry - rm; y <= ry + rm; ++y)
            for (x = rx - rm; x <= rx + rm; ++x) mask[x + (long)wR * y] = 1;
        int clean = 1;
        for (y = ry - ro; y <= ry + ro && clean; ++y)
            for (x = rx - ro; x <= rx + ro; ++x) {
                if (y >= ry - ri && y <= ry + ri && x >= rx - ri && x <= rx + ri) continue;
                if (F[x + (long)wR * y] > tol) { clean = 0; break; }  /* Hi_neib */
            }
        if (clean) {
            *rxo = rx;
            *ryo = ry;
            ok = 1;
        }
    }
    free(mask);
    return ok;
}

typedef struct { double f; long i; } fkey;

/* OrderF's sort (Correlation.cpp:2542-2546): decreasing F; exact ties by
 * increasing index. */
static int fkey_cmp(const void *a, const void *b)
{
    const fkey *p = (const fkey *)a, *q = (const fkey *)b;
    if (p->f > q->f) return -1;
    if (p->f < q->f) return 1;
    return p->i < q->i ? -1 : (p->i > q->i);
}

/* CCorImg::SetDims / MakeRandA / MakeSandRandA / MaskA / MakeF / OrderF /
 * FPeak / SimpleMax / ReturnR and CorrImagesF / CorrImagesR / CorrImagesS
 * (Correlation.cpp:1949-2370, 2412-2767, 2939-3122, 3144-3290).
 * mode: XC_MODE_R, XC_MODE_S or XC_MODE_F.  Optional outputs (each nR =
 * wR*hR): R_out, S_out (the S image in S mode, the F image in F mode),
 * A_out (F mode: with MakeF's guard band cleared).  On exact ties in
 * OrderF's sort the oracle takes the lowest centred-image index
 * (std::sort's order on exact ties is unspecified). */
int xo_corr_images(int mode,
                   const xc_point *ip1, const double *iv1, int n1,
                   const xc_point *ip2, const double *iv2, int n2,
                   const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                   double mincor, double nbmaxht,
                   int Ox, int Oy, int Rx, int Ry,
                   double *dx, double *dy, double *r_out, xc_result *res,
                   double *R_out, double *S_out, uint8_t *A_out, int cap)
{
    if (n1 <= 0 || n2 <= 0) return XC_E_ARG;
    if (mode != XC_MODE_R && mode != XC_MODE_S && mode != XC_MODE_F) return XC_E_ARG;
    ibox B1, B2;
    bbox_from_points(&B1, ip1, n1);
    bbox_from_points(&B2, ip2, n2);
    int w1 = B1.R - B1.L + 1, h1 = B1.T - B1.B + 1;
    int w2 = B2.R - B2.L + 1, h2 = B2.T - B2.B + 1;
    int nnegx = w1 - 1, nposx = w2, nnegy = h1 - 1, nposy = h2;
    int cx = nnegx, cy = nnegy, wR = nnegx + nposx, hR = nnegy + nposy;
    long nR = (long)wR * hR;
    int Nx = xo_fft_size(w1, w2), Ny = xo_fft_size(h1, h2);
    long Nxy = (long)Nx * Ny;
    const int mindelta = 11 / 2 + 1;

    if (nnegx < mindelta || nposx < mindelta || nnegy < mindelta || nposy < mindelta) {
        if (dx) *dx = 0.0;
        if (dy) *dy = 0.0;
        if (r_out) *r_out = 0.0;
        fill_result(res, XC_TOO_SMALL, 0, 0, 0, 0, 0, 0, 0, 0, Nx, Ny);
        return XC_TOO_SMALL;
    }

    double *i1 = (double *)malloc(sizeof(double) * Nxy);
    double *i2 = (double *)malloc(sizeof(double) * Nxy);
    double *rslt = (double *)malloc(sizeof(double) * Nxy);
    double *wrs = mode == XC_MODE_S ? (double *)malloc(sizeof(double) * Nxy) : NULL;
    xo_image_from_values_and_points(i1, Nx, Ny, iv1, ip1, n1, B1.L, B1.B);
    xo_image_from_values_and_points(i2, Nx, Ny, iv2, ip2, n2, B2.L, B2.B);
    cross_power_inverse(i1, i2, Nx, Ny, rslt, wrs);

    /* RCalc::Initialize (:1771-1791) */
    double *s1 = (double *)malloc(sizeof(double) * (long)w1 * h1);
    double *q1 = (double *)malloc(sizeof(double) * (long)w1 * h1);
    int    *z1 = (int *)malloc(sizeof(int) * (long)w1 * h1);
    double *s2 = (double *)malloc(sizeof(double) * (long)w2 * h2);
    double *q2 = (double *)malloc(sizeof(double) * (long)w2 * h2);
    int    *z2 = (int *)malloc(sizeof(int) * (long)w2 * h2);
    xo_integrate_image(s1, q1, z1, w1, h1, i1, Nx);
    xo_integrate_image(s2, q2, z2, w2, h2, i2, Nx);

    double  *R = (double *)malloc(sizeof(double) * nR);
    double  *S = mode == XC_MODE_S ? (double *)malloc(sizeof(double) * nR) : NULL;
    uint8_t *A = (uint8_t *)malloc(nR);

    for (int y = -nnegy; y < nposy; ++y) {
        long iy = (long)Nx * (y >= 0 ? y : Ny + y);
        for (int x = -nnegx; x < nposx; ++x) {
            long ix = (x >= 0 ? x : Nx + x);
            long ir = cx + x + (long)wR * (cy + y);
            /* RCalc::Valid (:1794-1826) */
            ibox OL1, OL2;
            boxes_from_shifts(&OL1, &OL2, w1, h1, w2, h2, x, y);
            int olw = OL1.R - OL1.L + 1, olh = OL1.T - OL1.B + 1;
            int ok = rgn_ok(rgn, olw, olh);
            if (cnt && cnt->enabled) {
                int i1c = table_i(z1, w1, &OL1), i2c = table_i(z2, w2, &OL2);
                if (!cnt_ok(cnt, i1c, i2c)) ok = 0;
            }
            A[ir] = (uint8_t)ok;
            /* RCalc::CalcR (:1829-1845) */
            double n = (double)(olw * olh);
            double im1sum = table_d(s1, w1, &OL1), im1sum2 = table_d(q1, w1, &OL1);
            double im2sum = table_d(s2, w2, &OL2), im2sum2 = table_d(q2, w2, &OL2);
            double num = n * rslt[ix + iy] / Nxy - im1sum * im2sum;
            double d1 = n * im1sum2 - im1sum * im1sum;
            double d2 = n * im2sum2 - im2sum * im2sum;
            double d = d1 * d2;
            double r = (d < n * n * 1.0E-9 ? 0.0 : num / sqrt(d));
            R[ir] = (r > -1.0 && r < 1.0 ? r : 0.0);
            /* CalcS (:1766-1767) */
            if (S) S[ir] = wrs[ix + iy] / Nxy;
        }
    }

    /* MaskA (:2011-2054) */
    if (Rx > 0 && Ry > 0) {
        int ox = Ox + B1.L - B2.L + cx, oy = Oy + B1.B - B2.B + cy;
        int xmin = ox - Rx, xmax = ox + Rx, ymin = oy - Ry, ymax = oy + Ry;
        double rx2 = Rx * Rx, asq = rx2 / ((double)Ry * Ry);
        for (long i = 0; i < nR; ++i) {
            int y = (int)(i / wR);
            if (y < ymin || y > ymax) { A[i] = 0; continue; }
            int x = (int)(i - (long)wR * y);
            if (x < xmin || x > xmax) { A[i] = 0; continue; }
            x -= ox;
            y -= oy;
            if (x * x + asq * y * y > rx2) A[i] = 0;
        }
    }

    const double *F = (mode == XC_MODE_S) ? S : R;
    long best = -1;
    double *Fimg = NULL;
    if (mode == XC_MODE_F) {
        /* MakeF (:2412-2492): LoG+I filter of R, guard band of 2 cleared in A
         * and F, then F / max(F, 1e-7) where F > 0, else 0 */
        static const double K[9] = {-1, -1, -1, -1, 9, -1, -1, -1, -1};
        const int grd = 2;
        Fimg = (double *)malloc(sizeof(double) * nR);
        xo_convolve(Fimg, R, wR, hR, K, 3, 3, 1, 1);
        for (long i = 0; i < (long)grd * wR; ++i) {
            A[i] = 0; A[nR - 1 - i] = 0;
            Fimg[i] = 0.0; Fimg[nR - 1 - i] = 0.0;
        }
        for (int y = grd; y < hR - grd; ++y)
            for (int x = 0; x < grd; ++x) {
                long i = x + (long)wR * y;
                A[i] = 0; Fimg[i] = 0.0;
                i = wR - 1 - x + (long)wR * y;
                A[i] = 0; Fimg[i] = 0.0;
            }
        double vmax = 1e-7;
        for (long i = 0; i < nR; ++i) if (Fimg[i] > vmax) vmax = Fimg[i];
        for (long i = 0; i < nR; ++i) Fimg[i] = Fimg[i] > 0.0 ? Fimg[i] / vmax : 0.0;
        /* OrderF (:2509-2569): A pixels (with R > mincor when mincor > 0) by decreasing F */
        fkey *keys = (fkey *)malloc(sizeof(fkey) * nR);
        long nc = 0;
        for (long i = 0; i < nR; ++i) {
            if (!A[i]) continue;
            if (mincor > 0.0 && !(R[i] > mincor)) continue;
            keys[nc].f = Fimg[i];
            keys[nc].i = i;
            ++nc;
        }
        qsort(keys, nc, sizeof(fkey), fkey_cmp);
        long *order = (long *)malloc(sizeof(long) * (nc ? nc : 1));
        for (long i = 0; i < nc; ++i) order[i] = keys[i].i;
        int rx, ry;
        if (nc > 0 && xo_fpeak(Fimg, R, wR, hR, order, nc, nbmaxht, &rx, &ry)) best = rx + (long)wR * ry;
        free(keys);
        free(order);
        F = R;  /* ReturnR( dx, dy, rx, ry, R, R ) */
    } else {
        /* OrderF (:2509-2569) + SimpleMax (:2939-2948) */
        for (long i = 0; i < nR; ++i) {
            if (!A[i]) continue;
            if (mincor > 0.0 && !(R[i] > mincor)) continue;
            if (best < 0 || F[i] > F[best]) best = i;
        }
    }

    int status = XC_OK;
    if (best < 0) {
        status = XC_NO_PEAK;
        if (dx) *dx = 0.0;
        if (dy) *dy = 0.0;
        if (r_out) *r_out = 0.0;
        fill_result(res, status, 0, 0, 0, 0, 0, 0, 0, 0, Nx, Ny);
    } else {
        int ry = (int)(best / wR), rx = (int)(best - (long)wR * ry);
        /* ReturnR (:3016-3043) */
        double bigR = F[best];
        double px = rx, py = ry;
        parab_peak(&px, &py, 1, F, wR, hR);
        double ox = px + (B2.L - B1.L - cx), oy = py + (B2.B - B1.B - cy);
        double rv = bigR;
        if (mode == XC_MODE_S) {
            /* Stats(S) then 10 (r - mn) / (sd sqrt(n)) (:3285-3289) */
            double mn, sd;
            xo_stats(S, (int)nR, &mn, &sd);
            rv = 10.0 * (bigR - mn) / (sd * sqrt((double)nR));
        }
        if (dx) *dx = ox;
        if (dy) *dy = oy;
        if (r_out) *r_out = rv;
        fill_result(res, status, rv, ox, oy, bigR, rx - cx, ry - cy, px - rx, py - ry, Nx, Ny);
    }
    if (cap >= nR) {
        if (R_out) memcpy(R_out, R, sizeof(double) * nR);
        if (S_out && S) memcpy(S_out, S, sizeof(double) * nR);
        if (S_out && Fimg) memcpy(S_out, Fimg, sizeof(double) * nR);
        if (A_out) memcpy(A_out, A, nR);
    }
    free(i1); free(i2); free(rslt); free(wrs);
    free(s1); free(q1); free(z1); free(s2); free(q2); free(z2);
    free(R); free(S); free(A); free(Fimg);
    return status;
}

/* ------------------------------------------------------------------ */
/* CThmScan::RFromAngle                                                */
/* ------------------------------------------------------------------ */

/* RFromAngle (CThmScan.cpp:329-366) with identity Tdfm and Tptwk and no
 * pretweak list: rotate A's points by R(deg*PI/180) (TAffine::NUSetRot,
 * TAffine.cpp:116-122; Apply_R_Part :407-416), then CorrImagesS with
 * BigEnough(olap1D, reqArea) and EnoughPoints(reqArea), mincor 0. */
int xo_rfromangle(double deg, const xc_point *ap, const double *av, int na,
                  const xc_point *bp, const double *bv, int nb,
                  int olap1D, long reqArea, int Ox, int Oy, int Rx, int Ry,
                  double *X, double *Y, double *Rv, xc_result *res)
{
    double c = cos(deg * REF_PI / 180.0), s = sin(deg * REF_PI / 180.0);
    double t0 = c, t1 = -s, t3 = s, t4 = c;
    xc_point *pts = (xc_point *)malloc(sizeof(xc_point) * na);
    for (int i = 0; i < na; ++i) {
        double x = ap[i].x * t0 + ap[i].y * t1;
        double y = ap[i].x * t3 + ap[i].y * t4;
        pts[i].x = x;
        pts[i].y = y;
    }
    xc_rgn_rule g;
    memset(&g, 0, sizeof g);
    g.enabled = 1;
    g.min_w = g.min_h = olap1D;
    g.min_area = reqArea;
    xc_cnt_rule k;
    memset(&k, 0, sizeof k);
    k.enabled = 1;
    k.k1 = k.k2 = reqArea;
    int st = xo_corr_images(XC_MODE_S, pts, av, na, bp, bv, nb, &g, &k, 0.0, 0.0,
                            Ox, Oy, Rx, Ry, X, Y, Rv, res, NULL, NULL, NULL, 0);
    free(pts);
    return st;
}

/* DecimateVector (Maths.cpp:1315-1364): bin points (corner at 0,0) into
 * lbin x lbin cells and average.  p/v are updated in place; returns the
 * new count. */
int xo_decimate_vector(xc_point *p, double *v, int np, int w, int h, int lbin)
{
    w = (w + lbin - 1) / lbin;
    h = (h + lbin - 1) / lbin;
    long nr = (long)w * h;
    double *vr = (double *)calloc(nr, sizeof(double));
    uint8_t *cr = (uint8_t *)calloc(nr, 1);
    for (int ip = 0; ip < np; ++ip) {
        int x = (int)(p[ip].x) / lbin, y = (int)(p[ip].y) / lbin;
        long ir = x + (long)w * y;
        vr[ir] += v[ip];
        ++cr[ir];
    }
    int k = 0;
    for (long ir = 0; ir < nr; ++ir) {
        int cnt = cr[ir];
        if (cnt) {
            int y = (int)(ir / w), x = (int)(ir - (long)w * y);
            v[k] = vr[ir] / cnt;
            p[k].x = x;
            p[k].y = y;
            ++k;
        }
    }
    free(vr);
    free(cr);
    return k;
}

/* CThmUtil::GetOlapBoxes (1_DMesh/CThmUtil.cpp:352-373): a's origin (tx, ty)
 * in b's system, scaled by xyconf / scl and truncated; BoxesFromShifts of two
 * ws x hs pictures.  Returns 1 when the boxes are too thin for Crop (:404-411). */
static int olap_boxes(ibox *Ba, ibox *Bb, int ws, int hs, double tx, double ty, double xyconf, double scl,
                      int olap1d)
{
    int dx = (int)(xyconf * tx / scl), dy = (int)(xyconf * ty / scl);
    boxes_from_shifts(Ba, Bb, ws, hs, ws, hs, dx, dy);
    int ow = Ba->R - Ba->L + 1, oh = Ba->T - Ba->B + 1, min1d = olap1d > 8 ? olap1d : 8;
    return ow < min1d || oh < min1d;
}

/* SubI_ThesePoints (CThmUtil.cpp:163-189): values v[int(x) + ws int(y)],
 * points relative to their bounding box's corner. */
static void subi_these_points(const double *v, int ws, const xc_point *p, int n, xc_point *op, double *ov,
                              xc_subimage *S)
{
    ibox B;
    bbox_from_points(&B, p, n);
    for (int i = 0; i < n; ++i) {
        ov[i] = v[(int)p[i].x + ws * (int)p[i].y];
        op[i].x = p[i].x - B.L;
        op[i].y = p[i].y - B.B;
    }
    S->Ox = B.L;
    S->Oy = B.B;
    S->w = B.R - B.L + 1;
    S->h = B.T - B.B + 1;
    S->n = n;
}

/* SubI_ThisBox (CThmUtil.cpp:195-235): the points inside the box (in list
 * order); 0 (the reference's false) when olap2d or fewer remain. */
static int subi_this_box(const double *v, int ws, const xc_point *p, int n, const ibox *Bx, long olap2d,
                         xc_point *op, double *ov, xc_subimage *S)
{
    xc_point *kept = (xc_point *)malloc(sizeof(xc_point) * (n > 0 ? n : 1));
    int k = 0;
    for (int i = 0; i < n; ++i)
        if (p[i].x >= Bx->L && p[i].x <= Bx->R && p[i].y >= Bx->B && p[i].y <= Bx->T) kept[k++] = p[i];
    int ok = k > olap2d;
    if (ok) subi_these_points(v, ws, kept, k, op, ov, S);
    free(kept);
    return ok;
}

/* CThmUtil::Crop (CThmUtil.cpp:385-430) with Olap_WholeImage (:243-258).
 * Returns XC_OK (*whole = 1 for whole images) or XC_TOO_SMALL (false). */
int xo_crop(const double *av, const double *bv, int ws, int hs, const xc_point *apts, int na, const xc_point *bpts,
            int nb, double tx, double ty, double xyconf, double scl, int olap1d, long olap2d, xc_point *ap,
            double *avo, xc_subimage *a, xc_point *bp, double *bvo, xc_subimage *b, int *whole)
{
    ibox Ba, Bb;
    if (xyconf == 0.0 || olap_boxes(&Ba, &Bb, ws, hs, tx, ty, xyconf, scl, olap1d)) {
        *whole = 1;
        subi_these_points(av, ws, apts, na, ap, avo, a);
        subi_these_points(bv, ws, bpts, nb, bp, bvo, b);
        return XC_OK;
    }
    *whole = 0;
    if (!subi_this_box(av, ws, apts, na, &Ba, olap2d, ap, avo, a) ||
        !subi_this_box(bv, ws, bpts, nb, &Bb, olap2d, bp, bvo, b))
        return XC_TOO_SMALL;
    return XC_OK;
}

/* CThmUtil::Crop_NoCR (CThmUtil.cpp:437-464) with Olap_TheseBoxes_NoCR
 * (:262-312), Olap_WholeImage_NoCR (:318-346) and MakeZeroBasedPoints
 * (CPoint.cpp:82-95). */
int xo_crop_nocr(const double *av, const double *bv, int ws, int hs, double tx, double ty, double xyconf, double scl,
                 int olap1d, double *avo, double *bvo, xc_point *p, xc_subimage *a, xc_subimage *b, int *whole)
{
    ibox Ba, Bb;
    if (xyconf == 0.0 || olap_boxes(&Ba, &Bb, ws, hs, tx, ty, xyconf, scl, olap1d)) {
        Ba.L = Bb.L = 0;
        Ba.R = Bb.R = ws - 1;
        Ba.B = Bb.B = 0;
        Ba.T = Bb.T = hs - 1;
        *whole = 1;
    } else {
        *whole = 0;
    }
    int ow = Ba.R - Ba.L + 1, oh = Ba.T - Ba.B + 1, np = ow * oh;
    for (int i = 0; i < np; ++i) {
        int y = i / ow, x = i - ow * y;
        avo[i] = av[Ba.L + x + ws * (Ba.B + y)];
        bvo[i] = bv[Bb.L + x + ws * (Bb.B + y)];
        if (p) {
            p[i].x = x;
            p[i].y = y;
        }
    }
    a->Ox = Ba.L;
    a->Oy = Ba.B;
    b->Ox = Bb.L;
    b->Oy = Bb.B;
    a->w = b->w = ow;
    a->h = b->h = oh;
    a->n = b->n = np;
    return XC_OK;
}

/* ------------------------------------------------------------------ */
/* CThmScan angle search                                               */
/* ------------------------------------------------------------------ */

/* LineFit (Maths.cpp:457-507): lincor = r / sqrt(varX varY) of the least-
 * squares line through (x[i], y[i]), i < n (NaN or Inf for a zero variance). */
double xo_linefit_lincor(const double *x, const double *y, int n)
{
    double r, varX, varY, nn = 0.0, sumX = 0.0, sumY = 0.0, sumXX = 0.0, sumYY = 0.0, sumXY = 0.0;
    for (int i = 0; i < n; ++i) {
        double X = x[i], Y = y[i];
        sumX += X;
        sumY += Y;
        sumXX += X * X;
        sumYY += Y * Y;
        sumXY += X * Y;
        ++nn;
    }
    r = nn * sumXY - sumX * sumY;
    varX = nn * sumXX - sumX * sumX;
    varY = nn * sumYY - sumY * sumY;
    return r / sqrt(varX * varY);
}

/* ROUND (Maths.h:19). */
static double round_ref(double x) { return (double)(x >= 0.0 ? (int)(x + 0.5) : (int)(x - 0.5)); }

/* The angle lists of AngleScanMaxR / AngleScanConstXY (CThmScan.cpp:444-470,
 * 543-544).  Returns the count; a (may be NULL) receives the angles. */
int xo_scan_angles(int mode, double center, double hlfwid, double step, int favor, double *a_out, int cap)
{
    int n = 0;
#define PUSH(v) do { if (a_out && n < cap) a_out[n] = (v); ++n; } while (0)
    if (mode == XC_SCAN_MAXR) {
        PUSH(center);
        if (favor) {
            double cwide = 2.0 * step, cstep = step / 3.0;
            for (double a = center - cwide; a <= center + cwide; a += cstep) PUSH(a);
            if (hlfwid >= 180) {
                double opc = center + 180;
                for (double a = opc - cwide; a <= opc + cwide; a += cstep) PUSH(a);
            }
        }
    }
    for (double a = center - hlfwid; a <= center + hlfwid; a += step) PUSH(a);
#undef PUSH
    return n;
}

/* The selections of AngleScanMaxR (:476-487) and AngleScanConstXY
 * (:554-613) over evaluated records vC[nC].  ConstXY's sort by decreasing R
 * keeps exact ties in index order (std::sort leaves them unspecified). */
void xo_scan_select(int mode, const xc_angle_rec *vC, int nC, xc_angle_rec *best)
{
    if (mode == XC_SCAN_MAXR) {
        *best = vC[0];
        for (int ic = 1; ic < nC; ++ic)
            if (vC[ic].R > best->R) *best = vC[ic];
        return;
    }
    best->X = best->Y = best->A = best->R = 0.0;
    int *order = (int *)malloc(sizeof(int) * (nC ? nC : 1));
    for (int i = 0; i < nC; ++i) order[i] = i;
    for (int i = 1; i < nC; ++i) {  /* stable insertion sort by decreasing R */
        int v = order[i], j = i - 1;
        while (j >= 0 && vC[order[j]].R < vC[v].R) { order[j + 1] = order[j]; --j; }
        order[j + 1] = v;
    }
    const double rthresh = 0.7;
    const int m = 3, M = 2 * m + 1;
    for (int i = 0; i < nC; ++i) {
        int ic = order[i];
        if (ic < m) continue;
        if (nC - 1 - ic < m) continue;
        double A[7], X[7], Y[7], lincor;
        int n = 0;
        for (int j = ic - m; j <= ic + m; ++j) {
            A[n] = vC[j].A;
            X[n] = round_ref(vC[j].X);
            Y[n] = round_ref(vC[j].Y);
            ++n;
        }
        (void)M;
        lincor = xo_linefit_lincor(A, X, n);
        if (isfinite(lincor) && fabs(lincor) < rthresh) continue;
        lincor = xo_linefit_lincor(A, Y, n);
        if (isfinite(lincor) && fabs(lincor) < rthresh) continue;
        *best = vC[ic];
        break;
    }
    free(order);
}

static void rfa_rec(double deg, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                    int nb, int olap1D, long reqArea, int Ox, int Oy, int Rx, int Ry, xc_angle_rec *C)
{
    C->A = deg;
    xo_rfromangle(deg, ap, av, na, bp, bv, nb, olap1D, reqArea, Ox, Oy, Rx, Ry, &C->X, &C->Y, &C->R, NULL);
}

/* AngleScanMaxR / AngleScanConstXY (CThmScan.cpp:431-621) over RFromAngle. */
int xo_angle_scan(int mode, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                  int nb, double center, double hlfwid, double step, int favor, int olap1D, long reqArea, int Ox,
                  int Oy, int Rx, int Ry, xc_angle_rec *best, xc_angle_rec *recs, int cap, int *nrec)
{
    int nC = xo_scan_angles(mode, center, hlfwid, step, favor, NULL, 0);
    double *a = (double *)malloc(sizeof(double) * (nC ? nC : 1));
    xc_angle_rec *vC = (xc_angle_rec *)malloc(sizeof(xc_angle_rec) * (nC ? nC : 1));
    xo_scan_angles(mode, center, hlfwid, step, favor, a, nC);
    for (int i = 0; i < nC; ++i) rfa_rec(a[i], ap, av, na, bp, bv, nb, olap1D, reqArea, Ox, Oy, Rx, Ry, &vC[i]);
    if (nC > 0) xo_scan_select(mode, vC, nC, best);
    else best->X = best->Y = best->A = best->R = 0.0;
    if (recs) memcpy(recs, vC, sizeof(xc_angle_rec) * (nC < cap ? nC : cap));
    if (nrec) *nrec = nC;
    free(a);
    free(vC);
    return XC_OK;
}

/* PeakHunt (CThmScan.cpp:676-732). */
int xo_peak_hunt(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv, int nb,
                 double hlfwid, int olap1D, long reqArea, int Ox, int Oy, int Rx, int Ry, xc_angle_rec *best, int *kout)
{
    xc_angle_rec C, B0 = *best;
    double L = best->A - hlfwid, R = best->A + hlfwid, M;
    int k = 1;
    rfa_rec(M = (L + R) / 2.0, ap, av, na, bp, bv, nb, olap1D, reqArea, Ox, Oy, Rx, Ry, best);
    while (R - L > 0.0001) {
        double a;
        ++k;
        rfa_rec(a = (L + M) / 2.0, ap, av, na, bp, bv, nb, olap1D, reqArea, Ox, Oy, Rx, Ry, &C);
        if (C.R >= best->R) {
            R = M;
            *best = C;
            M = a;
        } else
            L = a;
        rfa_rec(a = (M + R) / 2.0, ap, av, na, bp, bv, nb, olap1D, reqArea, Ox, Oy, Rx, Ry, &C);
        if (C.R >= best->R) {
            L = M;
            *best = C;
            M = a;
        } else
            R = a;
    }
    if (B0.R > best->R) *best = B0;
    if (kout) *kout = k;
    return XC_OK;
}

/* ---- Pretweaks ---------------------------------------------------- */

typedef struct { double t[6]; } xo_aff;

/* TAffine::operator* (TAffine.cpp:220-232). */
static xo_aff aff_mul(xo_aff L, xo_aff R)
{
    xo_aff r;
    r.t[0] = L.t[0] * R.t[0] + L.t[1] * R.t[3];
    r.t[1] = L.t[0] * R.t[1] + L.t[1] * R.t[4];
    r.t[2] = L.t[0] * R.t[2] + L.t[1] * R.t[5] + L.t[2];
    r.t[3] = L.t[3] * R.t[0] + L.t[4] * R.t[3];
    r.t[4] = L.t[3] * R.t[1] + L.t[4] * R.t[4];
    r.t[5] = L.t[3] * R.t[2] + L.t[4] * R.t[5] + L.t[5];
    return r;
}

/* TAffine::NUSelect (TAffine.cpp:76-135). */
static xo_aff nu_select(int sel, double a)
{
    xo_aff r = {{1, 0, 0, 0, 1, 0}};
    switch (sel) {
    case 1: r.t[0] = a; break;
    case 2: r.t[4] = a; break;
    case 3: r.t[1] = a; break;
    case 4: r.t[3] = a; break;
    case 5: { double c = cos(a), s = sin(a); r.t[0] = c; r.t[1] = -s; r.t[3] = s; r.t[4] = c; break; }
    default: r.t[0] = a; r.t[4] = a; break;
    }
    return r;
}

static xo_aff aff_of(const double *t)
{
    xo_aff r = {{1, 0, 0, 0, 1, 0}};
    if (t) memcpy(r.t, t, sizeof r.t);
    return r;
}

/* RFromAngle (CThmScan.cpp:329-366) with C.T = R * (Tdfm * T) applied to A's points
 * (Apply_R_Part, TAffine.cpp:407-423), then CorrImagesS with BigEnough/EnoughPoints. */
int xo_rfromangle_affine(double deg, const xc_point *ap, const double *av, int na, const xc_point *bp,
                         const double *bv, int nb, const double *Tdfm, const double *Ttw, int olap1D, long reqArea,
                         int Ox, int Oy, int Rx, int Ry, double *X, double *Y, double *Rv)
{
    xo_aff T = aff_mul(nu_select(5, deg * REF_PI / 180.0), aff_mul(aff_of(Tdfm), aff_of(Ttw)));
    xc_point *pts = (xc_point *)malloc(sizeof(xc_point) * na);
    for (int i = 0; i < na; ++i) {
        double x = ap[i].x * T.t[0] + ap[i].y * T.t[1];
        double y = ap[i].x * T.t[3] + ap[i].y * T.t[4];
        pts[i].x = x;
        pts[i].y = y;
    }
    xc_rgn_rule g;
    memset(&g, 0, sizeof g);
    g.enabled = 1;
    g.min_w = g.min_h = olap1D;
    g.min_area = reqArea;
    xc_cnt_rule k;
    memset(&k, 0, sizeof k);
    k.enabled = 1;
    k.k1 = k.k2 = reqArea;
    int st = xo_corr_images(XC_MODE_S, pts, av, na, bp, bv, nb, &g, &k, 0.0, 0.0, Ox, Oy, Rx, Ry, X, Y, Rv, NULL,
                            NULL, NULL, NULL, 0);
    free(pts);
    return st;
}

/* CThmScan::Pretweaks (CThmScan.cpp:205-323) with PTWInterp (:119-160) and
 * NewXFromParabola (:48-61). */
int xo_pretweaks(double deg, const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                 int nb, const double *Tdfm, double *Tptwk, double *bestR, int olap1D, long reqArea, int Ox, int Oy,
                 int Rx, int Ry, int *changed, int *n_evals)
{
    xo_aff tw = aff_of(Tptwk);
    int vused[5] = {0, 0, 0, 0, 0}, any = 0, evals = 0;
    const double astep = 0.02;
    const int nstep = 5;
    double bR = *bestR, X, Y, R;
#define EVAL(sel, a, out)                                                                                     \
    do {                                                                                                      \
        xo_aff W = aff_mul(nu_select((sel), (a)), tw);                                                        \
        ++evals;                                                                                              \
        xo_rfromangle_affine(deg, ap, av, na, bp, bv, nb, Tdfm, W.t, olap1D, reqArea, Ox, Oy, Rx, Ry, &X, &Y, \
                             &R);                                                                             \
        (out) = R;                                                                                            \
    } while (0)
    for (int itype = 0; itype < 5; ++itype) {
        int vsel[5 * 11];  /* 5 types x (2 nstep + 1) magnitudes */
        double va[5 * 11];
        int n = 0;
        for (int sel = 0; sel < 5; ++sel) {
            if (vused[sel]) continue;
            double abase = (sel >= 0 && sel <= 2) ? 1.0 : 0.0;
            for (int j = -nstep; j <= nstep; ++j) {
                vsel[n] = sel;
                va[n] = abase + j * astep;
                ++n;
            }
        }
        double rbest = 0;
        int ibest = -1;
        for (int ic = 0; ic < n; ++ic) {
            double r;
            EVAL(vsel[ic], va[ic], r);
            if (r > rbest) {
                rbest = r;
                ibest = ic;
            }
        }
        if (ibest == -1) break;
        int sel = vsel[ibest];
        double x1 = va[ibest], y1 = rbest, d = astep, y0, y2, ynew, xnew;
        EVAL(sel, x1 - d, y0);
        EVAL(sel, x1 + d, y2);
        double Dp = d * (y0 - y2) / (2 * (y0 + y2 - y1 - y1));
        xnew = x1;
        if (fabs(Dp) <= 2 * d) xnew += Dp;
        EVAL(sel, xnew, ynew);
        if (ynew < y1) {
            xnew = x1;
            ynew = y1;
        }
        rbest = ynew;
        if (rbest > bR) {
            tw = aff_mul(nu_select(sel, xnew), tw);
            vused[sel] = 1;
            bR = rbest;
            any = 1;
        } else
            break;
    }
#undef EVAL
    memcpy(Tptwk, tw.t, sizeof tw.t);
    *bestR = bR;
    if (changed) *changed = any;
    if (n_evals) *n_evals = evals;
    return XC_OK;
}

/* CThmScan::FinishAtFullRes (CThmScan.cpp:810-842). */
int xo_finish_at_full_res(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv, int nb,
                          const double *Tdfm, const double *Tptwk, int olap1D, long reqArea, xc_angle_rec *best,
                          int *used)
{
    xc_angle_rec b0 = *best, C;
    C.A = b0.A;
    xo_rfromangle_affine(b0.A, ap, av, na, bp, bv, nb, Tdfm, Tptwk, olap1D, reqArea, (int)b0.X, (int)b0.Y, 80, 80,
                         &C.X, &C.Y, &C.R);
    int ok = (fabs(C.X - b0.X) <= 20) && (fabs(C.Y - b0.Y) <= 20);
    if (ok) {
        *best = C;
        best->R = b0.R;
    } else
        *best = b0;
    if (used) *used = ok;
    return XC_OK;
}

/* ---- CThmScan's matching sequence (CThmScan.cpp:627-802) over an xc_thm_params block ---- */

typedef struct {
    const xc_point *ap, *bp;
    const double *av, *bv;
    int na, nb;
    xc_thm_params *p;
} xo_thm;

/* The search oval at angle deg (:351-355): newAngProc's when set, else the disc. */
static void thm_oval(const xc_thm_params *p, double deg, int *ox, int *oy, int *rx, int *ry)
{
    if (p->ang_proc) p->ang_proc(ox, oy, rx, ry, deg, p->ang_arg);
    else { *ox = p->Ox; *oy = p->Oy; *rx = p->Rx; *ry = p->Ry; }
}

/* RFromAngle (:329-366) with the block's Tdfm, Tptwk and oval. */
static void thm_rfa(xo_thm *t, double deg, xc_angle_rec *C)
{
    xc_thm_params *p = t->p;
    int ox, oy, rx, ry;
    thm_oval(p, deg, &ox, &oy, &rx, &ry);
    C->A = deg;
    ++p->n_evals;
    xo_rfromangle_affine(deg, t->ap, t->av, t->na, t->bp, t->bv, t->nb, p->Tdfm, p->Tptwk, p->olap1D, p->req_area,
                         ox, oy, rx, ry, &C->X, &C->Y, &C->R);
}

/* AngleScanSel (:627-641). */
static double thm_scan_sel(xo_thm *t, xc_angle_rec *best, double center, double hlfwid, double step, int favor)
{
    int mode = t->p->const_xy ? XC_SCAN_CONSTXY : XC_SCAN_MAXR;
    int nC = xo_scan_angles(mode, center, hlfwid, step, favor, NULL, 0);
    double *a = (double *)malloc(sizeof(double) * (nC ? nC : 1));
    xc_angle_rec *vC = (xc_angle_rec *)malloc(sizeof(xc_angle_rec) * (nC ? nC : 1));
    xo_scan_angles(mode, center, hlfwid, step, favor, a, nC);
    for (int i = 0; i < nC; ++i) thm_rfa(t, a[i], &vC[i]);
    if (nC > 0) xo_scan_select(mode, vC, nC, best);
    else best->X = best->Y = best->A = best->R = 0.0;
    free(a);
    free(vC);
    return best->R;
}

/* Pretweaks (:205-323) folding into the block's Tptwk. */
static int thm_pretweaks(xo_thm *t, double bestR, double deg)
{
    xc_thm_params *p = t->p;
    int changed = 0, ev = 0, ox, oy, rx, ry;
    thm_oval(p, deg, &ox, &oy, &rx, &ry);  /* every candidate of a round is evaluated at deg */
    xo_pretweaks(deg, t->ap, t->av, t->na, t->bp, t->bv, t->nb, p->Tdfm, p->Tptwk, &bestR, p->olap1D, p->req_area,
                 ox, oy, rx, ry, &changed, &ev);
    p->n_evals += ev;
    return changed;
}

/* AngleScanWithTweaks (:647-672). */
static double thm_with_tweaks(xo_thm *t, xc_angle_rec *best, double center, double hlfwid, double step)
{
    if (thm_scan_sel(t, best, center, hlfwid, step, 1) < t->p->rthresh && t->p->pretweak) {
        if (thm_pretweaks(t, best->R, best->R > 0.0 ? best->A : center))
            thm_scan_sel(t, best, center, hlfwid, step, 1);
    }
    return best->R;
}

/* PeakHunt (:676-732). */
static double thm_hunt(xo_thm *t, xc_angle_rec *best, double hlfwid)
{
    xc_angle_rec C, B0 = *best;
    double L = best->A - hlfwid, R = best->A + hlfwid, M;
    thm_rfa(t, M = (L + R) / 2.0, best);
    while (R - L > 0.0001) {
        double a;
        thm_rfa(t, a = (L + M) / 2.0, &C);
        if (C.R >= best->R) {
            R = M;
            *best = C;
            M = a;
        } else
            L = a;
        thm_rfa(t, a = (M + R) / 2.0, &C);
        if (C.R >= best->R) {
            L = M;
            *best = C;
            M = a;
        } else
            R = a;
    }
    if (B0.R > best->R) *best = B0;
    return best->R;
}

int xo_thm_scan_with_tweaks(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                            int nb, xc_thm_params *p, double center, double hlfwid, double step, xc_angle_rec *best)
{
    xo_thm t = {ap, bp, av, bv, na, nb, p};
    thm_with_tweaks(&t, best, center, hlfwid, step);
    return XC_OK;
}

/* UsePriorAngles (:734-758). */
int xo_thm_use_prior_angles(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                            int nb, xc_thm_params *p, double ang0, double hfangpr, xc_angle_rec *best, int *ok)
{
    xo_thm t = {ap, bp, av, bv, na, nb, p};
    *ok = !(thm_with_tweaks(&t, best, ang0, hfangpr, 0.1) < p->rthresh || thm_hunt(&t, best, 0.1 * 1.5) < p->rthresh);
    return XC_OK;
}

/* DenovoBestAngle (:764-790). */
int xo_thm_denovo_best_angle(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv,
                             int nb, xc_thm_params *p, double ang0, double hfangdn, double step, xc_angle_rec *best,
                             int *ok)
{
    xo_thm t = {ap, bp, av, bv, na, nb, p};
    *ok = !(thm_with_tweaks(&t, best, ang0, hfangdn, step) < p->rthresh ||
            thm_scan_sel(&t, best, best->A, step * 2.0, step * 0.05, 0) < p->rthresh ||
            thm_hunt(&t, best, step * 0.05 * 1.5) < p->rthresh);
    return XC_OK;
}

/* PostTweaks (:796-802). */
int xo_thm_post_tweaks(const xc_point *ap, const double *av, int na, const xc_point *bp, const double *bv, int nb,
                       xc_thm_params *p, xc_angle_rec *best)
{
    xo_thm t = {ap, bp, av, bv, na, nb, p};
    p->Ox = (int)best->X;
    p->Oy = (int)best->Y;
    p->Rx = p->Ry = 40;
    thm_pretweaks(&t, best->R, best->A);
    thm_rfa(&t, best->A, best);
    return XC_OK;
}

/* ---- PixPair::Load's raster preparation (CPixPair.cpp:276-460) ---- */

/* LegPolyCreate (Maths.cpp:962-997): L[o * n + s], o = 0..ord (ord <= 6). */
void xo_leg_poly_create(double *L, int ord, int n)
{
    double dn = n, n_1 = dn - 1;
    for (int s = 0; s < n; ++s) {
        double x = (2.0 * s - n_1) / dn;
        double v[7];
        v[0] = 1.0;
        v[1] = x;
        v[2] = 1.5 * x * x - 0.5;
        v[3] = 2.5 * x * x * x - 1.5 * x;
        v[4] = 0.125 * (35.0 * x * x * x * x - 30.0 * x * x + 3.0);
        v[5] = 0.125 * (63.0 * x * x * x * x * x - 70.0 * x * x * x + 15.0 * x);
        v[6] = 0.0625 * (231.0 * x * x * x * x * x * x - 315.0 * x * x * x * x + 105.0 * x * x - 5.0);
        for (int o = 0; o <= ord; ++o) L[(long)o * n + s] = v[o];
    }
}

/* LegPolyFlatten for 8-bit rasters (Maths.cpp:1006-1086): remove, term by term
 * in (xo, yo) order, the projection onto each Legendre product but (0,0), then
 * Normalize.  order 0..6. */
int xo_leg_poly_flatten(double *vals, const uint8_t *src, int w, int h, int order)
{
    long n = (long)w * h;
    if (order < 0 || order > 6) return XC_E_ARG;
    for (long i = 0; i < n; ++i) vals[i] = src[i];
    if (order > 0) {
        double *xp = (double *)malloc(sizeof(double) * (order + 1) * w);
        double *yp = (double *)malloc(sizeof(double) * (order + 1) * h);
        xo_leg_poly_create(xp, order, w);
        xo_leg_poly_create(yp, order, h);
        for (int xo = 0; xo <= order; ++xo) {
            const double *XP = xp + (long)xo * w;
            for (int yo = 0; yo <= order; ++yo) {
                if (xo + yo == 0) continue;
                const double *YP = yp + (long)yo * h;
                double coef = 0.0, norm = 0.0;
                for (long i = 0; i < n; ++i) {
                    int y = (int)(i / w), x = (int)(i - (long)w * y);
                    double F = XP[x] * YP[y];
                    coef += F * vals[i];
                    norm += F * F;
                }
                coef /= norm;
                for (long i = 0; i < n; ++i) {
                    int y = (int)(i / w), x = (int)(i - (long)w * y);
                    vals[i] -= coef * XP[x] * YP[y];
                }
            }
        }
        free(xp);
        free(yp);
    }
    xo_normalize(vals, (int)n);
    return XC_OK;
}

/* MakeDoGKernel (CPixPair.cpp:36-80): D x D, D = 6 r2 + 1; returns D. */
int xo_make_dog_kernel(double *DoG, int r1, int r2)
{
    int d = 3 * r2, D = 2 * d + 1;
    double sum1 = 0.0, sum2 = 0.0;
    for (int y = -d; y <= d; ++y)
        for (int x = -d; x <= d; ++x) {
            double R = x * x + y * y;
            sum1 += exp(-(R / (2 * r1 * r1)));
            sum2 += exp(-(R / (2 * r2 * r2)));
        }
    if (DoG)
        for (int y = -d; y <= d; ++y)
            for (int x = -d; x <= d; ++x) {
                double R = x * x + y * y;
                double v1 = exp(-(R / (2 * r1 * r1)));
                double v2 = exp(-(R / (2 * r2 * r2)));
                DoG[x + d + D * (y + d)] = v1 / sum1 - v2 / sum2;
            }
    return D;
}

/* PixPair::Downsample (CPixPair.cpp:90-113): scl x scl box means. */
void xo_downsample(double *dst, const double *src, int wf, int scl, int ws, int hs)
{
    int n = scl * scl;
    for (int iy = 0; iy < hs; ++iy)
        for (int ix = 0; ix < ws; ++ix) {
            double sum = 0.0;
            for (int dy = 0; dy < scl; ++dy)
                for (int dx = 0; dx < scl; ++dx) sum += src[ix * scl + dx + (long)wf * (iy * scl + dy)];
            dst[ix + (long)ws * iy] = sum / n;
        }
}

/* PixPair::Load's raster stage for one picture (CPixPair.cpp:276-298, 371-389,
 * 427-460): flatten; with dog, Convolve by the DoG kernel (symmetric,
 * pre-normalised) and Normalize; halve while w or h > 2048 and box-average.
 * Outputs may be NULL; the downsampled ones are written only when scl > 1.
 * XC_E_ARG for order outside 0..6, r1/r2 < 1, or sizes not multiples of scl. */
int xo_pixpair_prep(const uint8_t *ras, int w, int h, int order, int dog, int r1, int r2, double *vf, double *vfflt,
                    double *vs, double *vsflt, int *ws_out, int *hs_out, int *scl_out)
{
    int ws = w, hs = h, scl = 1;
    long n = (long)w * h;
    if (w < 1 || h < 1 || order < 0 || order > 6 || (dog && (r1 < 1 || r2 < 1))) return XC_E_ARG;
    while (ws > 2048 || hs > 2048) {
        ws /= 2;
        hs /= 2;
        scl *= 2;
    }
    if (ws * scl != w || hs * scl != h) return XC_E_ARG;
    double *f = (double *)malloc(sizeof(double) * n);
    xo_leg_poly_flatten(f, ras, w, h, order);
    double *g = NULL;
    if (dog) {
        int D = xo_make_dog_kernel(NULL, r1, r2);
        double *K = (double *)malloc(sizeof(double) * D * D);
        xo_make_dog_kernel(K, r1, r2);
        g = (double *)malloc(sizeof(double) * n);
        xo_convolve(g, f, w, h, K, D, D, 1, 1);
        xo_normalize(g, (int)n);
        free(K);
    }
    if (vf) memcpy(vf, f, sizeof(double) * n);
    if (vfflt && g) memcpy(vfflt, g, sizeof(double) * n);
    if (scl > 1) {
        if (vs) xo_downsample(vs, f, w, scl, ws, hs);
        if (vsflt && g) xo_downsample(vsflt, g, w, scl, ws, hs);
    }
    free(f);
    free(g);
    if (ws_out) *ws_out = ws;
    if (hs_out) *hs_out = hs;
    if (scl_out) *scl_out = scl;
    return XC_OK;
}

/* ---- FourierMatch (Metrics.cpp:517-658) ---- */

typedef struct { double L, R, B, T; } xo_dbox;

/* BBoxFromPoints, DBox form (Geometry.cpp:109-137). */
static void dbox_from_points(xo_dbox *B, const xc_point *p, int n)
{
    B->L = B->R = p[0].x;
    B->B = B->T = p[0].y;
    for (int i = 1; i < n; ++i) {
        double t;
        if ((t = p[i].x) < B->L) B->L = t;
        else if (t > B->R) B->R = t;
        if ((t = p[i].y) < B->B) B->B = t;
        else if (t > B->T) B->T = t;
    }
    B->L = floor(B->L);
    B->R = ceil(B->R);
    B->B = floor(B->B);
    B->T = ceil(B->T);
}

/* NUSetRot + Apply_R_Part (TAffine.cpp): x' = c x - s y, y' = s x + c y. */
static void rotate_points(xc_point *q, const xc_point *p, int n, double rad)
{
    double c = cos(rad), s = sin(rad);
    for (int i = 0; i < n; ++i) {
        double x = p[i].x, y = p[i].y;
        q[i].x = c * x + -s * y;
        q[i].y = s * x + c * y;
    }
}

/* TightestBBox (Geometry.cpp:187-267): the whole-degree angle in -45..45 whose
 * rotation of the region's row outline has the least bounding-box area. */
int xo_tightest_bbox(const xc_point *pts, int np)
{
    xo_dbox B;
    dbox_from_points(&B, pts, np);
    int ny = (int)(B.T - B.B) + 1;
    double *minx = (double *)malloc(sizeof(double) * ny), *maxx = (double *)malloc(sizeof(double) * ny);
    for (int i = 0; i < ny; ++i) {
        minx[i] = B.R + 1;
        maxx[i] = B.L - 1;
    }
    for (int i = 0; i < np; ++i) {
        int iy = (int)floor(pts[i].y - B.B);
        minx[iy] = fmin(minx[iy], pts[i].x);
        maxx[iy] = fmax(maxx[iy], pts[i].x);
    }
    xc_point *outline = (xc_point *)malloc(sizeof(xc_point) * 2 * ny);
    int no = 0;
    for (int iy = 0; iy < ny; ++iy)
        if (minx[iy] <= maxx[iy]) {
            outline[no].x = minx[iy];
            outline[no++].y = iy + B.B;
            outline[no].x = maxx[iy];
            outline[no++].y = iy + B.B;
        }
    const xc_point *set = outline;
    int ns = no;
    if (no >= np) {
        set = pts;
        ns = np;
    }
    double best_area = (B.R - B.L) * (B.T - B.B);
    int best_angle = 0;
    xc_point *P = (xc_point *)malloc(sizeof(xc_point) * (ns ? ns : 1));
    for (int angle = -45; angle <= 45; ++angle) {
        if (angle == 0) continue;
        xo_dbox box;
        rotate_points(P, set, ns, angle * REF_PI / 180);
        dbox_from_points(&box, P, ns);
        double area = (box.R - box.L) * (box.T - box.B);
        if (area < best_area) {
            best_area = area;
            best_angle = angle;
        }
    }
    free(P);
    free(outline);
    free(minx);
    free(maxx);
    return best_angle;
}

/* FFT_r2c_lookup (Metrics.cpp:420-435), the reference's form. */
static void r2c_lookup(const cd *c, int Nx, int Ny, int x, int y, double *re, double *im)
{
    int M = Nx / 2 + 1;
    if (x < 0) x += Nx;
    if (y < 0) y += Ny;
    if (x > Nx / 2) {
        *re = c[(Nx - x) + (long)M * y].re;
        *im = -c[(Nx - x) + (long)M * y].im;
        return;
    }
    *re = c[x + (long)M * y].re;
    *im = c[x + (long)M * y].im;
}

/* FourierMatch (Metrics.cpp:517-658) with MeanSqrDiff (:30-52),
 * SmallestFootprint (:68-89) and MakeMetricImagesFFT (:102-193). */
int xo_fourier_match(const xc_point *pts, const double *av, const double *bv, int n, int wvlen, double *dot_out,
                     xc_fourier_qa *qa)
{
    xc_fourier_qa q;
    memset(&q, 0, sizeof q);
    *dot_out = 0.0;
    if (!n) {
        if (qa) *qa = q;
        return XC_OK;
    }
    double sum = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = av[i] - bv[i];
        sum += d * d;
    }
    q.msd = sum / n;
    q.deg = xo_tightest_bbox(pts, n);
    xc_point *P = (xc_point *)malloc(sizeof(xc_point) * n);
    if (q.deg) rotate_points(P, pts, n, q.deg * REF_PI / 180);
    else memcpy(P, pts, sizeof(xc_point) * n);
    ibox B;
    bbox_from_points(&B, P, n);
    int Nx = xo_ceil_pow2(B.R - B.L + 1), Ny = xo_ceil_pow2(B.T - B.B + 1);
    long N2 = (long)Nx * Ny, M = Nx / 2 + 1;
    double *i1 = (double *)calloc(N2, sizeof(double)), *i2 = (double *)calloc(N2, sizeof(double));
    double *df = (double *)malloc(sizeof(double) * N2);
    for (int i = 0; i < n; ++i) {
        double x = P[i].x - B.L, y = P[i].y - B.B;
        distribute_pixel(x, y, av[i], i1, Nx, Ny);
        distribute_pixel(x, y, bv[i], i2, Nx, Ny);
    }
    for (long i = 0; i < N2; ++i) df[i] = i1[i] - i2[i];
    cd *f1 = (cd *)malloc(sizeof(cd) * M * Ny), *f2 = (cd *)malloc(sizeof(cd) * M * Ny);
    cd *fd = (cd *)malloc(sizeof(cd) * M * Ny);
    xo_fft2d_r2c(i1, Nx, Ny, (double *)f1);
    xo_fft2d_r2c(i2, Nx, Ny, (double *)f2);
    xo_fft2d_r2c(df, Nx, Ny, (double *)fd);
    double total = 0.0, dot = 0.0;
    int xlim = Nx / wvlen, ylim = Ny / wvlen;
    for (int x = -xlim; x <= xlim; ++x)
        for (int y = -ylim; y <= ylim; ++y) {
            double a, b, c, d;
            r2c_lookup(f1, Nx, Ny, x, y, &a, &b);
            r2c_lookup(f2, Nx, Ny, x, y, &c, &d);
            total += sqrt((a * a + b * b) * (c * c + d * d));
            dot += a * c + b * d;
        }
    dot /= total;
    int maxf = (int)(sqrt((double)(Nx * Nx + Ny * Ny)) / 2.0) + 1;
    double *ps = (double *)calloc(maxf, sizeof(double));
    for (int x = 1; x < M; ++x) {
        double wavex = (double)Nx / x;
        for (int y = 1; y < Ny; ++y) {
            double wavey = (double)Ny / (y > Ny / 2 ? Ny - y : y);
            double wave = sqrt(wavex * wavex + wavey * wavey);
            int iwave = (int)wave;
            if (iwave > maxf - 1) iwave = maxf - 1;
            cd v = fd[x + M * y];
            ps[iwave] += v.re * v.re + v.im * v.im;
        }
    }
    double tot = 0.0, cum = 0.0;
    int i;
    for (i = 0; i < maxf; ++i) tot += ps[i];
    for (i = 0; cum < tot * 0.50 && i < maxf; ++i) cum += ps[i];
    q.dot = dot;
    q.energy = total;
    q.tot_power = tot;
    q.half_frac = cum / tot;
    q.half_index = i;
    q.maxf = maxf;
    q.nx = Nx;
    q.ny = Ny;
    *dot_out = dot;
    if (qa) *qa = q;
    free(P); free(i1); free(i2); free(df); free(f1); free(f2); free(fd); free(ps);
    return XC_OK;
}
