// aln_corr_compat.cpp -- the reference's correlation signatures
// (0_GEN/Correlation.h:16-220) over the libxcorr C-ABI.
#include "../../include/aln_corr_compat.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>

namespace {

// One device context per calling thread.
struct Ctx {
    xc_ctx *h = nullptr;
    ~Ctx()
    {
        if (h) xc_ctx_destroy(h);
    }
};
thread_local Ctx tls;

xc_ctx *ctx()
{
    if (!tls.h) {
        const char *d = getenv("ALN_XC_DEVICE");
        const int st = xc_ctx_create(d ? atoi(d) : 0, &tls.h);
        if (st != XC_OK) {
            fprintf(stderr, "aln_corr: no HIP device (status %d)\n", st);
            exit(42);
        }
    }
    return tls.h;
}

[[noreturn]] void die(FILE *flog, const char *what, int st)
{
    fprintf(flog ? flog : stderr, "aln_corr: %s failed: status %d (%s)\n", what, st, xc_last_error(tls.h));
    exit(44);
}

void bbox(const vector<Point> &p, int &L, int &R, int &B, int &T)
{
    // BBoxFromPoints (Geometry.cpp:78-106)
    double xlo, xhi, ylo, yhi;
    xlo = xhi = p[0].x;
    ylo = yhi = p[0].y;
    for (size_t i = 1; i < p.size(); ++i) {
        double t;
        if ((t = p[i].x) < xlo) xlo = t;
        else if (t > xhi) xhi = t;
        if ((t = p[i].y) < ylo) ylo = t;
        else if (t > yhi) yhi = t;
    }
    L = (int)floor(xlo);
    R = (int)ceil(xhi);
    B = (int)floor(ylo);
    T = (int)ceil(yhi);
}

// LegalRgn tabulated over every possible overlap olw in [1, tw], olh in [1, th].
struct RgnTable {
    xc_rgn_rule rule;
    std::vector<uint8_t> tab;
    RgnTable(EvalType fn, void *arg, int tw, int th)
    {
        memset(&rule, 0, sizeof rule);
        if (!fn) return;
        tab.resize((size_t)tw * th);
        for (int h = 1; h <= th; ++h)
            for (int w = 1; w <= tw; ++w) tab[(size_t)(h - 1) * tw + (w - 1)] = fn(w, h, arg) ? 1 : 0;
        rule.enabled = 1;
        rule.min_w = rule.min_h = 0;
        rule.min_area = -1;
        rule.table_w = tw;
        rule.table_h = th;
        rule.table = tab.data();
    }
};

// LegalCnt as c1 > k1 && c2 > k2, found by bisection and verified.
bool cnt_rule(EvalType fn, void *arg, long long m1, long long m2, xc_cnt_rule &out)
{
    memset(&out, 0, sizeof out);
    if (!fn) return true;
    auto first_true = [&](bool first, long long hi, long long other) -> long long {
        long long lo = 0, r = hi + 1;
        long long a = 0, b = hi;
        while (a <= b) {
            const long long mid = a + (b - a) / 2;
            const bool ok = first ? fn((int)mid, (int)other, arg) : fn((int)other, (int)mid, arg);
            if (ok) {
                r = mid;
                b = mid - 1;
            } else {
                a = mid + 1;
            }
        }
        (void)lo;
        return r;
    };
    const long long c1 = first_true(true, m1, m2), c2 = first_true(false, m2, m1);
    out.enabled = 1;
    out.k1 = c1 - 1;
    out.k2 = c2 - 1;
    // verify on probe points against the callback itself
    const long long p1[] = {0, 1, out.k1, out.k1 + 1, m1 / 2, m1};
    const long long p2[] = {0, 1, out.k2, out.k2 + 1, m2 / 2, m2};
    for (long long a : p1)
        for (long long b : p2) {
            if (a < 0 || b < 0 || a > m1 || b > m2) continue;
            if (fn((int)a, (int)b, arg) != (a > out.k1 && b > out.k2)) return false;
        }
    return true;
}

// The fft2 cache of the FFTW build (Correlation.cpp:977, Correlation_fft_fftw.cpp:50):
// when its size is not M, it is filled with FFT_2D of B's padded raster.
void refresh_cache(vector<CD> &fft2, const vector<Point> &ip2, const vector<double> &iv2, int L, int B, int Nx,
                   int Ny)
{
    const size_t M = (size_t)Ny * (Nx / 2 + 1);
    if (fft2.size() == M) return;
    vector<double> img((size_t)Nx * Ny, 0.0);
    for (size_t i = 0; i < ip2.size(); ++i) {
        // DistributePixel (Maths.h:236-276)
        const double x = ip2[i].x - L, y = ip2[i].y - B, v = iv2[i];
        int ix = (int)floor(x), iy = (int)floor(y);
        const double a = x - ix, b = y - iy;
        long long k = ix + (long long)Nx * iy;
        if (ix >= 0 && ix < Nx) {
            if (iy >= 0 && iy < Ny) img[k] += (1.0 - a) * (1.0 - b) * v;
            if (iy + 1 >= 0 && iy + 1 < Ny) img[k + Nx] += (1.0 - a) * b * v;
        }
        if (ix + 1 >= 0 && ix + 1 < Nx) {
            if (iy >= 0 && iy < Ny) img[k + 1] += a * (1.0 - b) * v;
            if (iy + 1 >= 0 && iy + 1 < Ny) img[k + 1 + Nx] += a * b * v;
        }
    }
    FFT_2D(fft2, img, Nx, Ny, false);
}

const xc_point *pts(const vector<Point> &p) { return reinterpret_cast<const xc_point *>(p.data()); }

}  // namespace

static_assert(sizeof(Point) == sizeof(xc_point), "Point must be two doubles");
static_assert(sizeof(CD) == 2 * sizeof(double), "CD must be two doubles");

int FFTSize(int n1, int n2) { return xc_fft_size(n1, n2); }
int FFTSizeSQR(int w1, int h1, int w2, int h2) { return xc_fft_size_sqr(w1, h1, w2, h2); }

int FFT_2D(vector<CD> &out, const vector<double> &in, int Nfast, int Nslow, bool cached, FILE *flog)
{
    const int M = Nslow * (Nfast / 2 + 1);
    if (cached && (int)out.size() == M) return M;
    out.resize(M);
    const int st = xc_fft2d_r2c_f64(ctx(), in.data(), Nfast, Nslow, reinterpret_cast<double *>(out.data()), nullptr);
    if (st != XC_OK) die(flog, "FFT_2D", st);
    return M;
}

void IFT_2D(vector<double> &out, const vector<CD> &in, int Nfast, int Nslow, FILE *flog)
{
    out.resize((size_t)Nfast * Nslow);
    const int st = xc_ift2d_c2r_f64(ctx(), reinterpret_cast<const double *>(in.data()), Nfast, Nslow, out.data());
    if (st != XC_OK) die(flog, "IFT_2D", st);
}

// Convolve (Correlation.cpp:371-457): src (ws x hs) against the odd-sized kernel K (wk x hk) centred on the
// origin of the FFTSize canvas, by the device's FP64 transforms; a kfft of the canvas's spectrum size is taken
// as the kernel's transform from an earlier call, as the reference takes it (stale or not).
void Convolve(vector<double> &dst, const vector<double> &src, int ws, int hs, const double *K, int wk, int hk,
              bool kIsSymmetric, bool preNormK, vector<CD> &kfft, FILE *flog)
{
    const int Ns = ws * hs, Nk = wk * hk;
    const int Nx = FFTSize(ws, wk / (1 + kIsSymmetric)), Ny = FFTSize(hs, hk / (1 + kIsSymmetric));
    const int Nxy = Nx * Ny, M = Ny * (Nx / 2 + 1);
    if ((int)kfft.size() != M) {
        vector<double> KK((size_t)Nxy, 0.0);
        double nrm = 1.0;
        if (preNormK) {
            nrm = 0.0;
            for (int i = 0; i < Nk; ++i) nrm += K[i];
            if (!nrm) nrm = 1.0;
        }
        const int w2 = wk / 2, h2 = hk / 2;
        for (int i = -h2; i <= h2; ++i) {
            const int y = i >= 0 ? i : Ny + i;
            for (int j = -w2; j <= w2; ++j) {
                const int x = j >= 0 ? j : Nx + j;
                KK[x + (size_t)Nx * y] = K[j + w2 + (size_t)wk * (i + h2)] / nrm;
            }
        }
        FFT_2D(kfft, KK, Nx, Ny, false, flog);
    }
    vector<double> SS((size_t)Nxy, 0.0);
    for (int y = 0; y < hs; ++y)
        for (int x = 0; x < ws; ++x) SS[x + (size_t)Nx * y] = src[x + (size_t)ws * y];
    vector<CD> sfft;
    FFT_2D(sfft, SS, Nx, Ny, false, flog);
    for (int i = 0; i < M; ++i) sfft[i] *= kfft[i];
    IFT_2D(SS, sfft, Nx, Ny, flog);
    dst.resize(Ns);
    for (int y = 0; y < hs; ++y)
        for (int x = 0; x < ws; ++x) dst[x + (size_t)ws * y] = SS[x + (size_t)Nx * y] / Nxy;
}

void ParabPeakFFT(double &xpk, double &ypk, int d, const double *I, int nx, int ny)
{
    auto lk = [&](int x, int y) {
        if (x < 0) x += nx;
        if (y < 0) y += ny;
        return I[x + (long long)nx * y];
    };
    const int ix = (int)xpk, iy = (int)ypk;
    const double z1 = lk(ix, iy);
    double z0, z2;
    if (ix - d >= 0 && ix + d < nx && (z0 = lk(ix - d, iy)) != 0.0 && (z2 = lk(ix + d, iy)) != 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) xpk += z0;
    }
    if (iy - d >= 0 && iy + d < ny && (z0 = lk(ix, iy - d)) != 0.0 && (z2 = lk(ix, iy + d)) != 0.0) {
        z0 = d * (z0 - z2) / (2 * (z0 + z2 - z1 - z1));
        if (fabs(z0) <= 8 * d) ypk += z0;
    }
}

double CorrPatches(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, int Ox, int Oy,
                   int radius, EvalType LegalRgn, void *arglr, EvalType LegalCnt, void *arglc, vector<CD> &fft2)
{
    int L1, R1, B1, T1, L2, R2, B2, T2;
    bbox(ip1, L1, R1, B1, T1);
    bbox(ip2, L2, R2, B2, T2);
    const int w1 = R1 - L1 + 1, h1 = T1 - B1 + 1, w2 = R2 - L2 + 1, h2 = T2 - B2 + 1;
    if (verbose) {
        fprintf(flog, "NormCorr: region size is [%d %d] in x, [%d %d] in y.\n", L1, R1, B1, T1);
        fprintf(flog, "NormCorr: target size is [%d %d] in x, [%d %d] in y.\n", L2, R2, B2, T2);
        fprintf(flog, "NormCorr: Nx = %d, Ny = %d\n", FFTSize(w1, w2), FFTSize(h1, h2));
    }
    RgnTable rg(LegalRgn, arglr, std::min(w1, w2), std::min(h1, h2));
    xc_cnt_rule cr;
    if (!cnt_rule(LegalCnt, arglc, (long long)w1 * h1, (long long)w2 * h2, cr)) {
        fprintf(flog, "NormCorr: LegalCnt callback is not of the form c1 > k1 && c2 > k2.\n");
        exit(44);
    }
    double r = 0.0;
    xc_result res;
    const int st = xc_corr_patches_f64(ctx(), pts(ip1), iv1.data(), (int)ip1.size(), pts(ip2), iv2.data(),
                                       (int)ip2.size(), Ox, Oy, radius, &rg.rule, &cr, &dx, &dy, &r, &res);
    if (st == XC_E_ODD_R) {
        fprintf(flog, "NormCorr: Very odd - |r| > 1.0001\n");
        exit(44);
    }
    if (st != XC_OK && st != XC_NO_PEAK) die(flog, "CorrPatches", st);
    refresh_cache(fft2, ip2, iv2, L2, B2, res.nx, res.ny);
    if (st == XC_NO_PEAK) {
        if (verbose) {
            fprintf(flog, "NormCorr: No legal subregions at all...\n");
            fprintf(flog, "NormCorr: Maximum correlation of %f at [%d,%d].\n", 0.0, 0, 0);
        }
        return 0.0;
    }
    fprintf(flog, "NormCorr: Maximum correlation of %f at [%d,%d].\n", res.r, res.lag_x, res.lag_y);
    return r;
}

double CorrPatchToImage(double &dx, double &dy, const vector<Point> &ip1, const vector<double> &iv1,
                        const vector<double> &i2, int Ox, int Oy, int radius, bool bFilter)
{
    double r = 0.0;
    xc_result res;
    const int st = xc_corr_patch_to_image_f64(ctx(), pts(ip1), iv1.data(), (int)ip1.size(), i2.data(),
                                              (int)i2.size(), Ox, Oy, radius, bFilter ? 1 : 0, &dx, &dy, &r, &res);
    if (st == XC_E_ODD_R) {
        printf("FindCor: Very odd - norm-rslt x y %f %d %d\n", res.r, res.lag_x, res.lag_y);
        exit(44);
    }
    if (st != XC_OK) die(stdout, "CorrPatchToImage", st);
    printf("FindCor: Maximum correlation %f at (%d, %d).\n", res.r, res.lag_x, res.lag_y);
    printf("FindCor: Interpolated max at (%.3f, %.3f).\n", dx, dy);
    return r;
}

static double corr_images(int mode, FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                          const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2,
                          EvalType LegalRgn, void *arglr, EvalType LegalCnt, void *arglc, double mincor,
                          double nbmaxht, int Ox, int Oy, int Rx, int Ry, vector<CD> &fft2)
{
    int L1, R1, B1, T1, L2, R2, B2, T2;
    bbox(ip1, L1, R1, B1, T1);
    bbox(ip2, L2, R2, B2, T2);
    const int w1 = R1 - L1 + 1, h1 = T1 - B1 + 1, w2 = R2 - L2 + 1, h2 = T2 - B2 + 1;
    RgnTable rg(LegalRgn, arglr, std::min(w1, w2), std::min(h1, h2));
    xc_cnt_rule cr;
    if (!cnt_rule(LegalCnt, arglc, (long long)w1 * h1, (long long)w2 * h2, cr)) {
        fprintf(flog, "Corr: LegalCnt callback is not of the form c1 > k1 && c2 > k2.\n");
        exit(44);
    }
    double r = 0.0;
    xc_result res;
    const int st = xc_corr_images_f64(ctx(), mode, pts(ip1), iv1.data(), (int)ip1.size(), pts(ip2), iv2.data(),
                                      (int)ip2.size(), &rg.rule, &cr, mincor, nbmaxht, Ox, Oy, Rx, Ry, &dx, &dy, &r,
                                      &res);
    if (st != XC_OK && st != XC_NO_PEAK && st != XC_TOO_SMALL) die(flog, "CorrImages", st);
    if (st == XC_TOO_SMALL) {
        if (verbose) fprintf(flog, "Corr: Too small to evaluate.\n");
        return 0.0;
    }
    refresh_cache(fft2, ip2, iv2, L2, B2, res.nx, res.ny);
    if (st == XC_NO_PEAK) {
        if (verbose) fprintf(flog, "Corr: No peak candidates.\n");
        return 0.0;
    }
    if (verbose)
        fprintf(flog, "Corr: Max corr %f at [dx,dy]=[%d,%d].\n", res.peak, res.lag_x, res.lag_y);
    return r;
}

double CorrImagesF(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2)
{
    return corr_images(XC_MODE_F, flog, verbose, dx, dy, ip1, iv1, ip2, iv2, LegalRgn, arglr, LegalCnt, arglc, mincor,
                       nbmaxht, Ox, Oy, Rx, Ry, fft2);
}

double CorrImagesR(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2)
{
    return corr_images(XC_MODE_R, flog, verbose, dx, dy, ip1, iv1, ip2, iv2, LegalRgn, arglr, LegalCnt, arglc, mincor,
                       nbmaxht, Ox, Oy, Rx, Ry, fft2);
}

double CorrImagesS(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2)
{
    return corr_images(XC_MODE_S, flog, verbose, dx, dy, ip1, iv1, ip2, iv2, LegalRgn, arglr, LegalCnt, arglc, mincor,
                       nbmaxht, Ox, Oy, Rx, Ry, fft2);
}

void LegPolyFlatten(vector<double> &vals, const unsigned char *src, int w, int h, int maxOrder)
{
    vals.resize((size_t)w * h);
    const int st = xc_leg_poly_flatten_f64(ctx(), src, w, h, maxOrder, vals.data());
    if (st != XC_OK) die(stderr, "LegPolyFlatten", st);
}

double FourierMatch(const vector<Point> &pts, const vector<double> &av, const vector<double> &bv, int wvlen,
                    bool write_images, const char *msg, FILE *flog)
{
    double dot = 0.0;
    xc_fourier_qa qa;
    if (pts.empty()) {
        fprintf(flog, "Metrics: No data?\n");
        return 0.0;
    }
    const int st = xc_fourier_match_f64(ctx(), reinterpret_cast<const xc_point *>(pts.data()), av.data(), bv.data(),
                                        (int)pts.size(), wvlen, &dot, &qa);
    if (st != XC_OK) die(flog, "FourierMatch", st);
    if (write_images) fprintf(flog, "FourierMatch: metric images are not written by this build.\n");
    fprintf(flog, "Metrics: %s: Mean square difference %f\n", msg, qa.msd);
    fprintf(flog, "Metrics: %s: Smallest footprint deg=%d\n", msg, qa.deg);
    fprintf(flog, "FFT: norm-dot %f, energy %f\n", qa.dot, qa.energy);
    fprintf(flog, "FFT: %s: Cum power exceeds %2d%% of %f at index %d/%d (frac %f).\n", msg, 50, qa.tot_power,
            qa.half_index, qa.maxf, qa.half_frac);
    return dot;
}
