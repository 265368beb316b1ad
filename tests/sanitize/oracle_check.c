/* The oracle's entry points on small cases, built with AddressSanitizer and
 * UndefinedBehaviorSanitizer by tests/test_sanitizers.py (with the oracle's own
 * radix-2 FFT): each must run clean and give the planted answer. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/xcorr.h"

int xo_corr_patches(const xc_point *, const double *, int, const xc_point *, const double *, int, int, int, int,
                    const xc_rgn_rule *, const xc_cnt_rule *, double *, double *, double *, xc_result *, double *, int);
int xo_corr_patch_to_image(const xc_point *, const double *, int, const double *, int, int, int, int, int, double *,
                           double *, double *, xc_result *);
int xo_corr_images(int, const xc_point *, const double *, int, const xc_point *, const double *, int,
                   const xc_rgn_rule *, const xc_cnt_rule *, double, double, int, int, int, int, double *, double *,
                   double *, xc_result *, double *, double *, uint8_t *, int);
int xo_rfromangle(double, const xc_point *, const double *, int, const xc_point *, const double *, int, int, long,
                  int, int, int, int, double *, double *, double *, xc_result *);
int xo_decimate_vector(xc_point *, double *, int, int, int, int);
int xo_convolve(double *, const double *, int, int, const double *, int, int, int, int);
void xo_integrate_image(double *, double *, int *, int, int, const double *, int);
int xo_make_dog_kernel(double *, int, int);
int xo_fourier_match(const xc_point *, const double *, const double *, int, int, double *, xc_fourier_qa *);
double xo_normalize(double *, int);

#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                        \
        }                                                                   \
    } while (0)

/* a smooth deterministic field: sums of incommensurate sines */
static double field(double x, double y)
{
    return sin(0.37 * x + 0.11 * y) + 0.7 * sin(0.23 * y - 0.05 * x + 1.0) + 0.4 * sin(0.61 * x * 0.9 + 0.43 * y);
}

static void raster(int w, int h, int sx, int sy, xc_point *p, double *v)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            p[x + w * y].x = x;
            p[x + w * y].y = y;
            v[x + w * y] = field(x + sx, y + sy);
        }
}

int main(void)
{
    enum { W = 48, H = 40, N = W * H };
    xc_point *pa = malloc(sizeof(xc_point) * N), *pb = malloc(sizeof(xc_point) * N);
    double *va = malloc(sizeof(double) * N), *vb = malloc(sizeof(double) * N);
    raster(W, H, 0, 0, pa, va);
    raster(W, H, 5, -3, pb, vb); /* B(p) = A(p + (5, -3)) */
    xo_normalize(va, N);
    xo_normalize(vb, N);
    xc_rgn_rule rgn;
    memset(&rgn, 0, sizeof rgn);
    rgn.enabled = 1;
    rgn.min_w = 16;
    rgn.min_h = 16;
    rgn.min_area = 300;
    xc_cnt_rule cnt;
    memset(&cnt, 0, sizeof cnt);
    cnt.enabled = 1;
    cnt.k1 = cnt.k2 = 400;
    double dx, dy, r, *rs = malloc(sizeof(double) * 128 * 128);
    xc_result res;
    int st = xo_corr_patches(pa, va, N, pb, vb, N, 0, 0, 4096, &rgn, &cnt, &dx, &dy, &r, &res, rs, 128 * 128);
    CHECK(st == XC_OK && res.lag_x == -5 && res.lag_y == 3 && r > 0.999);
    for (int mode = XC_MODE_R; mode <= XC_MODE_F; ++mode) {
        const int cap = 200 * 200;
        double *R = malloc(sizeof(double) * cap), *S = malloc(sizeof(double) * cap);
        uint8_t *A = malloc(cap);
        st = xo_corr_images(mode, pa, va, N, pb, vb, N, &rgn, NULL, 0.0, 1.0, 0, 0, -1, -1, &dx, &dy, &r, &res, R, S,
                            A, cap);
        CHECK(st == XC_OK || st == XC_NO_PEAK);
        free(R);
        free(S);
        free(A);
    }
    /* CorrPatchToImage: A's points against a 64 x 64 image holding B's raster */
    double *img = calloc(64 * 64, sizeof(double));
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) img[x + 64 * y] = vb[x + W * y];
    for (int f = 0; f < 2; ++f) {
        st = xo_corr_patch_to_image(pa, va, N, img, 64 * 64, 0, 0, 64, f, &dx, &dy, &r, &res);
        CHECK(st == XC_OK || st == XC_E_ODD_R);
    }
    /* RFromAngle at a small angle */
    st = xo_rfromangle(0.5, pa, va, N, pb, vb, N, 16, 300, 0, 0, -1, -1, &dx, &dy, &r, &res);
    CHECK(st == XC_OK || st == XC_NO_PEAK);
    /* tables, DoG, convolution, decimation, Fourier match */
    double *S1 = malloc(sizeof(double) * N), *S2 = malloc(sizeof(double) * N);
    int *nz = malloc(sizeof(int) * N);
    xo_integrate_image(S1, S2, nz, W, H, va, W);
    CHECK(nz[N - 1] == N && fabs(S1[N - 1]) < 1e-9);
    const int D = xo_make_dog_kernel(NULL, 2, 4);
    double *K = malloc(sizeof(double) * D * D), *cv = malloc(sizeof(double) * N);
    CHECK(xo_make_dog_kernel(K, 2, 4) == D);
    CHECK(xo_convolve(cv, va, W, H, K, D, D, 1, 1) >= 0);
    xc_point *pd = malloc(sizeof(xc_point) * N);
    double *vd = malloc(sizeof(double) * N);
    memcpy(pd, pa, sizeof(xc_point) * N);
    memcpy(vd, va, sizeof(double) * N);
    CHECK(xo_decimate_vector(pd, vd, N, W, H, 4) >= 0);
    double dot;
    xc_fourier_qa qa;
    CHECK(xo_fourier_match(pa, va, vb, N, 8, &dot, &qa) == XC_OK);
    free(pa);
    free(pb);
    free(va);
    free(vb);
    free(rs);
    free(img);
    free(S1);
    free(S2);
    free(nz);
    free(K);
    free(cv);
    free(pd);
    free(vd);
    printf("oracle ok\n");
    return 0;
}
