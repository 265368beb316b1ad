// compat_kat.cpp -- runs the reference's retired harness testnormcorr
// (Retired_Proj/4_TestNormCorr/testnormcorr.cpp:20-63) through the drop-in
// signatures, plus a CorrImagesS call with EnoughPoints-style callbacks,
// FourierMatch, LegPolyFlatten and Convolve (with its kernel-transform
// cache), and prints one machine-readable line per case.
#include <stdio.h>

#include "../../include/aln_corr_compat.h"

static bool SomeArea(int sx, int sy, void *) { return sx > 1 && sy > 1; }
static bool BigEnough16(int sx, int sy, void *a) { return sx >= 16 && sy >= 16 && (long)sx * sy > (long)a; }
static bool EnoughPoints(int c1, int c2, void *a) { return c1 > (long)a && c2 > (long)a; }

int main()
{
    vector<Point> pts, pb;
    vector<double> vals, vb;
    pb.push_back(Point(0, 3)); vb.push_back(1);
    pb.push_back(Point(2, 3)); vb.push_back(2);
    pb.push_back(Point(3, 2)); vb.push_back(1);
    pb.push_back(Point(3, 3)); vb.push_back(3);
    pb.push_back(Point(3, 4)); vb.push_back(1);
    pb.push_back(Point(4, 3)); vb.push_back(2);
    pb.push_back(Point(5, 4)); vb.push_back(1);
    pts.push_back(Point(7.0, 2.0)); vals.push_back(1.0);
    pts.push_back(Point(8.0, 2.0)); vals.push_back(0.0);
    pts.push_back(Point(9.0, 2.0)); vals.push_back(1.0);
    pts.push_back(Point(7.0, 1.0)); vals.push_back(3.0);
    pts.push_back(Point(8.0, 1.0)); vals.push_back(2.0);
    double dx = 0, dy = 0;
    vector<CD> ftc;
    double c = CorrPatches(stdout, false, dx, dy, pts, vals, pb, vb, 0, 0, 4096, SomeArea, NULL, NULL, NULL, ftc);
    printf("KAT testnormcorr %.17g %.17g %.17g %zu\n", c, dx, dy, ftc.size());

    // a 64 x 48 pattern against itself shifted by (5, -3), CorrImagesS with callbacks
    vector<Point> a, b;
    vector<double> va, vbb;
    unsigned s = 12345;
    vector<double> f(80 * 64);
    for (size_t i = 0; i < f.size(); ++i) {
        s = s * 1103515245u + 12345u;
        f[i] = (double)((s >> 16) & 0x7fff) / 32768.0 - 0.5;
    }
    for (int y = 0; y < 48; ++y)
        for (int x = 0; x < 64; ++x) {
            a.push_back(Point(x, y));
            va.push_back(f[(y + 8) * 80 + (x + 8)]);
            b.push_back(Point(x, y));
            vbb.push_back(f[(y + 8 + 3) * 80 + (x + 8 - 5)]);
        }
    vector<CD> ftc2;
    const long req = 64 * 48 / 4;
    double r = CorrImagesS(stdout, false, dx, dy, a, va, b, vbb, BigEnough16, (void *)req, EnoughPoints, (void *)req,
                           0.0, 0.0, 0, 0, -1, -1, ftc2);
    printf("KAT imagesS %.17g %.17g %.17g %zu\n", r, dx, dy, ftc2.size());

    // FourierMatch of the same region's two value lists, wavelength 25
    double fm = FourierMatch(a, va, vbb, 25, false, "KAT", stdout);
    printf("KAT fourier %.17g\n", fm);

    // LegPolyFlatten of the 80 x 64 field as 8-bit pixels, order 2
    vector<unsigned char> ras(f.size());
    for (size_t i = 0; i < f.size(); ++i) ras[i] = (unsigned char)((f[i] + 0.5) * 255.0);
    vector<double> flat;
    LegPolyFlatten(flat, ras.data(), 80, 64, 2);
    printf("KAT legflat %zu %.17g %.17g %.17g\n", flat.size(), flat[0], flat[1234], flat.back());

    // Convolve of the 80 x 64 field with a 7 x 5 kernel (symmetric canvas, prenormalised); the second call
    // hands back the first call's kernel transform with another kernel: the transform in kfft is used
    double K[35], K2[35];
    for (int i = 0; i < 35; ++i) {
        K[i] = 1.0 + (i % 5);
        K2[i] = 100.0 - i;
    }
    vector<double> cv, cv2;
    vector<CD> kfft;
    Convolve(cv, f, 80, 64, K, 7, 5, true, true, kfft, stdout);
    Convolve(cv2, f, 80, 64, K2, 7, 5, true, true, kfft, stdout);
    printf("KAT convolve %zu %zu %.17g %.17g %.17g %.17g\n", cv.size(), kfft.size(), cv[0], cv[1234], cv.back(),
           cv2[1234]);
    return 0;
}
