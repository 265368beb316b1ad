/*
 * aln_corr_compat.h -- C++ drop-in for the correlation entry points of
 * billkarsh/Alignment_Projects (0_GEN/Correlation.h:16-220), implemented on
 * the C-ABI of libxcorr (xcorr.h).  Link libaln_corr.so instead of the
 * reference's Correlation.o; the signatures below are the reference's.
 *
 * Types: inside the reference tree build with -DALN_COMPAT_WITH_REFERENCE so
 * that Point, CD and the rest come from the reference's own GenDefs.h /
 * CPoint.h.  Standalone (tests, other callers) this header declares
 * layout-identical equivalents: Point {double x, y;}, CD = complex<double>.
 *
 * Legality callbacks: a LegalRgn callback depends only on (olw, olh, arg) and
 * is tabulated on the host for the pair's overlap range, so any callback works
 * exactly.  A LegalCnt callback is mapped to the threshold form c1 > k1 &&
 * c2 > k2 (EnoughPoints, CThmScan.cpp:39-42, and the always-true lc/lcTRUE):
 * the shim finds k1, k2 by bisection and verifies the inferred rule on probe
 * points; a callback that is not of that form is reported on flog and the call
 * fails as the reference fails on an odd result (exit 44).
 *
 * Threads: each calling thread gets its own device context (HIP stream); calls
 * from EZThreads workers run concurrently, as the MKL build of the reference
 * does.  Device: ALN_XC_DEVICE (default 0).
 */
#ifndef ALN_CORR_COMPAT_H
#define ALN_CORR_COMPAT_H

#include <stdio.h>

#include <complex>
#include <vector>

#include "xcorr.h"

#ifdef ALN_COMPAT_WITH_REFERENCE
#include "CPoint.h"
#include "GenDefs.h"
#else
using std::vector;
typedef std::complex<double> CD;
class Point {
public:
    double x, y;
    Point() : x(0.0), y(0.0) {}
    Point(double x, double y) : x(x), y(y) {}
};
#endif

/* FFT support: FFTSize (Correlation.cpp:344), FFTSizeSQR (:356). */
int FFTSize(int n1, int n2);
int FFTSizeSQR(int w1, int h1, int w2, int h2);

/* FFT_2D / IFT_2D (Correlation_fft_fftw.cpp:38-92), on the device in FP64. */
int FFT_2D(vector<CD> &out, const vector<double> &in, int Nfast, int Nslow, bool cached, FILE *flog = stderr);
void IFT_2D(vector<double> &out, const vector<CD> &in, int Nfast, int Nslow, FILE *flog = stderr);

/* ParabPeakFFT (Correlation.cpp:482-517). */
/* Convolve (Correlation.h:38-50, Correlation.cpp:371-457): by the device's FP64 transforms; a kfft of the
 * canvas's spectrum size is reused as the kernel's transform, as the reference reuses it. */
void Convolve(vector<double> &dst, const vector<double> &src, int ws, int hs, const double *K, int wk, int hk,
              bool kIsSymmetric, bool preNormK, vector<CD> &kfft, FILE *flog);
void ParabPeakFFT(double &xpk, double &ypk, int d, const double *I, int nx, int ny);

typedef bool (*EvalType)(int sx, int sy, void *v);

/* CorrPatches (Correlation.cpp:901-1170). */
double CorrPatches(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, int Ox, int Oy,
                   int radius, EvalType LegalRgn, void *arglr, EvalType LegalCnt, void *arglc, vector<CD> &fft2);

/* CorrPatchToImage (Correlation.cpp:1186-1355). */
double CorrPatchToImage(double &dx, double &dy, const vector<Point> &ip1, const vector<double> &iv1,
                        const vector<double> &i2, int Ox, int Oy, int radius, bool bFilter);

/* CorrImagesF / CorrImagesR / CorrImagesS (Correlation.cpp:3066-3122, 3144-3290). */
double CorrImagesF(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2);
double CorrImagesR(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2);
double CorrImagesS(FILE *flog, int verbose, double &dx, double &dy, const vector<Point> &ip1,
                   const vector<double> &iv1, const vector<Point> &ip2, const vector<double> &iv2, EvalType LegalRgn,
                   void *arglr, EvalType LegalCnt, void *arglc, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                   int Ry, vector<CD> &fft2);

/* LegPolyFlatten, 8-bit form (Maths.cpp:1006-1086, Maths.h:153-158), on the device in FP64. */
void LegPolyFlatten(vector<double> &vals, const unsigned char *src, int w, int h, int maxOrder);

/* FourierMatch (Metrics.cpp:517-658, Metrics.h:23-30), on the device in FP64.  flog receives the
 * mean-square-difference, norm-dot and half-power lines; write_images is not supported (noted on flog). */
double FourierMatch(const vector<Point> &pts, const vector<double> &av, const vector<double> &bv, int wvlen,
                    bool write_images, const char *msg, FILE *flog);

#endif /* ALN_CORR_COMPAT_H */
