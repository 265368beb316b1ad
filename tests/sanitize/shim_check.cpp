// Host-only entry points of the C++ drop-in (aln_corr_compat.cpp: FFTSize, FFTSizeSQR, ParabPeakFFT),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py against libxcorr.so.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "../../include/aln_corr_compat.h"

#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                        \
        }                                                                   \
    } while (0)

int main()
{
    CHECK(FFTSize(513, 512) == 1024 && FFTSize(1, 1) == 1 && FFTSize(2048, 2048) == 4096);
    CHECK(FFTSizeSQR(300, 200, 100, 400) == 1024);
    const int nx = 16, ny = 8;
    std::vector<double> I(nx * ny, 0.0);
    auto at = [&](int x, int y) -> double & { return I[((x + nx) % nx) + nx * ((y + ny) % ny)]; };
    // a peak at (5, 3) with its four neighbours: both fits run and move by less than one pixel
    at(5, 3) = 1.0;
    at(4, 3) = 0.6;
    at(6, 3) = 0.8;
    at(5, 2) = 0.7;
    at(5, 4) = 0.5;
    double x = 5, y = 3;
    ParabPeakFFT(x, y, 1, I.data(), nx, ny);
    CHECK(x > 5.0 && x < 6.0 && y > 2.0 && y < 3.0);
    // a peak on the border and one whose neighbours are zero: no fit, no access outside the image
    for (int px : {0, nx - 1})
        for (int py : {0, ny - 1}) {
            double a = px, b = py;
            ParabPeakFFT(a, b, 1, I.data(), nx, ny);
            CHECK(fabs(a - px) <= 8 && fabs(b - py) <= 8);
        }
    x = 10;
    y = 6;
    ParabPeakFFT(x, y, 2, I.data(), nx, ny);
    CHECK(x == 10 && y == 6);
    printf("shim ok\n");
    return 0;
}
