// Host-only checks of the ThmPair table writer and readers (xc_thmpair.hip, compiled as C++),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/xcorr.h"

#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                   \
        }                                                              \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    char path[4096];
    snprintf(path, sizeof path, "%s/ThmPair_1^0.txt", argv[1]);
    CHECK(xc_thmpair_read_all(path, nullptr, 0) == XC_E_ARG);  // no file yet
    CHECK(xc_thmpair_create(path) == XC_OK);
    xc_thmpair e;
    memset(&e, 0, sizeof e);
    for (int k = 0; k < 40; ++k) {
        e.atl = k;
        e.acr = k % 3;
        e.btl = k + 1;
        e.bcr = 0;
        e.err = k % 5 == 0;
        e.A = 0.25 * k - 3.0;
        e.R = 0.5 + k * 1e-3;
        for (int j = 0; j < 6; ++j) e.T[j] = 1e6 * (j - 2.5) + k;  // wide values: long lines
        CHECK(xc_thmpair_append(path, &e) == XC_OK);
    }
    e.atl = -1;
    CHECK(xc_thmpair_append(path, &e) == XC_OK);  // skipped, as the reference skips it
    xc_thmpair all[64];
    CHECK(xc_thmpair_read_all(path, all, 64) == 40);
    CHECK(xc_thmpair_read_all(path, all, 3) == 40);  // entries beyond cap are counted only
    CHECK(all[2].atl == 2 && all[2].btl == 3 && all[2].A == -2.5);
    xc_thmpair f;
    CHECK(xc_thmpair_find(path, 7, 1, 8, 0, &f) == XC_OK && f.R > 0.506 && f.R < 0.508);
    CHECK(xc_thmpair_find(path, 7, 2, 8, 0, &f) == XC_NO_PEAK);
    double ang = 99.0;
    CHECK(xc_thmpair_prior_angle(all, 3, &ang) == 0 && ang == 99.0);  // fewer than 4 error-free entries
    xc_thmpair m[64];
    const int n = xc_thmpair_read_all(path, m, 64);
    const int c = xc_thmpair_prior_angle(m, n, &ang);
    CHECK(c == 32 && ang > -3.0 && ang < 7.0);
    // a file of one very long line and a truncated header
    FILE *fp = fopen(path, "w");
    CHECK(fp != nullptr);
    for (int k = 0; k < 100000; ++k) fputc('x', fp);
    fclose(fp);
    const int rl = xc_thmpair_read_all(path, m, 64);
    CHECK(rl == 0 || rl == XC_E_ARG);
    printf("thmpair ok\n");
    return 0;
}
