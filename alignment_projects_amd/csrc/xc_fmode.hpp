// xc_fmode.hpp -- CorrImagesF's peak selection (MakeF, OrderF, FPeak,
// ReturnR) on the device, run after the R-mode pipeline.
#pragma once

#include "xc_host.hpp"
#include "xc_types.hpp"

namespace xc {

// Device workspace of one CorrImagesF call over a wR x hR R image.
struct FWork {
    double *R;      // centred R image (written by the c2r evaluation)
    uint8_t *cand;  // its candidate flags (A, oval, mincor); MakeF clears the guard band
    double *F;      // filtered, normalised image
    double *kin, *kout;  // sort keys (F of candidates, -1 otherwise)
    int *vin, *vout;     // pixel indices, sorted by decreasing key (ties: increasing index)
    uint8_t *mask;       // FPeak's core mask
    double *bmax;        // per-block maxima of F
    int2 *rec;           // per-candidate test results of one chunk
    int *ints;           // [0] candidate count, [1] accepted index or -1
    void *sort;          // radix-sort scratch
    size_t sort_bytes;
};

size_t fmode_bytes(long long nR);
void fmode_carve(char *base, long long nR, FWork &w);
// MakeF, OrderF, FPeak and ReturnR on stream s; writes the xc_result of
// pairs[0] (its `out`).  The R image and flags must be complete on s.
int fmode_select(xc_ctx *ctx, hipStream_t s, const PairDev *dP, int wR, int hR, double nbmaxht, const FWork &w);

}  // namespace xc
