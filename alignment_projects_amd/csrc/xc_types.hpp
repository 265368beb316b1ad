// xc_types.hpp -- device-side descriptors shared by the kernels and the host
// launch code of libxcorr.
#pragma once

#include <stdint.h>

namespace xc {

// Lag-row storage order.
//  RASTER : CorrPatches' wrapped scan order, iy = 0..Ny-1 (positive lags
//           first, then the negative zone), Correlation.cpp:1017-1070.
//  CENTRED: CCorImg's centred image, signed y ascending from -nnegy,
//           Correlation.cpp:2141-2161.
enum LagOrder { LAG_RASTER = 0, LAG_CENTRED = 1 };

enum PairMode { PM_NCC = 0, PM_R = 1, PM_S = 2 };

struct RuleDev {
    int32_t rgn_enabled, min_w, min_h, exact_w, exact_h, or_w_ge, or_h_gt, table_w, table_h;
    int32_t cnt_enabled;
    int64_t min_area;
    int64_t k1, k2;
    const uint8_t *table;  // device copy of an olw x olh legality table, or null
};

// Per-row partial of the peak search (one per stored lag row).
struct RowRec {
    double best;    // best value among candidates in this row
    double left;    // value at key-1 (same row)
    double right;   // value at key+1 (same row)
    double sum;     // S mode: sum of S over the row's centred lags
    double sum2;    //          and of S^2
    int32_t key;    // within-row index of the best (raster ix or centred x)
    int32_t flags;  // bit0: row has a candidate; bit1: |r| > 1.0001 seen
    int32_t pad[2];
};

// One correlation, as the kernels see it.  Images are compact w x h rasters
// (row stride w) whose origin is the point list's bounding-box corner; the
// zero padding to Nx x Ny is implicit.
struct PairDev {
    int32_t w1, h1, w2, h2;
    int32_t Nx, Ny;          // padded FFT size (Nx >= 2)
    int32_t Nx_ref, Ny_ref;  // the reference's FFTSize() (differs only when it is 1)
    int32_t mode;            // PairMode
    int32_t lag_order;       // LagOrder
    int32_t Ox, Oy, radius;  // NCC disc
    int32_t Rx, Ry;          // R/S oval
    int32_t offx, offy;      // B2.L - B1.L, B2.B - B1.B
    int32_t nnegx, nposx, nnegy, nposy, wR, hR;
    int32_t need_cnt;        // count predicate active (nz tables needed)
    int32_t pad0;
    double mincor;
    RuleDev rule;
    const void *imgA, *imgB;          // T rasters
    const double *SA, *QA, *SB, *QB;  // FP64 summed-area tables (S, S^2)
    const int32_t *ZA, *ZB;           // non-zero count tables
    void *specA, *specB;              // cpx<T> [h][Nx/2+1] row spectra
    void *lag;                        // cpx<T> [hR][Nx/2+1] lag rows
    RowRec *rows;                     // [hR]
    void *out;                        // xc_result (device)
};

}  // namespace xc
