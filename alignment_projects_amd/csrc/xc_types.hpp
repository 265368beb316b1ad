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
enum LagOrder { LAG_RASTER = 0, LAG_CENTRED = 1, LAG_ALL = 2 };  // LAG_ALL: every raster row stored

// PM_P2I: CorrPatchToImage (Correlation.cpp:1186-1355), raw correlation
// surface at N x N without zero padding.
enum PairMode { PM_NCC = 0, PM_R = 1, PM_S = 2, PM_P2I = 3 };

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

// A source image as the kernels read it: raw pixels of any supported type,
// an optional presence mask, and a device (mean, sd, 1/sd) triple applied
// as (v - mean) * (1/sd) before rounding to the transform precision
// (Normalize, Maths.cpp:192-212; (0, 1, 1) for raw values, which leaves
// them unchanged).  Absent pixels read as 0,
// exactly as ImageFromValuesAndPoints leaves them.
struct ImgView {
    const void *data;
    const uint8_t *mask;
    const double *msd;   // device (mean, sd, 1/sd)
    int32_t dtype;       // xc_dtype
    int32_t w, h;
    int32_t stride;      // elements per row of data
    int32_t mstride;     // bytes per row of mask
    int32_t pad;
};

// Forms of a summed-area table of x and x^2 (IntegrateImage, Correlation.cpp:49-105):
//  TK_F64: interleaved (S, S^2) FP64 pairs, 16 bytes a pixel (the reference keeps double tables);
//  TK_F32: interleaved FP32 pairs, each the FP64 sum rounded once (tiles that ask for them);
//  TK_X  : exact integer sums of an unmasked 8-bit image of at most 2^20 pixels, one uint64
//          (Sraw << 36) + Qraw a pixel, Sraw / Qraw the sums of the raw values / their squares.
enum TabKind { TK_F64 = 0, TK_F32 = 1, TK_X = 2 };

// Per-image preparation record (statistics and summed-area tables).
struct TileDev {
    ImgView v;
    double *msd;         // out: (mean, sd, 1/sd)
    void *SQ;            // out: summed-area tables of x and x^2 per pixel, w x h, in the form tabk names
    int32_t *Z;          // out: non-zero count table, w x h
    double *part;        // scratch: [TP_PARTS][3] statistics partials
    double *bandp;       // scratch: column sums of the table sweep's row bands (nbands > 1)
    int32_t normalize;
    int32_t nbands;      // row bands of the table sweep (1: one workgroup walks every row)
    int32_t tabk;        // TabKind of SQ
    int32_t pad;
};

// One rotated copy of an image's pixel grid (CThmScan::RFromAngle,
// CThmScan.cpp:329-366): points (x, y) -> (x t0 + y t1, x t3 + y t4)
// (TAffine::Apply_R_Part, TAffine.cpp:407-416), splatted bilinearly at
// (x' - L, y' - B) into a w x h raster (ImageFromValuesAndPoints).
struct RotDev {
    ImgView src;
    double t0, t1, t3, t4;  // the 2 x 2 map applied to A's points
    double u0, u1, u3, u4;  // its inverse (the transpose for a rotation)
    int32_t L, B;
    int32_t w, h;
    void *out;
    const double *nsrc;     // src's normalised values (raw - mean) / sd, src.w a row (RotNorm's dst)
};

// A rotated copy's source, normalised once per run for all of its copies.
struct RotNorm {
    ImgView src;
    double *dst;            // src.w x src.h values, src.w a row
};

// Work item of the 8-bit row r2c (FP32 transforms of u8 rows that fit the
// upper-half padding): ROWS rows of one image from src (row stride `stride`
// bytes, w pixels a row; presence mask rows at mask, mstride bytes apart, or
// null) into dst (rows of Nx/2 + 1 complex values), nrows of them present; the
// image's (mean, sd, 1/sd) at msd.
struct R2cU8 {
    const uint8_t *src;
    void *dst;
    const double *msd;
    const uint8_t *mask;
    int32_t nrows, w, stride, mstride;
};

constexpr int TP_PARTS = 16;     // statistics partials per image
constexpr int TP_MAX_W = 8192;   // widest image the table sweep handles

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
    int32_t filter;          // PM_P2I: Gaussian low-pass of the cross power (bFilter)
    int32_t ylo, yhi;        // signed lag rows computed; rows outside hold no legal lag (all 0, no candidate)
    int32_t tabk;            // TabKind of SQA and SQB
    double norm;             // PM_P2I: np1 * N^2
    double mincor;
    RuleDev rule;
    ImgView A, B;                     // source images
    const void *SQA, *SQB;            // summed-area tables of x and x^2 (form: tabk)
    const int32_t *ZA, *ZB;           // non-zero count tables
    void *specA, *specB;              // cpx<T> [h][Nx/2+1] row spectra
    void *lag;                        // cpx<T> [hR][Nx/2+1] lag rows
    RowRec *rows;                     // [hR]
    void *out;                        // xc_result (device)
    double *rimg;                     // R mode: centred wR x hR R image written by the c2r (or null)
    uint8_t *cimg;                    //         and its candidate flags (A, oval, mincor)
};

}  // namespace xc
