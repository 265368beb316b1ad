// xc_host.hpp -- host-side context, workspace and launch helpers of libxcorr.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/xcorr.h"
#include "xc_fft.hpp"
#include "xc_types.hpp"

namespace xc {

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace xc

// A device-resident image (see xc_tile_put).
struct xc_tile {
    xc_ctx *ctx = nullptr;
    int w = 0, h = 0;           // the view: bounding box of the present pixels
    int ox = 0, oy = 0;         // its corner in the picture
    int fw = 0, fh = 0;         // the whole picture
    int dtype = 0, flags = 0;
    xc::ImgView full;           // the whole picture (rotated copies read it)
    void *raw = nullptr;       // device pixels (owned unless adopted)
    bool own_raw = false;
    uint8_t *mask = nullptr;   // device mask (owned unless adopted)
    bool own_mask = false;
    xc::DevBuf mem;                // tables + scratch + descriptor
    xc::TileDev dev;               // host copy, device pointers
    xc::TileDev *d_dev = nullptr;  // device copy
    int64_t bytes = 0;
    uint64_t uid = 0;           // process-unique id (caches key on it, never on the address)
    std::vector<int> rx0, rx1;  // per picture row: first and last present x (-1: none)
};

// A tuned column-pass shape (see XC_PLAN_AUTOTUNE).
struct xc_col_tune {
    int c = 0, xcd = 1;
    double us = 0.0;
};

struct xc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream: table sweeps run here beside the transforms
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    std::mutex mu;
    std::string err;
    xc::cpx<double> *tw64 = nullptr;
    xc::cpx<float> *tw32 = nullptr;
    xc::DevBuf arena;       // per-call device workspace
    void *pinned = nullptr; // pinned host staging
    size_t pinned_bytes = 0;
    int num_cu = 0;
    size_t lds_max = 65536;  // LDS bytes one workgroup may allocate
    std::map<std::tuple<int, int, int>, xc_col_tune> tune;  // (NH, NY, whiten) -> column-pass shape
    // (NH, NY, kernel: 0 general, 1 8-bit, 2 8-bit general rows) -> FP32 row r2c shape (c: rows per thread,
    // xcd: work items per workgroup of the 8-bit kernel)
    std::map<std::tuple<int, int, int>, xc_col_tune> tune_r2c;
    std::string tune_file;  // XC_TUNE_FILE: tuned shapes persist here across processes
};

namespace xc {

// Status helpers ----------------------------------------------------------

int set_err(xc_ctx *ctx, int code, const char *fmt, ...);

#define XC_HIP(ctx, call)                                                                        \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ::xc::set_err((ctx), XC_E_DEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                 __FILE__, __LINE__);                                            \
    } while (0)

// Workspace layout: reserve sizes, then materialise once.
struct Layout {
    std::vector<size_t> off;
    size_t total = 0;
    size_t add(size_t bytes)
    {
        size_t o = (total + 255) & ~(size_t)255;
        off.push_back(o);
        total = o + bytes;
        return off.size() - 1;
    }
};

// slack: allocate a quarter more than asked, for buffers that grow with reuse (not for fixed-size ones).
int ensure_dev(xc_ctx *ctx, DevBuf &b, size_t bytes, bool slack = true);
int ensure_pinned(xc_ctx *ctx, size_t bytes);

// Host mirror of the reference geometry helpers.
struct IBox {
    int L, R, B, T;
};
void bbox_from_points(IBox &B, const xc_point *p, int n);  // Geometry.cpp:78-106
int ceil_pow2(int n);                                      // Maths.cpp:97-105

RuleDev make_rule(const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, const uint8_t *dev_table);

// ImageFromValuesAndPoints into a zeroed w x h raster, in the reference's order of
// additions (xc_splat.hip).  Off-grid lists need splat_scratch_bytes of device scratch.
bool points_off_grid(const xc_point *p, int n);
bool copy_points_off_grid(xc_point *dst, const xc_point *p, int n);  // the copy and the test in one pass
// The staging copy of n points (with the test) and of up to two value lists (da / db may be null), split over a
// few host threads for large lists; returns whether a point is off the grid.
bool stage_points(xc_point *dp, const xc_point *p, double *da, const double *a, double *db, const double *b, int n);
size_t splat_scratch_bytes(int n, long long npix);
int splat_points(xc_ctx *ctx, hipStream_t s, const xc_point *pts, const double *val, int n, double xmin, double ymin,
                 double *img, int w, int h, bool off_grid, void *scratch, size_t scratch_bytes);

// A device workspace a call must fit in (the fixed scratch of the
// allocation-free entry point).
struct FixedWs {
    void *dev = nullptr;
    size_t bytes = 0;
};

// One plain FP64 correlation of point lists (xc_plain.hip); mode PM_NCC / PM_R / PM_S, fmode: CorrImagesF's
// peak selection; fixed: run in that workspace and allocate nothing (null: the context's arena and staging).
int run_plain(xc_ctx *ctx, int mode, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
              const double *iv2, int n2, int Ox, int Oy, int radius, int Rx, int Ry, double mincor,
              const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, xc_result *out, bool fmode, double nbmaxht,
              const FixedWs *fixed);

// One batched FP32 sweep (CorrImagesS) of the 2 x 2 maps maps[4i..4i+3] applied to A's pixel grid against B,
// each with its own search oval ovals[4i..4i+3] = (Ox, Oy, Rx, Ry) (xc_batch.hip).
int sweep_affine_ovals(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *maps, const int *ovals,
                       int n_maps, const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, xc_result *results);

// Looks up the device address of the allocation-free entry point's fixed
// scratch on ctx->device (current), once per device (xc_noalloc.hip).
int noalloc_init(xc_ctx *ctx);

// Fill the derived geometry of a PairDev from (w1,h1,w2,h2) and the mode.
int setup_pair_geometry(PairDev &P, int w1, int h1, int w2, int h2, int mode);

// NCC pairs: shrink [ylo, yhi] to the hull of the lag rows where some lag
// can pass the disc and the rules (call after rule, Ox, Oy, radius are set).
// Outside it every lag is 0 and no candidate, as the reference's scan leaves
// it, so those rows are neither written nor evaluated.
void narrow_rows(PairDev &P);
// Start rows of the G-row c2r workgroups that cover the live rows.
void live_row_groups(const PairDev &P, int G, std::vector<int> &starts);
int live_rows(const PairDev &P);

// Power-of-two transform lengths the launchers dispatch on.
#define XC_POW2_CASES(F) \
    F(1) F(2) F(4) F(8) F(16) F(32) F(64) F(128) F(256) F(512) F(1024) F(2048) F(4096)

// Kernel launchers (dispatch on transform length).
// rpt: rows per thread of the r2c (0: default shape; see r2c_alt_rpt for the other one).
template <typename T>
int launch_row_r2c(xc_ctx *ctx, int NH, int nblocks, const PairDev *pd, const int4 *work, int rpt = 0);
// The 8-bit row r2c (FP32, u8 rows of w <= NH): work items of
// R2cCfg<float, NH>::ROWS rows, np of them per workgroup (1, 2 or 4); gen: some
// rows are masked or start at odd addresses (bytewise loads, FP64 normalisation).
int launch_row_r2c_u8(xc_ctx *ctx, int NH, int nitems, const xc::R2cU8 *items, int rpt, int np, bool gen);
constexpr int R2C_U8_NP = 1;  // default work items per workgroup of the 8-bit row r2c
bool r2c_u8_length(int NH);  // row length NH has the 8-bit row r2c
template <typename T> int r2c_default_rpt(int NH);  // rows per thread of the default r2c shape
template <typename T> int r2c_alt_rpt(int NH);      // the other shape's rows per thread (0: none)
// cwg: columns per workgroup (0: the default shape); xcd: XCD-aware workgroup order.
template <typename T>
int launch_col(xc_ctx *ctx, int NY, bool whiten, int cwg, int xcd, int nitems, const PairDev *pd, const int2 *work);
bool col_shape_ok(const xc_ctx *ctx, int NY, int cwg);  // an FP32 column pass of cwg columns fits
// cnt: some NCC pair of the launch has a count rule (selects the kernel variant that scans the count tables;
// the FP64 form always uses the general per-lag evaluation for those).  tabk: the TabKind of the pairs' tables
// (TK_F32 and TK_X with FP32 transforms only).
template <typename T>
int launch_row_c2r(xc_ctx *ctx, int NH, int nblocks, const PairDev *pd, const int2 *work, bool cnt = false,
                   int tabk = TK_F64);
// The TK_X launches (their own translation unit).
int launch_row_c2r_x(xc_ctx *ctx, int NH, int nblocks, const PairDev *pd, const int2 *work, bool cnt);
template <typename T> int launch_finalize(xc_ctx *ctx, int npairs, const PairDev *pd, int tabk = TK_F64);
template <typename T> int row_group(int NH);   // rows per workgroup of the row kernels
template <typename T> int r2c_rows(int NH);    // rows per work item of the row r2c
template <typename T> int col_group(int NY);   // columns per work item of the column kernel

template <typename T> const cpx<T> *twiddles(xc_ctx *ctx);

// Largest supported transform lengths for precision T.
template <typename T> constexpr int max_nh() { return sizeof(T) == 8 ? 2048 : 4096; }
template <typename T> constexpr int max_ny() { return 4096; }

}  // namespace xc
