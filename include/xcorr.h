/*
 * xcorr.h -- C-ABI of the MI355X-native FFT tile-pair cross-correlation library
 * (libxcorr.so).  Plain C types only: pointers, sizes, POD structs.  No torch,
 * no C++ types, so it can be bound from C, C++, ctypes or any FFI.
 *
 * Each entry point names the reference interface it replaces
 * (billkarsh/Alignment_Projects, 0_GEN/Correlation.h and friends).  The C++
 * drop-in that keeps the reference's exact signatures (std::vector<Point>,
 * EvalType callbacks, vector<CD> caches) is declared in aln_corr_compat.h and
 * is implemented on top of this header.
 *
 * Conventions
 *  - Every entry point returns an int status (XC_OK == 0, warnings > 0,
 *    errors < 0).  These replace the reference's exit(42)/exit(44).
 *  - "lag (x,y)" means B-coordinate = A-coordinate + (x,y), as in the
 *    reference's CorrPatches.
 *  - Padded FFT sizes follow FFTSize(): Nx = CeilPow2(w1+w2-1),
 *    Ny = CeilPow2(h1+h2-1) (Correlation.cpp:344-347).
 *  - Host pointers are plain host memory unless the name says "dev".
 *  - A context owns one HIP stream; calls on one context are serialised by
 *    the context.  Distinct contexts may be used from distinct threads.
 */
#ifndef XCORR_H
#define XCORR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */

enum xc_status {
    XC_OK            = 0,
    XC_NO_PEAK       = 1,   /* no legal lag: CorrPatches returns 0.0 and leaves dx,dy untouched
                               (Correlation.cpp:1074-1087); CorrImages* set dx=dy=0 (:3112-3119) */
    XC_TOO_SMALL     = 2,   /* CCorImg::SetDims: a lag half-range < 6 (Correlation.cpp:1991-2000) */
    XC_E_ODD_R       = -44, /* |r| > 1.0001: the reference calls exit(44) (Correlation.cpp:731-751, 1331-1336) */
    XC_E_ARG         = -2,  /* malformed argument (null pointer, empty point list, bad enum) */
    XC_E_SHAPE       = -3,  /* padded size outside what the kernels support */
    XC_E_DEVICE      = -4,  /* HIP runtime error */
    XC_E_NOMEM       = -5,  /* device allocation failed */
    XC_E_BUDGET      = -6,  /* serving admission: request exceeds the HBM budget */
    XC_E_UNSUPPORTED = -7,  /* feature not available on this path */
    XC_E_TIMEOUT     = -8,  /* serving: request deadline passed before completion */
    XC_E_CANCELLED   = -9,  /* serving: request cancelled by the caller */
    XC_E_NOTREADY    = 3    /* serving poll: result not ready yet */
};

/* ------------------------------------------------------------------ */
/* POD types                                                           */
/* ------------------------------------------------------------------ */

/* Same layout as the reference's Point {double x, y;} (CPoint.h:14-27). */
typedef struct xc_point {
    double x, y;
} xc_point;

/*
 * Legal-overlap rule, the device form of an EvalType LegalRgn(olw, olh, arg)
 * callback (Correlation.h:91).  A lag is legal iff every enabled clause holds:
 *   olw >= min_w && olh >= min_h && (int64)olw*olh > min_area
 *   && (exact_w <= 0 || olw == exact_w) && (exact_h <= 0 || olh == exact_h)
 *   && (or_w_ge <= 0 || olw >= or_w_ge || olh > or_h_gt)
 *   && (table == 0 || table[(olh-1)*table_w + (olw-1)] != 0)
 * enabled == 0 is the NULL callback (everything legal).
 * Production callbacks:
 *   BigEnough    CThmScan.cpp:28-33   min_w=min_h=olap1D, min_area=reqArea
 *   SomeArea     testnormcorr.cpp:11  min_w=min_h=2, min_area=-1
 *   InSectionLegal InSectionOverlap.cpp:41-47  min_area=30000, or_w_ge=1500, or_h_gt=1500
 *   lr / lr31x31 mos.cpp:1270, Inspect.cpp:631  exact_w=exact_h=21 / 31
 * Any other callback depends only on (olw, olh) and can be tabulated by the
 * caller into `table` (host pointer, table_w*table_h bytes, olw<=table_w,
 * olh<=table_h; lags outside the table are illegal).
 */
typedef struct xc_rgn_rule {
    int32_t        enabled;
    int32_t        min_w, min_h;
    int32_t        exact_w, exact_h;
    int32_t        or_w_ge, or_h_gt;
    int32_t        table_w, table_h;
    int32_t        _pad;
    int64_t        min_area;
    const uint8_t *table;
} xc_rgn_rule;

/*
 * Legal-count rule, the device form of EvalType LegalCnt(c1, c2, arg), where
 * c1,c2 are the non-zero pixel counts of the two overlap boxes:
 *   c1 > k1 && c2 > k2.   enabled == 0 is the NULL callback.
 *   EnoughPoints CThmScan.cpp:39-42  k1=k2=reqArea
 *   lc / lcTRUE  mos.cpp:1282, Inspect.cpp:640 -> enabled = 0
 */
typedef struct xc_cnt_rule {
    int32_t enabled;
    int32_t _pad;
    int64_t k1, k2;
} xc_cnt_rule;

/* Result record of one correlation. */
typedef struct xc_result {
    double  r;          /* value the reference returns (NCC r, or the S-mode score) */
    double  dx, dy;     /* displacement to add to A's points (reference output params) */
    double  peak;       /* raw value at the integer peak (NCC r for R/NCC, S for S-mode) */
    double  sub_x;      /* parabolic subpixel offsets that were applied */
    double  sub_y;
    int32_t status;     /* enum xc_status */
    int32_t lag_x;      /* integer peak lag (signed, before parabola and bbox offsets) */
    int32_t lag_y;
    int32_t nx, ny;     /* padded FFT size */
    int32_t _pad;
} xc_result;

/* Correlator modes (CorrImagesR / CorrImagesS, Correlation.cpp:3144, 3222). */
enum xc_mode {
    XC_MODE_NCC = 0,    /* CorrPatches: normalised cross-correlation, whole-grid wrapped scan */
    XC_MODE_R   = 1,    /* CorrImagesR: centred R image, A-mask, straight max */
    XC_MODE_S   = 2,    /* CorrImagesS: whitened cross-power S image, score 10(r-mn)/(sd sqrt(n)) */
    XC_MODE_F   = 3     /* CorrImagesF: R image, LoG+I filtered F image, best well-isolated F peak */
};

/* ------------------------------------------------------------------ */
/* Context                                                             */
/* ------------------------------------------------------------------ */

typedef struct xc_ctx xc_ctx;

/* Create a context on HIP device `device` (one stream, lazily grown workspace). */
int xc_ctx_create(int device, xc_ctx **out);
int xc_ctx_destroy(xc_ctx *ctx);
/* Last error message of this context (static storage, never NULL). */
const char *xc_last_error(const xc_ctx *ctx);
/* Library version string. */
const char *xc_version(void);

/* ------------------------------------------------------------------ */
/* Plain FP64 path: one pair per call, parity with the reference FFTW   */
/* path (integer peak bit-exact, r within 1e-9 relative).              */
/* ------------------------------------------------------------------ */

/* FFTSize (Correlation.cpp:344) and FFTSizeSQR (:356). */
int xc_fft_size(int n1, int n2);
int xc_fft_size_sqr(int w1, int h1, int w2, int h2);

/*
 * Replaces FFT_2D (Correlation_fft_fftw.cpp:38-56): forward 2-D r2c of the
 * row-major real array in[Nslow][Nfast], unnormalised, into
 * out[Nslow][Nfast/2+1] complex (interleaved re,im doubles).
 * Returns XC_OK; *M_out (may be NULL) receives Nslow*(Nfast/2+1).
 */
int xc_fft2d_r2c_f64(xc_ctx *ctx, const double *in, int Nfast, int Nslow,
                     double *out, int *M_out);

/* Replaces IFT_2D (Correlation_fft_fftw.cpp:67-92): unnormalised 2-D c2r. */
int xc_ift2d_c2r_f64(xc_ctx *ctx, const double *in, int Nfast, int Nslow,
                     double *out);

/*
 * Replaces CorrPatches (Correlation.cpp:901-1170).  Points/values are host
 * arrays.  On XC_OK, *dx,*dy receive the displacement and *r the maximum
 * NCC; on XC_NO_PEAK *r = 0 and *dx,*dy are left untouched, exactly as the
 * reference does.  `res` (may be NULL) receives the full record.
 * The padded raster is built on the device (bilinear splat of the points).
 */
int xc_corr_patches_f64(xc_ctx *ctx,
                        const xc_point *ip1, const double *iv1, int n1,
                        const xc_point *ip2, const double *iv2, int n2,
                        int Ox, int Oy, int radius,
                        const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                        double *dx, double *dy, double *r, xc_result *res);

/*
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  xc_corr_patches_f64 with the same arguments, kernels and
 * results, but it allocates nothing: no device or pinned memory, no library
 * routine that allocates.  Its device workspace is a fixed array of
 * XC_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends; the lists are copied from the caller's memory.
 * A pair whose workspace (lists, rasters, FP64 tables, spectra, lag rows)
 * exceeds it is refused with XC_E_SHAPE before any device work.  Every call
 * on one GPU shares that GPU's array and the array keeps the last call's
 * contents (each call clears what must start at zero); concurrent calls on
 * one GPU would race for it, so the library runs them one at a time.
 */
#define XC_NOALLOC_SCRATCH_BYTES (128ULL << 20)
int xc_corr_patches_noalloc_f64(xc_ctx *ctx,
                                const xc_point *ip1, const double *iv1, int n1,
                                const xc_point *ip2, const double *iv2, int n2,
                                int Ox, int Oy, int radius,
                                const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                                double *dx, double *dy, double *r, xc_result *res);

/*
 * Replaces CorrPatchToImage (Correlation.cpp:1186-1355): patch1 (integer
 * points) against the caller-embedded square N x N image i2 (N = sqrt(n_i2)).
 * bFilter applies the reference's Gaussian low-pass of the cross power.
 */
int xc_corr_patch_to_image_f64(xc_ctx *ctx,
                               const xc_point *ip1, const double *iv1, int n1,
                               const double *i2, int n_i2,
                               int Ox, int Oy, int radius, int bFilter,
                               double *dx, double *dy, double *r, xc_result *res);

/*
 * Replaces CorrImagesF / CorrImagesR / CorrImagesS (Correlation.cpp:
 * 3066-3122, 3144-3290) with mode XC_MODE_F / XC_MODE_R / XC_MODE_S.
 * mincor and nbmaxht keep their reference meaning (nbmaxht is used by
 * F mode only: a candidate F peak is rejected when its guard band holds a
 * pixel above nbmaxht times its height).  Rx<=0 or Ry<=0 disables the oval
 * search mask.  On XC_TOO_SMALL / XC_NO_PEAK, dx=dy=0 and r=0.
 */
int xc_corr_images_f64(xc_ctx *ctx, int mode,
                       const xc_point *ip1, const double *iv1, int n1,
                       const xc_point *ip2, const double *iv2, int n2,
                       const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                       double mincor, double nbmaxht,
                       int Ox, int Oy, int Rx, int Ry,
                       double *dx, double *dy, double *r, xc_result *res);

/* ------------------------------------------------------------------ */
/* Device-resident tiles (the fft2 / ThmRec.ftc cache, made explicit).  */
/* ------------------------------------------------------------------ */

typedef struct xc_tile xc_tile;

enum xc_dtype { XC_U8 = 0, XC_U16 = 1, XC_F32 = 2, XC_F64 = 3 };

enum xc_tile_flags {
    XC_TILE_NORMALIZE = 1,   /* Normalize() the valid pixels: mean 0, sd 1 (n-1), Maths.cpp:192-212 */
    XC_TILE_DEVICE_SRC = 2,  /* `data`/`mask` are device pointers (else host) */
    XC_TILE_F32_TABLES = 4,  /* keep the summed-area tables as FP32 (S, S^2) entries (each the FP64 sum
                                rounded once): half the table bytes, for callers that accept the rounding;
                                such tiles pair only with tiles of FP32 tables */
    XC_TILE_F64_TABLES = 8   /* keep FP64 (S, S^2) entries where exact integer tables would apply */
};

/*
 * Upload (or adopt from device memory) one w x h image with optional
 * validity mask (non-zero = pixel present in the point list; absent pixels
 * are 0 in the raster, as ImageFromValuesAndPoints leaves them).  `stride`
 * is the picture's row pitch in elements (0: w).  A host mask has rows of w
 * bytes; a device mask (XC_TILE_DEVICE_SRC) has rows `stride` bytes apart, so
 * that a caller can pad both to 16-byte multiples, which the table sweep
 * reads fastest (uploaded 8-bit pictures and masks are stored that way).  The tile
 * keeps its pixels, (mean, sd) and its summed-area tables of S and S^2
 * (IntegrateImage, Correlation.cpp:49-105) resident in HBM, plus the int32
 * non-zero count table.  The tables' form:
 *   - an unmasked XC_U8 tile of at most 2^20 pixels keeps exact integer
 *     sums, 8 bytes a pixel (the raw values' sum < 2^28 and their squares'
 *     sum < 2^36 packed in one uint64): every box sum is exact, and GetCorr's
 *     variance terms n q - s^2 = (n Qraw - Sraw^2) / sd^2 are formed in
 *     64-bit integers, so they carry no rounding at all until the final
 *     scaling -- the FP64 tables' content without their rounding;
 *   - any other tile (or one flagged XC_TILE_F64_TABLES) keeps interleaved
 *     FP64 pairs, the reference's double tables, 16 bytes a pixel;
 *   - XC_TILE_F32_TABLES keeps FP32 pairs (each the FP64 sum rounded once).
 * A plan pairs exact and FP64 tables by deriving an FP64 table of the exact
 * tile (also for a rotated A, whose copy has FP64 tables); the server needs
 * both tiles of a request in one form (else XC_E_ARG).  The tile's raster is the bounding box of its
 * present pixels, as the reference's raster spans the bounding box of the
 * point list (BBoxFromPoints, Geometry.cpp:78-106): a mask whose edge rows
 * or columns are empty shrinks it; displacements stay in the picture's
 * coordinates.  XC_E_ARG for a mask with no present pixel.
 */
int xc_tile_put(xc_ctx *ctx, const void *data, int dtype, int w, int h, int stride,
                const uint8_t *mask, int flags, xc_tile **out);
int xc_tile_release(xc_ctx *ctx, xc_tile *tile);
/* Bytes of device memory the tile holds (values, tables, cached spectra). */
int64_t xc_tile_bytes(const xc_tile *tile);

/* ------------------------------------------------------------------ */
/* Batched FP32 path                                                   */
/* ------------------------------------------------------------------ */

/*
 * One pair correlation of a batch.  mode XC_MODE_NCC follows CorrPatches
 * (disc Ox,Oy,radius), XC_MODE_S follows CorrImagesS (oval Ox,Oy,Rx,Ry).
 * angle_deg != 0 rotates A's pixel grid about the origin before the
 * correlation, as CThmScan::RFromAngle does (CThmScan.cpp:329-366).
 */
typedef struct xc_pair {
    const xc_tile *a;
    const xc_tile *b;
    int32_t        mode;
    int32_t        Ox, Oy;
    int32_t        radius;      /* NCC disc radius */
    int32_t        Rx, Ry;      /* S/R oval semi-axes (<= 0: none) */
    double         angle_deg;
    xc_rgn_rule    rgn;
    xc_cnt_rule    cnt;
} xc_pair;

/*
 * Correlate n pairs in FP32 (FFTs, multiply) with the tiles' summed-area
 * tables (exact integer or FP64 sums; FP32 ones for tiles made with
 * XC_TILE_F32_TABLES) and FP64 (exact-integer for exact tables) variance
 * terms; the final num / sqrt(d) of GetCorr is taken in FP32.  Blocks until
 * results[0..n) are written.
 */
int xc_batch_corr_f32(xc_ctx *ctx, const xc_pair *pairs, int n, xc_result *results);

/*
 * Batch plans: a fixed batch of pairs laid out once (descriptors, work
 * lists, spectra, lag rows) and replayed as one HIP graph.  This is the
 * batched counterpart of the reference's per-thread CorrPatches calls
 * (CThmScan::TCDGet, CThmScan.cpp:77-107), without per-call planning.
 * XC_PLAN_REFRESH_TILES re-runs the tiles' statistics and summed-area tables
 * inside every run (for tiles whose pixels are rewritten between runs).
 * XC_PLAN_AUTOTUNE times the candidate shapes of the column pass (columns
 * per workgroup 1..16, XCD-aware workgroup order on/off) and of the row r2c
 * (two rows per thread, or one row per thread group with twice the groups)
 * on each of the plan's size buckets and keeps the fastest; choices are
 * cached per context (and in the file named by XC_TUNE_FILE, if set).  Shapes
 * never change results, only speed; the radix schedule is not tuned (another
 * order of passes would change the rounding).  (The reference plans each FFT per call with
 * FFTW_ESTIMATE, Correlation_fft_fftw.cpp:20-26.)
 */
typedef struct xc_plan xc_plan;

enum xc_plan_flags { XC_PLAN_REFRESH_TILES = 1, XC_PLAN_AUTOTUNE = 2 };

/* One size bucket of a plan: transform shape, pairs, the shapes of its launches. */
typedef struct xc_bucket_info {
    int32_t nh, ny, whiten, pairs;
    int32_t col_c, col_xcd;  /* columns per workgroup (0: default), XCD order */
    double col_us;           /* tuned column-pass time, microseconds (0: not tuned) */
    int32_t r2c_rpt;         /* rows per thread of the row r2c */
    int32_t r2c_np;          /* work items per workgroup of the 8-bit row r2c (0: the general r2c) */
    double r2c_us;           /* tuned row r2c time, microseconds (0: not tuned) */
} xc_bucket_info;

/* Pipeline stages reported by xc_plan_stage_times. */
enum xc_stage { XC_STAGE_PREP = 0, XC_STAGE_R2C = 1, XC_STAGE_COL = 2, XC_STAGE_C2R = 3, XC_STAGE_FIN = 4,
                XC_STAGE_COUNT = 5 };

int xc_plan_create(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, xc_plan **out);
/* Launch the plan's graph on the context stream (asynchronous). */
int xc_plan_launch(xc_plan *plan);
/* Wait for the last launch; copy results[n] back if results != NULL. */
int xc_plan_wait(xc_plan *plan, xc_result *results);
/* launch + wait. */
int xc_plan_run(xc_plan *plan, xc_result *results);
/* Run each stage separately `iters` times between HIP events on the
 * context stream; ms[XC_STAGE_COUNT] receives the mean duration of each
 * stage, bytes[XC_STAGE_COUNT] its algorithmic HBM bytes per run. */
int xc_plan_stage_times(xc_plan *plan, int iters, double *ms, double *bytes);
/* Algorithmic HBM bytes of one run of the whole plan (inputs once; row
 * spectra, lag rows and the tables, 8 B per exact integer entry, 16 B per
 * FP64 (S, S^2) entry or 8 B per FP32 one, written once and read once). */
double xc_plan_model_bytes(const xc_plan *plan);
/* Copies up to cap bucket records; returns the plan's bucket count. */
int xc_plan_buckets(const xc_plan *plan, xc_bucket_info *out, int cap);
/* Set a bucket's launch shapes by hand (the autotuner's candidates): column
 * pass columns per workgroup (0: default; 1, 2, 4, 8, 16 where the device's
 * LDS holds them) and XCD order, row r2c rows per thread (0: default; 1 or 2
 * for FP32 rows of 512 points or more whose work item holds two rows, else
 * only the default), and for a bucket on the 8-bit row r2c (unmasked u8
 * images) its work items per workgroup (0: default; 1, 2 or 4; ignored
 * otherwise).  xc_plan_buckets reports the shapes in use.  Shapes change
 * speed only, never results.  Waits for a launch in flight, then records the
 * plan's graph again.  XC_E_SHAPE for a shape the bucket does not have. */
int xc_plan_set_shapes(xc_plan *plan, int bucket, int col_c, int col_xcd, int r2c_rpt, int r2c_np);
/* The share of each CU that a plan's table sweeps may hold while they run
 * beside the row transforms (plans that prepare their tiles or sweep rotated
 * copies in every run): at most `waves` wavefronts per CU, the rest of the
 * CU's wave slots left to the transforms; 0 = one workgroup per tile, no
 * limit.  The default is 16; with XC_PLAN_AUTOTUNE the plan times replays of
 * its whole graph with 0, 16 and 8 and keeps the fastest (*us: that time in
 * microseconds, 0 when not tuned).  Changes speed only, never results.  (No
 * counterpart in the reference, which prepares tables per call on the CPU.) */
int xc_plan_sweep_waves(const xc_plan *plan, int32_t *waves, double *us);
/* Set the share by hand (0..32, else XC_E_ARG); waits for a launch in
 * flight, then records the plan's graph again. */
int xc_plan_set_sweep_waves(xc_plan *plan, int waves);
int xc_plan_destroy(xc_plan *plan);

/*
 * Rotation sweep (CThmScan::AngleScan*, CThmScan.cpp:431-621): correlate
 * A rotated by each of n_angles angles against B in one batched launch, in
 * S mode with the given legality rules.  results[n_angles].
 */
int xc_sweep_angles_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b,
                        const double *angles_deg, int n_angles,
                        const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                        int Ox, int Oy, int Rx, int Ry, xc_result *results);

/* ------------------------------------------------------------------ */
/* Angle search around the rotation sweep (CThmScan)                   */
/* ------------------------------------------------------------------ */

/* One evaluated angle: CorRec (CThmScan.h:36-44) without its affine:
 * A = angle (deg), R = CorrImagesS' score, X, Y = its displacement. */
typedef struct xc_angle_rec {
    double A, R, X, Y;
} xc_angle_rec;

enum xc_scan_mode {
    XC_SCAN_MAXR    = 0,  /* AngleScanMaxR (CThmScan.cpp:431-505): largest R, the centre first */
    XC_SCAN_CONSTXY = 1   /* AngleScanConstXY (:514-621): by decreasing R, the first angle whose
                           * +-3 neighbours' rounded X and Y both fit a line with |lincor| >= 0.7 */
};

/*
 * CThmScan::RFromAngle (CThmScan.cpp:329-366) in FP64: A's points rotated
 * by deg (TAffine::NUSetRot with PI = 3.14159265358, Apply_R_Part), then
 * CorrImagesS against B with BigEnough(olap1D, req_area) and
 * EnoughPoints(req_area), oval (Ox, Oy, Rx, Ry).  rec->R = 0 and X = Y = 0
 * when CorrImagesS finds no peak.  Returns XC_OK or an error.
 */
int xc_rfromangle_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                      const xc_point *bp, const double *bv, int nb, double deg,
                      int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                      xc_angle_rec *rec);

/*
 * AngleScanMaxR / AngleScanConstXY (mode): RFromAngle over
 * a = center - hlfwid, += step, while a <= center + hlfwid; MAXR with
 * favor_center first evaluates the centre and a band of step/3 within
 * 2 step of it (and of centre + 180 when hlfwid >= 180).  *best receives
 * the selection (CONSTXY: R = A = X = Y = 0 when no angle qualifies; exact
 * ties of R keep the earlier angle).  recs (may be NULL) receives up to cap
 * evaluated angles in evaluation order, *n_recs their count.
 * The f64 form evaluates each angle with xc_rfromangle_f64 (the
 * reference's arithmetic); the f32 form evaluates all angles in one
 * batched sweep of tiles (xc_sweep_angles_f32).
 */
int xc_angle_scan_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                      const xc_point *bp, const double *bv, int nb, int mode,
                      double center, double hlfwid, double step, int favor_center,
                      int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                      xc_angle_rec *best, xc_angle_rec *recs, int cap, int *n_recs);
int xc_angle_scan_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, int mode,
                      double center, double hlfwid, double step, int favor_center,
                      int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                      xc_angle_rec *best, xc_angle_rec *recs, int cap, int *n_recs);

/*
 * The same sweep over general 2 x 2 maps: maps[4 i .. 4 i + 3] = {t0, t1, t3, t4}
 * is the linear part applied to A's points (RFromAngle's C.T = R * (Tdfm * T),
 * CThmScan.cpp:348-349, with a deformation or pretweak), splatted by the
 * inverse-map gather; an identity map uses A as it is.  XC_E_ARG for a
 * singular map.
 */
int xc_sweep_affine_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *maps,
                        int n_maps, const xc_rgn_rule *rgn, const xc_cnt_rule *cnt,
                        int Ox, int Oy, int Rx, int Ry, xc_result *results);

/*
 * CThmScan::PeakHunt (CThmScan.cpp:676-732): bracket search for the best
 * angle within best->A +- hlfwid, probing the left then the right half
 * until the bracket is narrower than 1e-4 deg; the entry *best is kept if
 * nothing beats it.  *n_evals receives the reference's probe counter k.
 */
int xc_peak_hunt_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                     const xc_point *bp, const double *bv, int nb, double hlfwid,
                     int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                     xc_angle_rec *best, int *n_evals);
int xc_peak_hunt_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, double hlfwid,
                     int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                     xc_angle_rec *best, int *n_evals);

/*
 * RFromAngle with a deformation and a pretweak (CThmScan.cpp:329-366):
 * A's points mapped by the 2x2 part of R(deg) * (Tdfm * Ttw), with the
 * affines as TAffine's t[6] (t0 t1 X t3 t4 Y) and products as
 * TAffine::operator* (TAffine.cpp:220-232).  Tdfm or Ttw NULL = identity.
 */
int xc_rfromangle_affine_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                             const xc_point *bp, const double *bv, int nb, double deg,
                             const double *Tdfm, const double *Ttw,
                             int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                             xc_angle_rec *rec);

/*
 * CThmScan::Pretweaks (CThmScan.cpp:205-323) in FP64: up to five rounds; in
 * each, sweep every unused near-unity transform type (scale, x-scale,
 * y-scale, x-skew, y-skew; TAffine::NUSelect, TAffine.cpp:76-135) over
 * base +- 5 x 0.02 at angle deg, refine the round's best magnitude with
 * PTWInterp's parabola (:119-160), and compose it into Tptwk (in/out, t[6])
 * when it beats *bestR (in/out).  *changed = 1 when Tptwk changed;
 * *n_evals receives the number of RFromAngle evaluations.
 */
int xc_pretweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                     const xc_point *bp, const double *bv, int nb, double deg,
                     const double *Tdfm, double *Tptwk, double *bestR,
                     int olap1D, int64_t req_area, int Ox, int Oy, int Rx, int Ry,
                     int *changed, int *n_evals);
/* The same over resident tiles: each round's candidate maps (5 x 11, then
 * fewer) in one batched FP32 sweep (xc_sweep_affine_f32), then the
 * parabola's two neighbours in one and its vertex; the decisions are the
 * FP64 form's, taken on the FP32 R values. */
int xc_pretweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, double deg,
                     const double *Tdfm, double *Tptwk, double *bestR, int olap1D,
                     int64_t req_area, int Ox, int Oy, int Rx, int Ry, int *changed,
                     int *n_evals);

/*
 * CThmScan::FinishAtFullRes (CThmScan.cpp:810-842): RFromAngle at best->A
 * on the full-resolution point lists with the oval ((int)X, (int)Y, 80, 80)
 * about the thumbnail offset; when the new X and Y both lie within 20 of
 * it they replace it (R stays the thumbnail R), else *best is kept.
 * *used receives 1 or 0.
 */
int xc_finish_at_full_res_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                              const xc_point *bp, const double *bv, int nb,
                              const double *Tdfm, const double *Tptwk,
                              int olap1D, int64_t req_area, xc_angle_rec *best, int *used);
/* The same over full-resolution resident tiles, one FP32 sweep of one map. */
int xc_finish_at_full_res_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *Tdfm,
                              const double *Tptwk, int olap1D, int64_t req_area, xc_angle_rec *best,
                              int *used);

/*
 * CThmScan's per-angle search-oval hook (T_NewAngProc, CThmScan.h:84,
 * 143-144): when set, RFromAngle asks it for the oval's centre and
 * semi-axes at each angle (in degrees) instead of using the disc
 * (CThmScan.cpp:351-355); 1_Scapeops sets StripAngProc (scapeops.cpp:
 * 476-491, 542), which centres the oval on the rotated strip.  `arg` is
 * handed back unchanged (the reference's hook reads globals instead).
 */
typedef void (*xc_ang_proc)(int *Ox, int *Oy, int *Rx, int *Ry, double deg, void *arg);

/*
 * CThmScan's matcher state for one pair (CThmScan.h:85-97): thresholds,
 * sweep switches, the disc, the two affines and the per-angle oval hook.
 * The entry points below read it and update it as the reference's methods
 * update their object: Pretweaks folds into Tptwk, PostTweaks sets the
 * disc. n_evals counts the RFromAngle evaluations made (the caller sets
 * it; each call adds). xc_thm_params_default gives the constructor's
 * values (CThmScan.cpp:160-177): rthresh 0.30, ConstXY and Pretweak sweeps
 * on, no disc (0, 0, -1, -1), identity affines, no hook.
 */
typedef struct xc_thm_params {
    double  rthresh;
    int32_t const_xy;     /* swpConstXY: AngleScanConstXY, else AngleScanMaxR */
    int32_t pretweak;     /* swpPretweak */
    int32_t olap1D;
    int32_t Ox, Oy, Rx, Ry;
    int32_t n_evals;
    int64_t req_area;
    double  Tdfm[6];
    double  Tptwk[6];
    xc_ang_proc ang_proc; /* newAngProc (NULL: the disc above for every angle) */
    void   *ang_arg;
} xc_thm_params;

void xc_thm_params_default(xc_thm_params *p);

/* AngleScanWithTweaks (CThmScan.cpp:647-672): AngleScanSel(favorCenter);
 * below rthresh and with the Pretweak sweep on, Pretweaks at best->A (or
 * center when best->R <= 0) and, when that changed Tptwk, the scan again. */
int xc_thm_scan_with_tweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                double center, double hlfwid, double step, xc_angle_rec *best);
/* UsePriorAngles (CThmScan.cpp:734-758): the tweaked scan about ang0 at
 * step 0.1, then PeakHunt(0.15); *ok = 0 when either R < rthresh. */
int xc_thm_use_prior_angles_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                double ang0, double hfangpr, xc_angle_rec *best, int *ok);
/* DenovoBestAngle (CThmScan.cpp:764-790): the tweaked scan, AngleScanSel
 * about best->A over +-2 step at step/20, PeakHunt(0.075 step); *ok = 0
 * at the first R < rthresh (the later stages are then skipped). */
int xc_thm_denovo_best_angle_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                                 const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                                 double ang0, double hfangdn, double step, xc_angle_rec *best,
                                 int *ok);
/* PostTweaks (CThmScan.cpp:796-802): disc ((int)X, (int)Y, 40, 40),
 * Pretweaks at best->A, RFromAngle at best->A. */
int xc_thm_post_tweaks_f64(xc_ctx *ctx, const xc_point *ap, const double *av, int na,
                           const xc_point *bp, const double *bv, int nb, xc_thm_params *p,
                           xc_angle_rec *best);

/* The same sequence over resident tiles in FP32: every scan's angles (with
 * the current Tdfm and Tptwk) and every Pretweaks round are one batched
 * sweep (xc_sweep_affine_f32); PeakHunt's probes stay one at a time. */
int xc_thm_scan_with_tweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                double center, double hlfwid, double step, xc_angle_rec *best);
int xc_thm_use_prior_angles_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                double ang0, double hfangpr, xc_angle_rec *best, int *ok);
int xc_thm_denovo_best_angle_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                                 double ang0, double hfangdn, double step, xc_angle_rec *best,
                                 int *ok);
int xc_thm_post_tweaks_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, xc_thm_params *p,
                           xc_angle_rec *best);

/* ------------------------------------------------------------------ */
/* Thumbnails                                                          */
/* ------------------------------------------------------------------ */

/*
 * DecimateVector (Maths.cpp:1315-1364), in place on host arrays: the n
 * points (bounding-box corner at (0, 0), inside w x h) summed in point order
 * into lbin x lbin cells and counted in an 8-bit counter (a cell of 256
 * points wraps to 0 and is dropped, as in the reference); the cells with a
 * non-zero count, in cell order, replace the list as (cell x, cell y) and
 * their mean.  *n_out receives the new length.  XC_E_ARG for a point outside
 * the box.
 */
int xc_decimate_vector_f64(xc_ctx *ctx, xc_point *p, double *v, int n, int w, int h, int lbin,
                           int *n_out);

/*
 * CThmUtil::MakeThumbs (1_DMesh/CThmUtil.cpp:470-523) on both sides, in place:
 * DecimateVector by decfactor (when > 1; *olap1D and *req_area are divided as
 * the reference divides them), then Normalize each value list.  *na_out and
 * *nb_out receive the lengths.  XC_TOO_SMALL when A keeps fewer than
 * *req_area points, XC_NO_PEAK when a side's sd is 0 (the reference's two
 * failures).
 */
int xc_make_thumbs_f64(xc_ctx *ctx, xc_point *ap, double *av, int na, int aw, int ah,
                       xc_point *bp, double *bv, int nb, int bw, int bh, int decfactor,
                       int *olap1D, int64_t *req_area, int *na_out, int *nb_out);

/*
 * The thumbnail of a whole device picture (every pixel a point, raster order):
 * lbin x lbin box means (DecimateVector's values, cells of 1..15 pixels so the
 * count never wraps) into out[ceil(h/lbin)][ceil(w/lbin)] (device, FP64), then
 * Normalize; *sd (may be NULL) receives the natural sd.  dtype: XC_U8, XC_U16,
 * XC_F32, XC_F64; stride in pixels.
 */
int xc_make_thumb_raster_dev_f64(xc_ctx *ctx, const void *ras, int dtype, int w, int h, int stride,
                                 int lbin, double *out, double *sd);

/* One side of the overlap CThmUtil::Crop hands to MakeThumbs (SubI, 1_DMesh/CThmUtil.h:14-20). */
typedef struct xc_subimage {
    double  Ox, Oy;  /* S.O: the subimage's corner in the picture */
    int32_t w, h;    /* S.w, S.h */
    int64_t n;       /* points (and values) written */
} xc_subimage;       /* 32 bytes */

/*
 * CThmUtil::Crop (1_DMesh/CThmUtil.cpp:385-430) with GetOlapBoxes (:352-373),
 * SubI_ThisBox (:195-235), SubI_ThesePoints (:163-189) and Olap_WholeImage
 * (:243-258).  av, bv: the two ws x hs FP64 pictures in device memory (the
 * scaled pictures of xc_pixpair_prep_dev_f64); apts / bpts: the connected
 * regions' points (host).  (tx, ty): a's origin in b's system (the
 * translation of Tab); the box offset is int(xyconf tx / scl), int(xyconf
 * ty / scl).  With xyconf = 0, or an overlap box narrower or lower than
 * max(olap1d, 8), every point is kept (*whole = 1); otherwise the points
 * inside each overlap box (*whole = 0).  The kept points, relative to their
 * bounding box's corner, go to ap_out / bp_out and their picture values,
 * gathered on the device, to av_out / bv_out (room for na and nb).
 * XC_TOO_SMALL when a box keeps olap2d points or fewer (the reference's
 * false; A is tried first and B is then not written); XC_E_ARG for an empty
 * list or a point whose truncated (x, y) lies outside the picture.
 */
int xc_crop_f64(xc_ctx *ctx, const double *av, const double *bv, int ws, int hs,
                const xc_point *apts, int na, const xc_point *bpts, int nb,
                double tx, double ty, double xyconf, double scl, int olap1d, int64_t olap2d,
                xc_point *ap_out, double *av_out, xc_subimage *a,
                xc_point *bp_out, double *bv_out, xc_subimage *b, int *whole);

/*
 * CThmUtil::Crop_NoCR (1_DMesh/CThmUtil.cpp:437-464) with
 * Olap_TheseBoxes_NoCR (:262-312) and Olap_WholeImage_NoCR (:318-346): the
 * same boxes without connected regions; the values of the overlap box (or of
 * the whole pictures), row by row, are copied from the device pictures to
 * av_out / bv_out (room for ws * hs each).  The points are the box's
 * MakeZeroBasedPoints (CPoint.cpp:82-95), written to p_out when it is not
 * NULL (room for ws * hs; the same list for both sides).
 */
int xc_crop_nocr_f64(xc_ctx *ctx, const double *av, const double *bv, int ws, int hs,
                     double tx, double ty, double xyconf, double scl, int olap1d,
                     double *av_out, double *bv_out, xc_point *p_out,
                     xc_subimage *a, xc_subimage *b, int *whole);

/* ------------------------------------------------------------------ */
/* PixPair::Load's raster preparation (the pair's pre-path)            */
/* ------------------------------------------------------------------ */

/*
 * One picture of PixPair::Load (CPixPair.cpp:276-460) from its 8-bit
 * raster (w x h, row-major): LegPolyFlatten (Maths.cpp:1006-1086) of the
 * given order (0..6) and Normalize -> vf; with dog, Convolve by
 * MakeDoGKernel(r1, r2) (CPixPair.cpp:36-80; symmetric, pre-normalised by
 * its sum) and Normalize -> vfflt; while w or h exceeds 2048 both halve
 * (scl doubles), and vs / vsflt receive the scl x scl box means of vf /
 * vfflt (PixPair::Downsample, CPixPair.cpp:90-113) -- written only when
 * scl > 1 (the reference then uses vf / vfflt as they are).  *ws, *hs, *scl
 * receive the downsampled geometry.  XC_E_ARG: order outside 0..6, r1 < 1,
 * r2 outside 1..18, or w, h not multiples of scl (the reference's
 * "Dimensions not multiple of scale" failure).
 * The _f64 form takes and returns host arrays (any output may be NULL);
 * the _dev_f64 form takes device pointers (vf required, vfflt required
 * with dog) and synchronises the context stream before returning.
 */
int xc_pixpair_prep_f64(xc_ctx *ctx, const uint8_t *ras, int w, int h, int order,
                        int dog, int r1, int r2, double *vf, double *vfflt,
                        double *vs, double *vsflt, int *ws, int *hs, int *scl);
int xc_pixpair_prep_dev_f64(xc_ctx *ctx, const uint8_t *ras, int w, int h, int order,
                            int dog, int r1, int r2, double *vf, double *vfflt,
                            double *vs, double *vsflt, int *ws, int *hs, int *scl);

/* LegPolyFlatten for 8-bit rasters alone (Maths.cpp:1006-1086, with its
 * final Normalize): host arrays, any w x h, order 0..6. */
int xc_leg_poly_flatten_f64(xc_ctx *ctx, const uint8_t *src, int w, int h, int order, double *vals);

/* ------------------------------------------------------------------ */
/* Fourier QA metric                                                   */
/* ------------------------------------------------------------------ */

typedef struct xc_fourier_qa {
    double  dot;         /* the metric: norm-dot of the low-frequency coefficients */
    double  energy;      /* its normaliser, sum sqrt(|v1|^2 |v2|^2) */
    double  msd;         /* MeanSqrDiff: mean of (av - bv)^2 */
    double  tot_power;   /* difference spectrum: total power */
    double  half_frac;   /* cumulative fraction at the half-power bin */
    int32_t half_index;  /* first bin count whose cumulative power reaches half */
    int32_t maxf;        /* wavelength bins */
    int32_t deg;         /* SmallestFootprint's rotation (TightestBBox) */
    int32_t nx, ny;      /* the metric images' size */
    int32_t _pad;
} xc_fourier_qa;

/*
 * FourierMatch (Metrics.cpp:517-658): rotate the region to its smallest
 * footprint (TightestBBox, Geometry.cpp:187-267: whole degrees -45..45),
 * splat av and bv into CeilPow2 x CeilPow2 images (MakeMetricImagesFFT,
 * Metrics.cpp:102-193), and return in *dot the normalised dot product of
 * their Fourier coefficients with |kx| <= Nx/wvlen and |ky| <= Ny/wvlen,
 * looked up as FFT_r2c_lookup does (Metrics.cpp:420-435: negative kx
 * read as the conjugate at (-kx, ky)).  qa (may be NULL) also receives the
 * figures the reference logs: the mean square difference, and the
 * difference image's power by integer wavelength, with the bin where the
 * cumulative power first reaches half.  n == 0 gives *dot = 0 (the
 * reference's "No data?").  Host arrays; wvlen >= 2 (1 reads past the
 * spectrum in the reference); images up to 4096 x 4096.
 * The wavelength bins are summed with device atomics: their last bits
 * depend on the order the additions land in.
 */
int xc_fourier_match_f64(xc_ctx *ctx, const xc_point *pts, const double *av, const double *bv,
                         int n, int wvlen, double *dot, xc_fourier_qa *qa);

/* ------------------------------------------------------------------ */
/* ThmPair tables (the thumbnail matcher's result file)                */
/* ------------------------------------------------------------------ */

/* One entry of a ThmPair_za^zb.txt table (ThmPair, PipeFiles.h:145-152):
 * tiles, error code, angle (deg), R, and the affine T[6] (T0 T1 X T3 T4 Y). */
typedef struct xc_thmpair {
    int32_t atl, acr, btl, bcr, err, _pad;
    double A, R;
    double T[6];
} xc_thmpair;

/* WriteThmPairHdr (PipeFiles.cpp:1284-1289): create or truncate `path`
 * with the header line.  No device needed. */
int xc_thmpair_create(const char *path);
/* WriteThmPair (PipeFiles.cpp:1319-1355): append one entry in the
 * reference's format ("%d\t" x 5, then "%f\t" x 8 with six decimals) in
 * one write; an entry with atl < 0 or btl < 0 is skipped, as there. */
int xc_thmpair_append(const char *path, const xc_thmpair *e);
/* ReadAllThmPair (PipeFiles.cpp:1200-1262): every entry after the header.
 * Returns the entry count (entries beyond cap are counted, not stored) or
 * XC_E_ARG when the file cannot be opened or has no header line. */
int xc_thmpair_read_all(const char *path, xc_thmpair *out, int cap);
/* ReadThmPair (PipeFiles.cpp:1132-1195): the first entry for (atl, acr,
 * btl, bcr): XC_OK, XC_NO_PEAK when there is none, XC_E_ARG as above. */
int xc_thmpair_find(const char *path, int atl, int acr, int btl, int bcr, xc_thmpair *out);
/* The mode-Y starting angle of CThmUtil::SetStartingAngle (1_DMesh/
 * CThmUtil.cpp:80-133): the median (MedianVal, Maths.cpp:436-449) of the
 * error-free entries' angles when there are at least 4, with |ang0| < 0.001
 * forced to 0.  Returns that count, or 0 (no prior; *ang0 untouched). */
int xc_thmpair_prior_angle(const xc_thmpair *e, int n, double *ang0);

/* ------------------------------------------------------------------ */
/* Serving loop: continuous batching of pair requests                 */
/* ------------------------------------------------------------------ */

/*
 * A server owns a scheduler thread, `streams` HIP streams and one device
 * arena of `hbm_budget` bytes.  Requests (xc_pair over resident tiles, no
 * rotation) are admitted in arrival order while their workspace fits in the
 * arena; admitted requests are launched as a batch on a free stream as soon
 * as one is free, so batches overlap on the device and new arrivals join
 * the next batch while earlier ones run.  Row spectra of tiles are kept
 * resident in the arena keyed by (tile, Nx) and reused by later requests
 * (least recently used first out when memory is needed).  A request's
 * result does not depend on what it was batched with.  Tiles must outlive
 * the requests that name them; a released tile's cached spectra are never
 * reused (the cache keys on a per-tile id, not its address).  Tabulated
 * LegalRgn rules and rotated pairs are XC_E_UNSUPPORTED here.
 * This has no single counterpart in the reference, whose pair jobs are
 * independent processes (make -j on grid-engine nodes); it replaces that
 * job-level parallelism on one GPU.
 */
typedef struct xc_server xc_server;

typedef struct xc_serve_cfg {
    int64_t hbm_budget;     /* bytes of device memory for workspaces + resident spectra */
    int32_t max_batch;      /* requests per launched batch */
    int32_t streams;        /* batches in flight at once */
    int32_t max_wait_us;    /* a partial batch waits at most this long for company (0: launch at once) */
    int32_t deadline_us;    /* a request not launched within this long of its submission completes with
                               XC_E_TIMEOUT (0: no deadline); once launched it runs to its result */
} xc_serve_cfg;

typedef struct xc_serve_stats {
    int64_t submitted, completed, rejected, batches;
    int64_t spectra_hits, spectra_misses;
    int64_t arena_bytes, arena_peak;
    int64_t timed_out, cancelled;
    int64_t failed;         /* completed with an error of the device path (status < 0) */
} xc_serve_stats;

int xc_server_create(xc_ctx *ctx, const xc_serve_cfg *cfg, xc_server **out);
/* Submit one request; *ticket identifies it.  XC_E_BUDGET if the request
 * alone can never fit the budget.  Thread-safe. */
int xc_server_submit(xc_server *s, const xc_pair *pair, uint64_t *ticket);
/* XC_OK with *res filled, or XC_E_NOTREADY. */
int xc_server_poll(xc_server *s, uint64_t ticket, xc_result *res);
/* Cancel a request that is still queued: it completes with XC_E_CANCELLED
 * (XC_OK returned).  XC_E_NOTREADY when it is already on the device or done
 * (its result stands); XC_E_ARG for an unknown ticket. */
int xc_server_cancel(xc_server *s, uint64_t ticket);
/* Block until the request completes; *latency_us (may be NULL) receives
 * submit-to-completion time as measured by the server. */
int xc_server_wait(xc_server *s, uint64_t ticket, xc_result *res, double *latency_us);
int xc_server_stats(xc_server *s, xc_serve_stats *out);
/* Drain outstanding requests and stop. */
int xc_server_destroy(xc_server *s);

#ifdef __cplusplus
}
#endif

#endif /* XCORR_H */
