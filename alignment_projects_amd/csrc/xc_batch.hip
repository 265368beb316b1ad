// xc_batch.hip -- device-resident tiles and batched FP32 plans.
//
// A tile is one image kept in HBM with its summed-area tables (exact integer
// sums for unmasked 8-bit tiles of at most 2^20 pixels, else FP64 entries, or
// FP32 ones when the tile asks for them) and its (mean, sd); it plays the role of the reference's fft2 / ThmRec.ftc cache
// (Correlation_fft_fftw.cpp:50-51, CThmScan.h:27-34), made explicit.  A plan
// is a fixed batch of pair correlations laid out once (descriptors, work
// lists, spectra, lag rows) and replayed as one HIP graph.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <vector>

#include "xc_exec.hpp"
#include "xc_kernels.hpp"

using namespace xc;


struct xc_plan {
    xc_ctx *ctx = nullptr;
    int n = 0;
    DevBuf ws;
    Exec ex;
    xc_result *d_res = nullptr;
    xc_result *h_res = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double stage_bytes[ST_COUNT] = {0, 0, 0, 0, 0};
    double model_bytes = 0.0;  // whole-pipeline algorithmic bytes
    double sweep_us = 0.0;     // tuned whole-plan time with the sweeps' share in use (0: not tuned)
    std::vector<int> order;
};

static int elem_size(int dtype)
{
    switch (dtype) {
    case XC_U8: return 1;
    case XC_U16: return 2;
    case XC_F32: return 4;
    case XC_F64: return 8;
    }
    return 0;
}

extern "C" int xc_tile_put(xc_ctx *ctx, const void *data, int dtype, int w, int h, int stride, const uint8_t *mask,
                           int flags, xc_tile **out)
{
    if (!ctx || !out || !data) return XC_E_ARG;
    *out = nullptr;
    const int es = elem_size(dtype);
    if (!es || w <= 0 || h <= 0) return set_err(ctx, XC_E_ARG, "bad tile dtype or extent");
    if (stride <= 0) stride = w;
    if (stride < w) return set_err(ctx, XC_E_ARG, "stride < width");
    if (w > TP_MAX_W) return set_err(ctx, XC_E_SHAPE, "tile wider than %d", TP_MAX_W);
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const bool dev_src = (flags & XC_TILE_DEVICE_SRC) != 0;
    // per-row present extents, and the bounding box of the present pixels: the raster the reference builds
    // from the point list spans that box (BBoxFromPoints, Geometry.cpp:78-106), so the tile's view does too
    std::vector<int> rx0(h, 0), rx1(h, w - 1);
    int L = 0, R = w - 1, Bt = 0, T = h - 1;
    if (mask) {
        std::vector<uint8_t> hm;
        const uint8_t *m = mask;
        if (dev_src) {  // a device mask's rows are `stride` bytes apart (the picture's stride)
            hm.resize((size_t)w * h);
            if (hipMemcpy2D(hm.data(), w, mask, stride, w, h, hipMemcpyDeviceToHost) != hipSuccess)
                return set_err(ctx, XC_E_DEVICE, "tile mask download failed");
            m = hm.data();
        }
        L = w;
        R = -1;
        Bt = -1;
        for (int y = 0; y < h; ++y) {
            int a = -1, b = -1;
            for (int x = 0; x < w; ++x)
                if (m[(size_t)y * w + x]) {
                    if (a < 0) a = x;
                    b = x;
                }
            rx0[y] = a;
            rx1[y] = b;
            if (a >= 0) {
                L = std::min(L, a);
                R = std::max(R, b);
                if (Bt < 0) Bt = y;
                T = y;
            }
        }
        if (R < 0) return set_err(ctx, XC_E_ARG, "tile has no present pixel");
    }
    const int vw = R - L + 1, vh = T - Bt + 1;
    // table form: FP32 on request; exact integer sums for unmasked u8 tiles of at most 2^20 pixels (their box
    // sums and variance terms are then exact); FP64 pairs otherwise or on request
    int tabk = TK_F64;
    if (flags & XC_TILE_F32_TABLES) tabk = TK_F32;
    else if (dtype == XC_U8 && !mask && (long long)vw * vh <= TABX_MAX_PX && !(flags & XC_TILE_F64_TABLES)) tabk = TK_X;
    static std::atomic<uint64_t> next_uid{1};
    xc_tile *t = new xc_tile;
    t->uid = next_uid++;
    t->ctx = ctx;
    t->w = vw;
    t->h = vh;
    t->ox = L;
    t->oy = Bt;
    t->fw = w;
    t->fh = h;
    t->dtype = dtype;
    t->flags = flags;
    t->rx0.swap(rx0);
    t->rx1.swap(rx1);
    const TileSizes ts = tile_sizes(vw, vh, tabk);
    // an uploaded 8-bit picture (and any mask) is stored with rows padded to 16-byte multiples, so that the
    // table sweep can stage whole 16-byte pieces of every row
    const int pitch = (!dev_src && es == 1) ? (w + 15) / 16 * 16 : stride;
    const int mpitch = dev_src ? stride : mask ? (w + 15) / 16 * 16 : w;
    Layout lay;
    size_t iS = lay.add(ts.SQ), iZ = lay.add(ts.Z), iP = lay.add(ts.part);
    size_t iM = lay.add(ts.msd), iD = lay.add(sizeof(TileDev)), iL = lay.add(sizeof(int));
    size_t iRaw = dev_src ? 0 : lay.add((size_t)es * pitch * h);
    size_t iMask = (!dev_src && mask) ? lay.add((size_t)mpitch * h) : 0;
    int st = ensure_dev(ctx, t->mem, lay.total, false);  // a tile's memory never grows
    if (st) {
        delete t;
        return st;
    }
    char *base = (char *)t->mem.p;
    if (dev_src) {
        t->raw = const_cast<void *>(data);
        t->mask = const_cast<uint8_t *>(mask);
    } else {
        t->raw = base + lay.off[iRaw];
        t->own_raw = true;
        if (hipMemcpy2D(t->raw, (size_t)es * pitch, data, (size_t)es * stride, (size_t)es * w, h,
                        hipMemcpyHostToDevice) != hipSuccess) {
            xc_tile_release(ctx, t);
            return set_err(ctx, XC_E_DEVICE, "tile upload failed");
        }
        if (mask) {
            t->mask = (uint8_t *)(base + lay.off[iMask]);
            t->own_mask = true;
            if (hipMemcpy2D(t->mask, mpitch, mask, w, w, h, hipMemcpyHostToDevice) != hipSuccess) {
                xc_tile_release(ctx, t);
                return set_err(ctx, XC_E_DEVICE, "tile mask upload failed");
            }
        }
    }
    // the whole picture (rotated copies map the points' own coordinates) ...
    ImgView &F = t->full;
    memset(&F, 0, sizeof F);
    F.data = t->raw;
    F.mask = t->mask;
    F.dtype = dtype;
    F.w = w;
    F.h = h;
    F.stride = dev_src ? stride : pitch;
    F.mstride = mpitch;
    F.msd = (double *)(base + lay.off[iM]);
    // ... and the bounding-box view the correlations and the tables use
    TileDev &D = t->dev;
    memset(&D, 0, sizeof D);
    D.v = F;
    D.v.data = (const char *)t->raw + (size_t)es * ((size_t)Bt * F.stride + L);
    D.v.mask = t->mask ? t->mask + (size_t)Bt * F.mstride + L : nullptr;
    D.v.w = vw;
    D.v.h = vh;
    D.msd = (double *)(base + lay.off[iM]);
    D.SQ = base + lay.off[iS];
    D.Z = (int32_t *)(base + lay.off[iZ]);
    D.part = (double *)(base + lay.off[iP]);
    D.nbands = 1;  // swept unbanded, as a plan's batched refresh sweeps it
    D.normalize = (flags & XC_TILE_NORMALIZE) ? 1 : 0;
    D.tabk = tabk;
    t->d_dev = (TileDev *)(base + lay.off[iD]);
    const int zero = 0;
    if (hipMemcpy(t->d_dev, &D, sizeof D, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(base + lay.off[iL], &zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess) {
        xc_tile_release(ctx, t);
        return set_err(ctx, XC_E_DEVICE, "tile descriptor upload failed");
    }
    t->bytes = (int64_t)t->mem.bytes;
    // prepare: statistics + summed-area tables of the FP32 values
    Exec ex;
    ex.ntiles = 1;
    ex.d_tiles = t->d_dev;
    SatGroup sg;
    sg.cpt = sat_cpt(vw);
    sg.threads = sat_threads(vw);
    sg.count = 1;
    sg.tabk = tabk;
    sg.stage = dtype == XC_U8;
    sg.list = (const int *)(base + lay.off[iL]);
    ex.sat_groups.push_back(sg);
    ex.fp64 = false;
    st = enqueue(ctx, ex, ST_PREP, ST_PREP + 1);
    if (st == XC_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) st = set_err(ctx, XC_E_DEVICE, "tile prep");
    if (st) {
        xc_tile_release(ctx, t);
        return st;
    }
    *out = t;
    return XC_OK;
}

extern "C" int xc_tile_release(xc_ctx *ctx, xc_tile *t)
{
    if (!t) return XC_E_ARG;
    if (!ctx) ctx = t->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (t->mem.p) (void)hipFree(t->mem.p);
    delete t;
    return XC_OK;
}

extern "C" int64_t xc_tile_bytes(const xc_tile *t) { return t ? t->bytes : 0; }

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------

extern "C" int xc_plan_destroy(xc_plan *p);

// Host side of the rotated bounding box (BBoxFromPoints over the rotated
// present pixels, Geometry.cpp:78-106).  x' and y' are monotone in x along a
// row, so each row's first and last present pixels carry its extremes.
#pragma clang fp contract(off)
static void rotated_bbox(const xc_tile *t, double t0, double t1, double t3, double t4, int &L, int &R, int &B, int &T)
{
    double xlo = 1e300, xhi = -1e300, ylo = 1e300, yhi = -1e300;
    for (int y = 0; y < t->fh; ++y) {
        const int a = t->rx0[y], b = t->rx1[y];
        if (a < 0) continue;
        for (int x : {a, b}) {
            const double xr = (double)x * t0 + (double)y * t1, yr = (double)x * t3 + (double)y * t4;
            xlo = std::min(xlo, xr);
            xhi = std::max(xhi, xr);
            ylo = std::min(ylo, yr);
            yhi = std::max(yhi, yr);
        }
    }
    L = (int)floor(xlo);
    R = (int)ceil(xhi);
    B = (int)floor(ylo);
    T = (int)ceil(yhi);
}
#pragma clang fp contract(on)

// GenDefs.h:16: the reference's deg -> rad constant.
static const double REF_PI = 3.14159265358;

struct RotKey {  // a tile and the 2 x 2 map of its copy
    const xc_tile *t;
    double m[4];
    bool operator<(const RotKey &o) const
    {
        if (t != o.t) return t < o.t;
        for (int k = 0; k < 4; ++k)
            if (m[k] != o.m[k]) return m[k] < o.m[k];
        return false;
    }
};

// Autotuner of the plan's shapes -- the shapes that change speed only, never
// results: for each size bucket it times every candidate the device accepts on
// the plan's own data and keeps the fastest.  Column pass: columns per
// workgroup (1..16, i.e. 64..1024 threads) x XCD-aware workgroup order; row
// r2c: rows per thread (two rows per thread sharing barriers and twiddles, or
// one row per thread group with twice the groups per workgroup).  The radix
// schedule is not a candidate: another order of passes changes the rounding.
// A context remembers its choices (and XC_TUNE_FILE keeps them across
// processes).
static void tune_save(xc_ctx *ctx, const char *tag, int a, int b, int c, const xc_col_tune &t)
{
    if (ctx->tune_file.empty()) return;
    FILE *f = fopen(ctx->tune_file.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %d %d %d %d %d %.3f\n", tag, a, b, c, t.c, t.xcd, t.us);
    fclose(f);
}

// Smallest of 3 timed launches of `launch` on the context stream, in us (< 0: the device refused it).
template <typename F> static double time_launch(xc_ctx *ctx, hipEvent_t e0, hipEvent_t e1, F &&launch)
{
    float tmin = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        const bool ok = hipEventRecord(e0, ctx->stream) == hipSuccess && launch() == XC_OK &&
                        hipEventRecord(e1, ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float t = 0.f;
        if (!ok) {
            (void)hipGetLastError();
            return -1.0;
        }
        if (hipEventElapsedTime(&t, e0, e1) == hipSuccess) tmin = std::min(tmin, t);
    }
    return 1e3 * tmin;
}

static int tune_shapes(xc_ctx *ctx, xc_plan *p)
{
    int st = enqueue(ctx, p->ex, ST_PREP, ST_COL);  // real spectra for the candidates to transform
    if (st) return st;
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipEvent_t e0, e1;
    XC_HIP(ctx, hipEventCreate(&e0));
    XC_HIP(ctx, hipEventCreate(&e1));
    for (Bucket &b : p->ex.buckets) {
        // row r2c: rows per thread, and for the 8-bit kernel work items per workgroup
        // (the bucket's own rows: one tuning per (NH, NY, kernel), like the column pass's)
        const int kern = b.d_r2c_u8 ? (b.r2c_gen ? 2 : 1) : 0;  // general, 8-bit, 8-bit general rows
        const auto rkey = std::make_tuple(b.NH, b.NY, kern);
        auto ir = ctx->tune_r2c.find(rkey);
        if (ir == ctx->tune_r2c.end()) {
            xc_col_tune best;
            best.us = 1e300;
            best.xcd = 0;
            for (int rpt : {r2c_default_rpt<float>(b.NH), r2c_alt_rpt<float>(b.NH)}) {
                if (rpt == 0) continue;
                for (int np : {1, 2, 4}) {
                    if (!b.d_r2c_u8 && np > 1) break;
                    const double us = time_launch(ctx, e0, e1, [&] {
                        return b.d_r2c_u8 ? launch_row_r2c_u8(ctx, b.NH, b.n_r2c, b.d_r2c_u8, rpt, np, b.r2c_gen)
                                          : launch_row_r2c<float>(ctx, b.NH, b.n_r2c, p->ex.d_pairs, b.d_r2c, rpt);
                    });
                    if (us >= 0.0 && us < best.us) {
                        best.c = rpt;
                        best.xcd = b.d_r2c_u8 ? np : 0;
                        best.us = us;
                    }
                }
            }
            if (best.c == 0) best.us = 0.0;
            ir = ctx->tune_r2c.emplace(rkey, best).first;
            tune_save(ctx, "r2c", b.NH, b.NY, kern, best);
        }
        b.r2c_rpt = ir->second.c;
        b.r2c_np = b.d_r2c_u8 ? ir->second.xcd : 0;
        b.r2c_us = ir->second.us;
        // column pass: columns per workgroup x XCD order
        const auto key = std::make_tuple(b.NH, b.NY, b.whiten ? 1 : 0);
        auto it = ctx->tune.find(key);
        if (it == ctx->tune.end()) {
            xc_col_tune best;
            best.us = 1e300;
            for (int c : {1, 2, 4, 8, 16}) {
                if (!col_shape_ok(ctx, b.NY, c)) continue;
                for (int x : {1, 0}) {
                    const double us = time_launch(ctx, e0, e1, [&] {
                        return launch_col<float>(ctx, b.NY, b.whiten, c, x, b.n_col, p->ex.d_pairs, b.d_col);
                    });
                    if (us >= 0.0 && us < best.us) {
                        best.c = c;
                        best.xcd = x;
                        best.us = us;
                    }
                }
            }
            if (best.c == 0) best.us = 0.0;  // nothing ran: keep the default shape
            it = ctx->tune.emplace(key, best).first;
            tune_save(ctx, "col", std::get<0>(key), std::get<1>(key), std::get<2>(key), best);
        }
        b.col_c = it->second.c;
        b.col_xcd = it->second.xcd;
        b.col_us = it->second.us;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->err.clear();
    return XC_OK;
}

// (Re)record the plan's launches as its graph.
static int plan_capture(xc_ctx *ctx, xc_plan *p)
{
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    XC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue(ctx, p->ex);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ce != hipSuccess) return set_err(ctx, XC_E_DEVICE, "graph capture: %s", hipGetErrorString(ce));
    p->graph = g;
    XC_HIP(ctx, hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    return XC_OK;
}

// The table sweeps' share of each CU while they run beside the row
// transforms (wavefronts per CU, 0: one workgroup per item): each candidate
// is recorded as the plan's graph and timed over replays of the whole plan,
// and the fastest stays (the caller records the graph again).  A plan whose
// sweeps do not run beside the transforms has nothing to tune.
static int tune_sweep_share(xc_ctx *ctx, xc_plan *p)
{
    if (!enqueue_forks(ctx, p->ex)) return XC_OK;
    hipEvent_t e0, e1;
    XC_HIP(ctx, hipEventCreate(&e0));
    XC_HIP(ctx, hipEventCreate(&e1));
    double best = 1e300;
    int bw = p->ex.sat_waves, st = XC_OK;
    for (int w : {0, 16, 8}) {
        p->ex.sat_waves = w;
        if ((st = plan_capture(ctx, p))) break;
        auto replay = [&] { return hipGraphLaunch(p->exec, ctx->stream) == hipSuccess ? XC_OK : XC_E_DEVICE; };
        if ((st = replay())) break;  // warm-up
        const double us = time_launch(ctx, e0, e1, replay);
        if (us >= 0.0 && us < best) {
            best = us;
            bw = w;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    p->ex.sat_waves = bw;
    p->sweep_us = best < 1e300 ? best : 0.0;
    ctx->err.clear();
    return XC_OK;
}

// aff4 (may be null): per pair the 2 x 2 map {t0, t1, t3, t4} applied to A's points in place
// of the pair's angle (an identity map means no copy, as 0 degrees does).
static int plan_build(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, xc_plan *p, const double *aff4)
{
    const bool refresh = (flags & XC_PLAN_REFRESH_TILES) != 0;
    // ---- pair geometry, rotated copies ----
    std::vector<PairDev> P(n);
    std::vector<const xc_tile *> tiles;
    std::map<const xc_tile *, int> tile_idx;
    std::vector<RotDev> rots;
    std::vector<int> rot_src;                      // per copy: its source (index into nsrcs)
    std::map<const xc_tile *, int> nsrc_idx;
    std::vector<const xc_tile *> nsrcs;            // the copies' distinct sources
    std::map<RotKey, int> rot_idx;
    std::vector<int> pair_rot(n, -1);
    std::vector<bool> rot_cnt;  // the rotated copy needs a count table
    std::map<const xc_tile *, int> der_idx;  // exact tiles that some pair reads through a derived FP64 table
    std::vector<const xc_tile *> ders;
    std::vector<int> derA(n, -1), derB(n, -1);
    for (int i = 0; i < n; ++i) {
        const xc_pair &q = pairs[i];
        const xc_tile *a = q.a, *b = q.b;
        if (!a || !b) return set_err(ctx, XC_E_ARG, "pair %d: null tile", i);
        if (q.mode != XC_MODE_NCC && q.mode != XC_MODE_R && q.mode != XC_MODE_S)
            return set_err(ctx, XC_E_ARG, "pair %d: bad mode", i);
        if (q.rgn.enabled && q.rgn.table)
            return set_err(ctx, XC_E_UNSUPPORTED, "pair %d: tabulated LegalRgn in batched form", i);
        for (const xc_tile *t : {a, b})
            if (!tile_idx.count(t)) {
                tile_idx[t] = (int)tiles.size();
                tiles.push_back(t);
            }
        // B2.L - B1.L, B2.B - B1.B: the corners of the two views' bounding boxes
        int wa = a->w, ha = a->h, offx = b->ox - a->ox, offy = b->oy - a->oy;
        RotKey key{a, {1.0, 0.0, 0.0, 1.0}};
        bool inv_t = true;  // a rotation: its inverse is its transpose
        if (aff4) {
            for (int k = 0; k < 4; ++k) key.m[k] = aff4[4 * i + k];
            inv_t = false;
        } else if (q.angle_deg != 0.0) {
            // RFromAngle: R.NUSetRot(deg * PI / 180) (TAffine.cpp:116-122); Tdfm = Tptwk = identity
            const double r = q.angle_deg * REF_PI / 180.0, c = cos(r), s = sin(r);
            key.m[0] = c;
            key.m[1] = -s;
            key.m[2] = s;
            key.m[3] = c;
        }
        if (!(key.m[0] == 1.0 && key.m[1] == 0.0 && key.m[2] == 0.0 && key.m[3] == 1.0)) {
            auto it = rot_idx.find(key);
            int ri;
            if (it == rot_idx.end()) {
                RotDev R;
                memset(&R, 0, sizeof R);
                R.src = a->full;  // the rotation maps A's points in their own coordinates
                R.t0 = key.m[0];
                R.t1 = key.m[1];
                R.t3 = key.m[2];
                R.t4 = key.m[3];
                if (inv_t) {
                    R.u0 = R.t0;
                    R.u1 = R.t3;
                    R.u3 = R.t1;
                    R.u4 = R.t4;
                } else {
                    const double det = R.t0 * R.t4 - R.t1 * R.t3;
                    if (!(fabs(det) > 1e-12)) return set_err(ctx, XC_E_ARG, "pair %d: singular map", i);
                    R.u0 = R.t4 / det;
                    R.u1 = -R.t1 / det;
                    R.u3 = -R.t3 / det;
                    R.u4 = R.t0 / det;
                }
                int L, Rr, B, T;
                rotated_bbox(a, R.t0, R.t1, R.t3, R.t4, L, Rr, B, T);
                if (L > Rr) return set_err(ctx, XC_E_ARG, "pair %d: tile A has no present pixel", i);
                R.L = L;
                R.B = B;
                R.w = Rr - L + 1;
                R.h = T - B + 1;
                if (R.w > TP_MAX_W) return set_err(ctx, XC_E_SHAPE, "pair %d: rotated A wider than %d", i, TP_MAX_W);
                ri = (int)rots.size();
                rots.push_back(R);
                if (!nsrc_idx.count(a)) {
                    nsrc_idx[a] = (int)nsrcs.size();
                    nsrcs.push_back(a);
                }
                rot_src.push_back(nsrc_idx[a]);
                rot_cnt.push_back(false);
                rot_idx[key] = ri;
            } else {
                ri = it->second;
            }
            pair_rot[i] = ri;
            if (q.cnt.enabled) rot_cnt[ri] = true;
            wa = rots[ri].w;
            ha = rots[ri].h;
            offx = b->ox - rots[ri].L;  // B2.L - B1.L
            offy = b->oy - rots[ri].B;
        }
        PairDev &D = P[i];
        memset(&D, 0, sizeof D);
        const int mode = q.mode == XC_MODE_NCC ? PM_NCC : (q.mode == XC_MODE_R ? PM_R : PM_S);
        setup_pair_geometry(D, wa, ha, b->w, b->h, mode);
        if (mode != PM_NCC) {
            const int md = 11 / 2 + 1;
            if (D.nnegx < md || D.nposx < md || D.nnegy < md || D.nposy < md)
                return set_err(ctx, XC_E_SHAPE, "pair %d: too small for CorrImages (SetDims)", i);
        }
        if (D.Nx / 2 > max_nh<float>() || D.Ny > max_ny<float>())
            return set_err(ctx, XC_E_SHAPE, "pair %d: padded size %d x %d too large", i, D.Nx, D.Ny);
        D.Ox = q.Ox;
        D.Oy = q.Oy;
        D.radius = q.radius;
        D.Rx = q.Rx;
        D.Ry = q.Ry;
        D.offx = offx;
        D.offy = offy;
        D.rule = make_rule(&q.rgn, &q.cnt, nullptr);
        D.need_cnt = D.rule.cnt_enabled;
        // table forms: both sides read one form; an unrotated exact (TK_X) side beside an FP64 one -- or
        // beside a rotated copy, whose tables are FP64 -- reads an FP64 table the plan derives from its tile
        // (S mode reads no (S, S^2) table)
        const int ka = pair_rot[i] >= 0 ? (a->dev.tabk == TK_F32 ? TK_F32 : TK_F64) : a->dev.tabk, kb = b->dev.tabk;
        if (mode == PM_S || ka == kb) {
            D.tabk = ka == kb ? ka : TK_F64;
        } else if (ka != TK_F32 && kb != TK_F32) {
            D.tabk = TK_F64;
            for (int side = 0; side < 2; ++side) {
                const xc_tile *t = side ? b : a;
                if ((side ? kb : ka) != TK_X) continue;
                if (!der_idx.count(t)) {
                    der_idx[t] = (int)ders.size();
                    ders.push_back(t);
                }
                (side ? derB : derA)[i] = der_idx[t];
            }
        } else {
            return set_err(ctx, XC_E_ARG, "pair %d: FP32 tables pair only with FP32 tables", i);
        }
        narrow_rows(D);
        D.A = a->dev.v;
        D.B = b->dev.v;
        D.SQA = a->dev.SQ;
        D.ZA = a->dev.Z;
        D.SQB = b->dev.SQ;
        D.ZB = b->dev.Z;
    }
    std::vector<Bucket> buckets;
    make_buckets(P, 0, p->order, buckets);

    // ---- shared row spectra: one per (image, Nx); image = tile or rotated copy ----
    struct SpecKey {
        const void *img;
        int nx;
        bool operator<(const SpecKey &o) const { return img < o.img || (img == o.img && nx < o.nx); }
    };
    std::map<SpecKey, int> spec_idx;
    std::vector<size_t> spec_bytes;
    std::vector<int> specA(n), specB(n);
    std::vector<bool> ownA(n, false), ownB(n, false);  // this pair's r2c produces the spectrum
    const size_t cs = sizeof(cpx<float>);
    for (int k = 0; k < n; ++k) {
        const int i = p->order[k];
        const PairDev &D = P[i];
        const size_t H = D.Nx / 2 + 1;
        const void *ia = pair_rot[i] >= 0 ? (const void *)&rots[pair_rot[i]] : (const void *)pairs[i].a;
        const void *ib = (const void *)pairs[i].b;
        for (int side = 0; side < 2; ++side) {
            const SpecKey key{side ? ib : ia, D.Nx};
            auto it = spec_idx.find(key);
            int si;
            if (it == spec_idx.end()) {
                si = (int)spec_bytes.size();
                spec_idx[key] = si;
                spec_bytes.push_back(cs * H * (side ? D.h2 : D.h1));
                (side ? ownB : ownA)[i] = true;
            } else {
                si = it->second;
            }
            (side ? specB : specA)[i] = si;
        }
    }

    // ---- layout ----
    Layout lay;
    std::vector<size_t> i_spec(spec_bytes.size());
    for (size_t s = 0; s < spec_bytes.size(); ++s) i_spec[s] = lay.add(spec_bytes[s]);
    std::vector<size_t> i_lag(n), i_rows(n);
    for (int k = 0; k < n; ++k) {
        const PairDev &D = P[p->order[k]];
        i_lag[k] = lay.add(cs * (size_t)(D.Nx / 2 + 1) * D.hR);
        i_rows[k] = lay.add(sizeof(RowRec) * D.hR);
    }
    // rotated copies: raster + tables
    std::vector<size_t> i_rimg(rots.size()), i_rS(rots.size()), i_rZ(rots.size());
    std::vector<bool> rot_sq(rots.size(), false);
    // a rotated copy's tables: FP32 pairs for a source with FP32 tables, else FP64 pairs (its values are
    // not integers); then the derived FP64 tables, swept in the same launches
    std::vector<char> dkind(rots.size() + ders.size(), (char)TK_F64);
    for (int i = 0; i < n; ++i)
        if (pair_rot[i] >= 0) {
            if (pairs[i].mode != XC_MODE_S) rot_sq[pair_rot[i]] = true;
            dkind[pair_rot[i]] = (char)(pairs[i].a->dev.tabk == TK_F32 ? TK_F32 : TK_F64);
        }
    for (size_t r = 0; r < rots.size(); ++r) {
        const size_t px = (size_t)rots[r].w * rots[r].h;
        i_rimg[r] = lay.add(sizeof(float) * px);
        i_rS[r] = lay.add(rot_sq[r] ? (dkind[r] == TK_F32 ? sizeof(float2) : sizeof(double2)) * px : 16);
        i_rZ[r] = lay.add(rot_cnt[r] ? sizeof(int32_t) * px : 8);
    }
    std::vector<size_t> i_nsrc(nsrcs.size());
    for (size_t k = 0; k < nsrcs.size(); ++k)
        i_nsrc[k] = lay.add(sizeof(double) * (size_t)nsrcs[k]->full.w * nsrcs[k]->full.h);
    std::vector<size_t> i_derS(ders.size());
    for (size_t d = 0; d < ders.size(); ++d) i_derS[d] = lay.add(sizeof(double2) * (size_t)ders[d]->w * ders[d]->h);
    size_t i_res = lay.add(sizeof(xc_result) * n);
    // uploaded block: constants, pairs, tiles, rotations, work lists
    size_t i_msd = lay.add(sizeof(double) * 3);
    size_t i_pairs = lay.add(sizeof(PairDev) * n);
    size_t i_tiles = lay.add(sizeof(TileDev) * (refresh ? tiles.size() : 0) + 16);
    std::vector<SatGroup> sgroups, dgroups;
    std::vector<int> sflat, soff, dflat, doff;
    if (refresh) {
        std::vector<int> widths;
        std::vector<char> kinds, u8;
        for (const xc_tile *t : tiles) {
            widths.push_back(t->w);
            kinds.push_back((char)t->dev.tabk);
            u8.push_back((char)(t->dtype == XC_U8));
        }
        make_sat_groups(widths, kinds, u8, sgroups, sflat, soff);
    }
    {
        std::vector<int> widths;
        std::vector<char> u8(rots.size(), 0);  // rotated copies are FP32 rasters; derived tables read u8 tiles
        for (const RotDev &R : rots) widths.push_back(R.w);
        for (const xc_tile *t : ders) {
            widths.push_back(t->w);
            u8.push_back((char)(t->dtype == XC_U8));
        }
        make_sat_groups(widths, dkind, u8, dgroups, dflat, doff);
    }
    size_t i_slist = lay.add(sizeof(int) * sflat.size() + 16);
    size_t i_dlist = lay.add(sizeof(int) * dflat.size() + 16);
    size_t i_rot = lay.add(sizeof(RotDev) * rots.size() + 16);
    size_t i_dtiles = lay.add(sizeof(TileDev) * (rots.size() + ders.size()) + 16);
    int nrw = 0;
    for (const RotDev &R : rots) nrw += (int)((rot_splat_units(R.w, R.h) + 255) / 256);
    size_t i_rw = lay.add(sizeof(int2) * nrw + 16);
    size_t i_rn = lay.add(sizeof(RotNorm) * nsrcs.size() + 16);
    int nnw = 0;
    for (const xc_tile *t : nsrcs) nnw += (int)(((long long)t->full.w * t->full.h + 255) / 256);
    size_t i_rnw = lay.add(sizeof(int2) * nnw + 16);
    std::vector<size_t> i_wr(buckets.size()), i_wc(buckets.size()), i_wc2(buckets.size()), i_w8(buckets.size(), 0);
    std::vector<bool> r2c8(buckets.size(), false), r2cg(buckets.size(), false);
    std::vector<int> starts;
    // the 8-bit row r2c takes u8 images that fit the upper-half padding; rows it cannot load as pixel
    // pairs (masked, or starting at odd addresses) need its general form
    auto u8_rows = [](const ImgView &V, int NH) { return V.dtype == XC_U8 && V.w <= NH; };
    auto u8_pairs = [](const ImgView &V) { return !V.mask && ((uintptr_t)V.data & 1) == 0 && (V.stride & 1) == 0; };
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        Bucket &B = buckets[bi];
        const int G = row_group<float>(B.NH), C = col_group<float>(B.NY), GR = r2c_rows<float>(B.NH);
        int nr = 0, nc = 0, nc2 = 0;
        bool u8 = r2c_u8_length(B.NH);
        for (int k = B.first; k < B.first + B.count; ++k) {
            const int i = p->order[k];
            const PairDev &D = P[i];
            if (ownA[i]) {
                nr += (D.h1 + GR - 1) / GR;
                u8 = u8 && pair_rot[i] < 0 && u8_rows(D.A, B.NH);
                if (!u8_pairs(D.A)) r2cg[bi] = true;
            }
            if (ownB[i]) {
                nr += (D.h2 + GR - 1) / GR;
                u8 = u8 && u8_rows(D.B, B.NH);
                if (!u8_pairs(D.B)) r2cg[bi] = true;
            }
            nc += (D.Nx / 2 + 1 + C - 1) / C;
            live_row_groups(D, G, starts);
            nc2 += (int)starts.size();
        }
        B.n_r2c = nr;
        B.n_col = nc;
        B.n_c2r = nc2;
        r2c8[bi] = u8 && nr > 0;
        if (r2c8[bi]) i_w8[bi] = lay.add(sizeof(R2cU8) * nr + 16);
        i_wr[bi] = lay.add(sizeof(int4) * nr + 16);
        i_wc[bi] = lay.add(sizeof(int2) * nc);
        i_wc2[bi] = lay.add(sizeof(int2) * nc2);
    }
    const size_t up_lo = lay.off[i_msd];
    const size_t up_hi = lay.total;
    int st = ensure_dev(ctx, p->ws, lay.total, false);  // a plan is laid out once
    if (st) return st;
    char *base = (char *)p->ws.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };

    // ---- host copy of the uploaded block ----
    std::vector<char> up(up_hi - up_lo);
    auto hat = [&](size_t i) { return up.data() + (lay.off[i] - up_lo); };
    double *hmsd = (double *)hat(i_msd);
    hmsd[0] = 0.0;
    hmsd[1] = 1.0;
    hmsd[2] = 1.0;
    const double *dmsd = (const double *)at(i_msd);
    // rotated copies and their table descriptors
    RotDev *hr = (RotDev *)hat(i_rot);
    TileDev *hd = (TileDev *)hat(i_dtiles);
    int2 *hrw = (int2 *)hat(i_rw);
    int w = 0;
    {
        RotNorm *hn = (RotNorm *)hat(i_rn);
        int2 *hnw = (int2 *)hat(i_rnw);
        int wn = 0;
        for (size_t k = 0; k < nsrcs.size(); ++k) {
            hn[k].src = nsrcs[k]->full;
            hn[k].dst = (double *)at(i_nsrc[k]);
            const long long np = ((long long)nsrcs[k]->full.w * nsrcs[k]->full.h + 255) / 256;
            for (long long b = 0; b < np; ++b) hnw[wn++] = make_int2((int)k, (int)b);
        }
    }
    for (size_t r = 0; r < rots.size(); ++r) {
        RotDev R = rots[r];
        R.out = at(i_rimg[r]);
        R.nsrc = (const double *)at(i_nsrc[rot_src[r]]);
        hr[r] = R;
        for (int b = 0; b < (int)((rot_splat_units(R.w, R.h) + 255) / 256); ++b) hrw[w++] = make_int2((int)r, b);
        TileDev T;
        memset(&T, 0, sizeof T);
        T.v.data = R.out;
        T.v.dtype = XC_F32;
        T.v.w = R.w;
        T.v.h = R.h;
        T.v.stride = R.w;
        T.v.mstride = R.w;
        T.v.msd = dmsd;
        T.msd = nullptr;
        T.SQ = rot_sq[r] ? (void *)at(i_rS[r]) : nullptr;
        T.Z = rot_cnt[r] ? (int32_t *)at(i_rZ[r]) : nullptr;
        T.nbands = 1;
        T.tabk = dkind[r];
        hd[r] = T;
        rots[r] = R;
    }
    for (size_t d = 0; d < ders.size(); ++d) {
        // the tile's own view and statistics, an FP64 table of the plan's (its count table is the tile's own)
        TileDev T = ders[d]->dev;
        T.msd = nullptr;
        T.SQ = at(i_derS[d]);
        T.Z = nullptr;
        T.nbands = 1;
        T.tabk = TK_F64;
        hd[rots.size() + d] = T;
    }
    if (!dflat.empty()) memcpy(hat(i_dlist), dflat.data(), sizeof(int) * dflat.size());
    for (size_t gi = 0; gi < dgroups.size(); ++gi) dgroups[gi].list = (const int *)at(i_dlist) + doff[gi];

    PairDev *hp = (PairDev *)hat(i_pairs);
    p->d_res = (xc_result *)at(i_res);
    for (int k = 0; k < n; ++k) {
        const int i = p->order[k];
        PairDev D = P[i];
        if (pair_rot[i] >= 0) {
            const TileDev &T = hd[pair_rot[i]];
            D.A = T.v;
            D.SQA = T.SQ;
            D.ZA = T.Z;
        }
        if (derA[i] >= 0) D.SQA = hd[rots.size() + derA[i]].SQ;
        if (derB[i] >= 0) D.SQB = hd[rots.size() + derB[i]].SQ;
        D.specA = at(i_spec[specA[i]]);
        D.specB = at(i_spec[specB[i]]);
        D.lag = at(i_lag[k]);
        D.rows = (RowRec *)at(i_rows[k]);
        D.out = p->d_res + i;
        hp[k] = D;
    }
    if (refresh) {
        // count tables only for images some pair tests with a LegalCnt rule
        std::map<const xc_tile *, bool> needz;
        for (int i = 0; i < n; ++i)
            if (pairs[i].cnt.enabled) {
                needz[pairs[i].b] = true;
                if (pair_rot[i] < 0) needz[pairs[i].a] = true;
            }
        TileDev *ht = (TileDev *)hat(i_tiles);
        for (size_t ti = 0; ti < tiles.size(); ++ti) {
            ht[ti] = tiles[ti]->dev;
            if (!needz[tiles[ti]]) ht[ti].Z = nullptr;
        }
        if (!sflat.empty()) memcpy(hat(i_slist), sflat.data(), sizeof(int) * sflat.size());
        for (size_t gi = 0; gi < sgroups.size(); ++gi) sgroups[gi].list = (const int *)at(i_slist) + soff[gi];
    }
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        Bucket &B = buckets[bi];
        const int G = row_group<float>(B.NH), C = col_group<float>(B.NY), GR = r2c_rows<float>(B.NH);
        int4 *wr = (int4 *)hat(i_wr[bi]);
        int2 *wc = (int2 *)hat(i_wc[bi]);
        int2 *wc2 = (int2 *)hat(i_wc2[bi]);
        R2cU8 *w8 = r2c8[bi] ? (R2cU8 *)hat(i_w8[bi]) : nullptr;
        // an 8-bit work item: rows y .. y + GR - 1 of view V into its spectrum
        auto item8 = [&](const ImgView &V, void *spec, int nh, int y) {
            R2cU8 it;
            it.src = (const uint8_t *)V.data + (long long)y * V.stride;
            it.dst = (cpx<float> *)spec + (long long)y * (nh + 1);
            it.msd = V.msd;
            it.mask = V.mask ? V.mask + (long long)y * V.mstride : nullptr;
            it.nrows = std::min(GR, V.h - y);
            it.w = V.w;
            it.stride = V.stride;
            it.mstride = V.mstride;
            return it;
        };
        int a = 0, c = 0, d = 0;
        for (int k = B.first; k < B.first + B.count; ++k) {
            const int i = p->order[k];
            const PairDev &D = hp[k];
            if (ownA[i])
                for (int y = 0; y < D.h1; y += GR) {
                    if (w8) w8[a] = item8(D.A, D.specA, B.NH, y);
                    wr[a++] = make_int4(k, 0, y, 0);
                }
            if (ownB[i])
                for (int y = 0; y < D.h2; y += GR) {
                    if (w8) w8[a] = item8(D.B, D.specB, B.NH, y);
                    wr[a++] = make_int4(k, 1, y, 0);
                }
            for (int x = 0; x < D.Nx / 2 + 1; x += C) wc[c++] = make_int2(k, x);
            live_row_groups(D, G, starts);
            for (int y : starts) wc2[d++] = make_int2(k, y);
        }
        B.d_r2c = (const int4 *)at(i_wr[bi]);
        B.d_r2c_u8 = r2c8[bi] ? (const R2cU8 *)at(i_w8[bi]) : nullptr;
        B.r2c_gen = r2c8[bi] && r2cg[bi];
        B.d_col = (const int2 *)at(i_wc[bi]);
        B.d_c2r = (const int2 *)at(i_wc2[bi]);
    }
    XC_HIP(ctx, hipMemcpy(base + up_lo, up.data(), up.size(), hipMemcpyHostToDevice));

    p->ex.fp64 = false;
    p->ex.ntiles = refresh ? (int)tiles.size() : 0;
    p->ex.d_tiles = (const TileDev *)at(i_tiles);
    p->ex.sat_groups = sgroups;
    p->ex.d_pairs = (const PairDev *)at(i_pairs);
    p->ex.buckets = buckets;
    p->ex.d_rot = (const RotDev *)at(i_rot);
    p->ex.d_rot_work = (const int2 *)at(i_rw);
    p->ex.n_rot_work = nrw;
    p->ex.d_rnorm = (const RotNorm *)at(i_rn);
    p->ex.d_rnorm_work = (const int2 *)at(i_rnw);
    p->ex.n_rnorm_work = nnw;
    p->ex.d_dtiles = (const TileDev *)at(i_dtiles);
    p->ex.dsat_groups = dgroups;

    // ---- algorithmic bytes ----
    for (int s = 0; s < ST_COUNT; ++s) p->stage_bytes[s] = 0.0;
    std::map<const void *, bool> prepped, specced;
    p->model_bytes = 0.0;
    for (const xc_tile *t : ders) {  // derived tables: pixels read, FP64 pairs written
        const double b = (1.0 + sizeof(double2)) * t->w * t->h;
        p->stage_bytes[ST_PREP] += b;
        p->model_bytes += b;
    }
    for (int i = 0; i < n; ++i) {
        const xc_pair &q = pairs[i];
        const PairDev &D = P[i];
        const bool rotated = pair_rot[i] >= 0;
        const int ea = rotated ? 4 : elem_size(q.a->dtype), eb = elem_size(q.b->dtype);
        double b[ST_COUNT];
        const bool pa = (refresh || rotated) && !prepped[rotated ? (const void *)&rots[pair_rot[i]] : q.a];
        prepped[rotated ? (const void *)&rots[pair_rot[i]] : q.a] = true;
        const bool pb = refresh && !prepped[q.b];
        prepped[q.b] = true;
        pair_stage_bytes(D, (int)cs, ea, eb, pa, pb, b);
        // a shared row spectrum is produced (and its input read) once
        const double H = D.Nx / 2 + 1;
        if (!ownA[i]) b[ST_R2C] -= ea * (double)D.w1 * D.h1 + cs * H * D.h1;
        if (!ownB[i]) b[ST_R2C] -= eb * (double)D.w2 * D.h2 + cs * H * D.h2;
        for (int s = 0; s < ST_COUNT; ++s) p->stage_bytes[s] += b[s];
        p->model_bytes += b[ST_PREP] + b[ST_R2C] + b[ST_COL] + b[ST_C2R] -
                          (pa ? ea * (double)D.w1 * D.h1 : 0.0) - (pb ? eb * (double)D.w2 * D.h2 : 0.0);
    }

    if (flags & XC_PLAN_AUTOTUNE) {
        st = tune_shapes(ctx, p);
        if (st) return st;
    }

    // ---- pinned results + graph ----
    XC_HIP(ctx, hipHostMalloc((void **)&p->h_res, sizeof(xc_result) * n, hipHostMallocDefault));
    if (flags & XC_PLAN_AUTOTUNE) {
        st = tune_sweep_share(ctx, p);
        if (st) return st;
    }
    return plan_capture(ctx, p);
}

static int plan_create(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, const double *aff4, xc_plan **out)
{
    if (!ctx || !pairs || n <= 0 || !out) return XC_E_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    xc_plan *p = new xc_plan;
    p->ctx = ctx;
    p->n = n;
    int st = plan_build(ctx, pairs, n, flags, p, aff4);
    if (st) {
        xc_plan_destroy(p);
        return st;
    }
    *out = p;
    return XC_OK;
}

extern "C" int xc_plan_create(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, xc_plan **out)
{
    return plan_create(ctx, pairs, n, flags, nullptr, out);
}

extern "C" int xc_plan_buckets(const xc_plan *p, xc_bucket_info *out, int cap)
{
    if (!p) return XC_E_ARG;
    const int nb = (int)p->ex.buckets.size();
    for (int i = 0; i < nb && i < cap && out; ++i) {
        const Bucket &b = p->ex.buckets[i];
        xc_bucket_info &o = out[i];
        o.nh = b.NH;
        o.ny = b.NY;
        o.whiten = b.whiten ? 1 : 0;
        o.pairs = b.count;
        o.col_c = b.col_c;
        o.col_xcd = b.col_xcd;
        o.col_us = b.col_us;
        o.r2c_rpt = b.r2c_rpt ? b.r2c_rpt : r2c_default_rpt<float>(b.NH);
        o.r2c_np = b.d_r2c_u8 ? (b.r2c_np ? b.r2c_np : R2C_U8_NP) : 0;
        o.r2c_us = b.r2c_us;
    }
    return nb;
}

extern "C" int xc_plan_set_shapes(xc_plan *p, int bucket, int col_c, int col_xcd, int r2c_rpt, int r2c_np)
{
    if (!p || bucket < 0 || bucket >= (int)p->ex.buckets.size()) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    Bucket &b = p->ex.buckets[bucket];
    if (col_c != 0 && !col_shape_ok(ctx, b.NY, col_c))
        return set_err(ctx, XC_E_SHAPE, "no column pass of %d columns per workgroup for %d-point columns", col_c, b.NY);
    if (r2c_rpt != 0 && r2c_rpt != r2c_default_rpt<float>(b.NH) && r2c_rpt != r2c_alt_rpt<float>(b.NH))
        return set_err(ctx, XC_E_SHAPE, "no row r2c with %d rows per thread for %d-point rows", r2c_rpt, 2 * b.NH);
    if (r2c_np != 0 && r2c_np != 1 && r2c_np != 2 && r2c_np != 4)
        return set_err(ctx, XC_E_SHAPE, "no 8-bit row r2c with %d work items per workgroup", r2c_np);
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // a launch in flight keeps its graph
    b.col_c = col_c;
    b.col_xcd = col_xcd ? 1 : 0;
    b.r2c_rpt = r2c_rpt;
    b.r2c_np = b.d_r2c_u8 ? r2c_np : 0;
    b.col_us = 0.0;
    b.r2c_us = 0.0;
    return plan_capture(ctx, p);
}

extern "C" int xc_plan_sweep_waves(const xc_plan *p, int32_t *waves, double *us)
{
    if (!p || !waves) return XC_E_ARG;
    *waves = p->ex.sat_waves;
    if (us) *us = p->sweep_us;
    return XC_OK;
}

extern "C" int xc_plan_set_sweep_waves(xc_plan *p, int waves)
{
    if (!p) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    if (waves < 0 || waves > 32) return set_err(ctx, XC_E_ARG, "sweep share of %d wavefronts per CU (0..32)", waves);
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // a launch in flight keeps its graph
    p->ex.sat_waves = waves;
    p->sweep_us = 0.0;
    return plan_capture(ctx, p);
}

extern "C" int xc_plan_launch(xc_plan *p)
{
    if (!p) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    XC_HIP(ctx, hipGraphLaunch(p->exec, ctx->stream));
    return XC_OK;
}

extern "C" int xc_plan_wait(xc_plan *p, xc_result *results)
{
    if (!p) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    if (results) XC_HIP(ctx, hipMemcpyAsync(p->h_res, p->d_res, sizeof(xc_result) * p->n, hipMemcpyDeviceToHost, ctx->stream));
    XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (results) memcpy(results, p->h_res, sizeof(xc_result) * p->n);
    return XC_OK;
}

extern "C" int xc_plan_run(xc_plan *p, xc_result *results)
{
    if (!p) return XC_E_ARG;
    std::lock_guard<std::mutex> lock(p->ctx->mu);
    int st = xc_plan_launch(p);
    if (st) return st;
    return xc_plan_wait(p, results);
}

extern "C" int xc_plan_stage_times(xc_plan *p, int iters, double *ms, double *bytes)
{
    if (!p || iters <= 0) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipEvent_t ev[ST_COUNT + 1];
    for (auto &e : ev) XC_HIP(ctx, hipEventCreate(&e));
    double acc[ST_COUNT] = {0, 0, 0, 0, 0};
    int st = XC_OK;
    for (int it = 0; it < iters && !st; ++it) {
        XC_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
        for (int s = 0; s < ST_COUNT && !st; ++s) {
            st = enqueue(ctx, p->ex, s, s + 1);
            XC_HIP(ctx, hipEventRecord(ev[s + 1], ctx->stream));
        }
        XC_HIP(ctx, hipEventSynchronize(ev[ST_COUNT]));
        for (int s = 0; s < ST_COUNT; ++s) {
            float t = 0.f;
            XC_HIP(ctx, hipEventElapsedTime(&t, ev[s], ev[s + 1]));
            acc[s] += t;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (st) return st;
    for (int s = 0; s < ST_COUNT; ++s) {
        if (ms) ms[s] = acc[s] / iters;
        if (bytes) bytes[s] = p->stage_bytes[s];
    }
    return XC_OK;
}

extern "C" double xc_plan_model_bytes(const xc_plan *p) { return p ? p->model_bytes : 0.0; }

extern "C" int xc_plan_destroy(xc_plan *p)
{
    if (!p) return XC_E_ARG;
    xc_ctx *ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->h_res) (void)hipHostFree(p->h_res);
    if (p->ws.p) (void)hipFree(p->ws.p);
    delete p;
    return XC_OK;
}

extern "C" int xc_batch_corr_f32(xc_ctx *ctx, const xc_pair *pairs, int n, xc_result *results)
{
    if (!results) return XC_E_ARG;
    xc_plan *p = nullptr;
    int st = xc_plan_create(ctx, pairs, n, 0, &p);
    if (st) return st;
    st = xc_plan_run(p, results);
    xc_plan_destroy(p);
    return st;
}

extern "C" int xc_sweep_angles_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *angles_deg,
                                   int n_angles, const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, int Ox, int Oy,
                                   int Rx, int Ry, xc_result *results)
{
    if (!ctx || !a || !b || !angles_deg || n_angles <= 0 || !results) return XC_E_ARG;
    std::vector<xc_pair> pairs(n_angles);
    for (int i = 0; i < n_angles; ++i) {
        xc_pair &q = pairs[i];
        memset(&q, 0, sizeof q);
        q.a = a;
        q.b = b;
        q.mode = XC_MODE_S;
        q.Ox = Ox;
        q.Oy = Oy;
        q.Rx = Rx;
        q.Ry = Ry;
        q.angle_deg = angles_deg[i];  // R(0) is the identity: 0 deg needs no rotated copy
        if (rgn) q.rgn = *rgn;
        if (cnt) q.cnt = *cnt;
    }
    xc_plan *p = nullptr;
    int st = xc_plan_create(ctx, pairs.data(), n_angles, 0, &p);
    if (st) return st;
    st = xc_plan_run(p, results);
    xc_plan_destroy(p);
    return st;
}

namespace xc {

// One batched sweep of the 2 x 2 maps maps[4i..4i+3] applied to A's pixel grid, each correlation (CorrImagesS)
// with its own search oval ovals[4i..4i+3] = (Ox, Oy, Rx, Ry).
int sweep_affine_ovals(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *maps, const int *ovals,
                       int n_maps, const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, xc_result *results)
{
    if (!ctx || !a || !b || !maps || !ovals || n_maps <= 0 || !results) return XC_E_ARG;
    std::vector<xc_pair> pairs(n_maps);
    for (int i = 0; i < n_maps; ++i) {
        xc_pair &q = pairs[i];
        memset(&q, 0, sizeof q);
        q.a = a;
        q.b = b;
        q.mode = XC_MODE_S;
        q.Ox = ovals[4 * i + 0];
        q.Oy = ovals[4 * i + 1];
        q.Rx = ovals[4 * i + 2];
        q.Ry = ovals[4 * i + 3];
        if (rgn) q.rgn = *rgn;
        if (cnt) q.cnt = *cnt;
    }
    xc_plan *p = nullptr;
    int st = plan_create(ctx, pairs.data(), n_maps, 0, maps, &p);
    if (st) return st;
    st = xc_plan_run(p, results);
    xc_plan_destroy(p);
    return st;
}

}  // namespace xc

extern "C" int xc_sweep_affine_f32(xc_ctx *ctx, const xc_tile *a, const xc_tile *b, const double *maps, int n_maps,
                                   const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, int Ox, int Oy, int Rx, int Ry,
                                   xc_result *results)
{
    if (!ctx || !a || !b || !maps || n_maps <= 0 || !results) return XC_E_ARG;
    std::vector<int> ovals(4 * (size_t)n_maps);
    for (int i = 0; i < n_maps; ++i) {
        ovals[4 * i + 0] = Ox;
        ovals[4 * i + 1] = Oy;
        ovals[4 * i + 2] = Rx;
        ovals[4 * i + 3] = Ry;
    }
    return sweep_affine_ovals(ctx, a, b, maps, ovals.data(), n_maps, rgn, cnt, results);
}
