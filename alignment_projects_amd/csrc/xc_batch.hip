// xc_batch.hip -- device-resident tiles and batched FP32 plans.
//
// A tile is one image kept in HBM with its FP64 summed-area tables and its
// (mean, sd); it plays the role of the reference's fft2 / ThmRec.ftc cache
// (Correlation_fft_fftw.cpp:50-51, CThmScan.h:27-34), made explicit.  A plan
// is a fixed batch of pair correlations laid out once (descriptors, work
// lists, spectra, lag rows) and replayed as one HIP graph.
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "xc_exec.hpp"
#include "xc_kernels.hpp"

using namespace xc;

struct xc_tile {
    xc_ctx *ctx = nullptr;
    int w = 0, h = 0, dtype = 0, flags = 0;
    void *raw = nullptr;       // device pixels (owned unless adopted)
    bool own_raw = false;
    uint8_t *mask = nullptr;   // device mask (owned unless adopted)
    bool own_mask = false;
    DevBuf mem;                // tables + scratch + descriptor
    TileDev dev;               // host copy, device pointers
    TileDev *d_dev = nullptr;  // device copy
    int64_t bytes = 0;
};

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
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    xc_tile *t = new xc_tile;
    t->ctx = ctx;
    t->w = w;
    t->h = h;
    t->dtype = dtype;
    t->flags = flags;
    const bool dev_src = (flags & XC_TILE_DEVICE_SRC) != 0;
    const TileSizes ts = tile_sizes(w, h);
    if (w > TP_MAX_W) {
        delete t;
        return set_err(ctx, XC_E_SHAPE, "tile wider than %d", TP_MAX_W);
    }
    Layout lay;
    size_t iS = lay.add(ts.S), iQ = lay.add(ts.Q), iZ = lay.add(ts.Z), iP = lay.add(ts.part);
    size_t iM = lay.add(ts.msd), iD = lay.add(sizeof(TileDev)), iL = lay.add(sizeof(int));
    size_t iRaw = dev_src ? 0 : lay.add((size_t)es * stride * h);
    size_t iMask = (!dev_src && mask) ? lay.add((size_t)w * h) : 0;
    int st = ensure_dev(ctx, t->mem, lay.total);
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
        if (hipMemcpy(t->raw, data, (size_t)es * stride * h, hipMemcpyHostToDevice) != hipSuccess) {
            xc_tile_release(ctx, t);
            return set_err(ctx, XC_E_DEVICE, "tile upload failed");
        }
        if (mask) {
            t->mask = (uint8_t *)(base + lay.off[iMask]);
            t->own_mask = true;
            if (hipMemcpy(t->mask, mask, (size_t)w * h, hipMemcpyHostToDevice) != hipSuccess) {
                xc_tile_release(ctx, t);
                return set_err(ctx, XC_E_DEVICE, "tile mask upload failed");
            }
        }
    }
    TileDev &D = t->dev;
    memset(&D, 0, sizeof D);
    D.v.data = t->raw;
    D.v.mask = t->mask;
    D.v.dtype = dtype;
    D.v.w = w;
    D.v.h = h;
    D.v.stride = stride;
    D.v.mstride = w;
    D.msd = (double *)(base + lay.off[iM]);
    D.v.msd = D.msd;
    D.S = (double *)(base + lay.off[iS]);
    D.Q = (double *)(base + lay.off[iQ]);
    D.Z = (int32_t *)(base + lay.off[iZ]);
    D.part = (double *)(base + lay.off[iP]);
    D.normalize = (flags & XC_TILE_NORMALIZE) ? 1 : 0;
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
    sg.cpt = sat_cpt(w);
    sg.threads = sat_threads(w);
    sg.count = 1;
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

static int plan_build(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, xc_plan *p)
{
    // ---- pair geometry ----
    std::vector<PairDev> P(n);
    std::vector<const xc_tile *> tiles;
    std::map<const xc_tile *, int> tile_idx;
    for (int i = 0; i < n; ++i) {
        const xc_pair &q = pairs[i];
        const xc_tile *a = q.a, *b = q.b;
        if (!a || !b) return set_err(ctx, XC_E_ARG, "pair %d: null tile", i);
        if (q.angle_deg != 0.0) return set_err(ctx, XC_E_UNSUPPORTED, "pair %d: rotated pairs go through xc_sweep_angles_f32", i);
        if (q.mode != XC_MODE_NCC && q.mode != XC_MODE_R && q.mode != XC_MODE_S)
            return set_err(ctx, XC_E_ARG, "pair %d: bad mode", i);
        PairDev &D = P[i];
        memset(&D, 0, sizeof D);
        const int mode = q.mode == XC_MODE_NCC ? PM_NCC : (q.mode == XC_MODE_R ? PM_R : PM_S);
        setup_pair_geometry(D, a->w, a->h, b->w, b->h, mode);
        if (mode != PM_NCC) {
            const int md = 11 / 2 + 1;
            if (D.nnegx < md || D.nposx < md || D.nnegy < md || D.nposy < md)
                return set_err(ctx, XC_E_SHAPE, "pair %d: too small for CorrImages (SetDims)", i);
        }
        if (D.Nx / 2 > max_nh<float>() || D.Ny > max_ny<float>())
            return set_err(ctx, XC_E_SHAPE, "pair %d: padded size %d x %d too large", i, D.Nx, D.Ny);
        if (q.rgn.enabled && q.rgn.table)
            return set_err(ctx, XC_E_UNSUPPORTED, "pair %d: tabulated LegalRgn in batched form", i);
        D.Ox = q.Ox;
        D.Oy = q.Oy;
        D.radius = q.radius;
        D.Rx = q.Rx;
        D.Ry = q.Ry;
        D.rule = make_rule(&q.rgn, &q.cnt, nullptr);
        D.need_cnt = D.rule.cnt_enabled;
        D.A = a->dev.v;
        D.B = b->dev.v;
        D.SA = a->dev.S;
        D.QA = a->dev.Q;
        D.ZA = a->dev.Z;
        D.SB = b->dev.S;
        D.QB = b->dev.Q;
        D.ZB = b->dev.Z;
        for (const xc_tile *t : {a, b})
            if (!tile_idx.count(t)) {
                tile_idx[t] = (int)tiles.size();
                tiles.push_back(t);
            }
    }
    std::vector<Bucket> buckets;
    make_buckets(P, 0, p->order, buckets);

    // ---- layout ----
    const bool refresh = (flags & 1) != 0;
    Layout lay;
    std::vector<size_t> i_spA(n), i_spB(n), i_lag(n), i_rows(n);
    const size_t cs = sizeof(cpx<float>);
    for (int k = 0; k < n; ++k) {
        const PairDev &D = P[p->order[k]];
        const size_t H = D.Nx / 2 + 1;
        i_spA[k] = lay.add(cs * H * D.h1);
        i_spB[k] = lay.add(cs * H * D.h2);
        i_lag[k] = lay.add(cs * H * D.hR);
        i_rows[k] = lay.add(sizeof(RowRec) * D.hR);
    }
    size_t i_res = lay.add(sizeof(xc_result) * n);
    // uploaded block: pairs, tiles, work lists
    size_t i_pairs = lay.add(sizeof(PairDev) * n);
    size_t i_tiles = lay.add(sizeof(TileDev) * (refresh ? tiles.size() : 0) + 16);
    std::vector<SatGroup> sgroups;
    std::vector<int> sflat, soff;
    if (refresh) {
        std::vector<int> widths;
        for (const xc_tile *t : tiles) widths.push_back(t->w);
        make_sat_groups(widths, sgroups, sflat, soff);
    }
    size_t i_slist = lay.add(sizeof(int) * sflat.size() + 16);
    std::vector<size_t> i_wr(buckets.size()), i_wc(buckets.size()), i_wc2(buckets.size());
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
        Bucket &B = buckets[bi];
        const int G = row_group<float>(B.NH), C = col_group<float>(B.NY);
        int nr = 0, nc = 0, nc2 = 0;
        for (int k = B.first; k < B.first + B.count; ++k) {
            const PairDev &D = P[p->order[k]];
            nr += (D.h1 + G - 1) / G + (D.h2 + G - 1) / G;
            nc += (D.Nx / 2 + 1 + C - 1) / C;
            nc2 += (D.hR + G - 1) / G;
        }
        B.n_r2c = nr;
        B.n_col = nc;
        B.n_c2r = nc2;
        i_wr[bi] = lay.add(sizeof(int4) * nr);
        i_wc[bi] = lay.add(sizeof(int2) * nc);
        i_wc2[bi] = lay.add(sizeof(int2) * nc2);
    }
    const size_t up_lo = lay.off[i_pairs];
    const size_t up_hi = lay.total;
    int st = ensure_dev(ctx, p->ws, lay.total);
    if (st) return st;
    char *base = (char *)p->ws.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };

    // ---- host copy of the uploaded block ----
    std::vector<char> up(up_hi - up_lo);
    auto hat = [&](size_t i) { return up.data() + (lay.off[i] - up_lo); };
    PairDev *hp = (PairDev *)hat(i_pairs);
    p->d_res = (xc_result *)at(i_res);
    for (int k = 0; k < n; ++k) {
        PairDev D = P[p->order[k]];
        D.specA = at(i_spA[k]);
        D.specB = at(i_spB[k]);
        D.lag = at(i_lag[k]);
        D.rows = (RowRec *)at(i_rows[k]);
        D.out = p->d_res + p->order[k];
        hp[k] = D;
    }
    if (refresh) {
        // count tables only for images some pair tests with a LegalCnt rule
        std::map<const xc_tile *, bool> needz;
        for (int i = 0; i < n; ++i)
            if (pairs[i].cnt.enabled) needz[pairs[i].a] = needz[pairs[i].b] = true;
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
        const int G = row_group<float>(B.NH), C = col_group<float>(B.NY);
        int4 *wr = (int4 *)hat(i_wr[bi]);
        int2 *wc = (int2 *)hat(i_wc[bi]);
        int2 *wc2 = (int2 *)hat(i_wc2[bi]);
        int a = 0, c = 0, d = 0;
        for (int k = B.first; k < B.first + B.count; ++k) {
            const PairDev &D = hp[k];
            for (int y = 0; y < D.h1; y += G) wr[a++] = make_int4(k, 0, y, 0);
            for (int y = 0; y < D.h2; y += G) wr[a++] = make_int4(k, 1, y, 0);
            for (int x = 0; x < D.Nx / 2 + 1; x += C) wc[c++] = make_int2(k, x);
            for (int y = 0; y < D.hR; y += G) wc2[d++] = make_int2(k, y);
        }
        B.d_r2c = (const int4 *)at(i_wr[bi]);
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

    // ---- algorithmic bytes ----
    for (int s = 0; s < ST_COUNT; ++s) p->stage_bytes[s] = 0.0;
    std::map<const xc_tile *, bool> prepped;
    p->model_bytes = 0.0;
    for (int i = 0; i < n; ++i) {
        const xc_pair &q = pairs[i];
        double b[ST_COUNT];
        const bool pa = refresh && !prepped[q.a];
        prepped[q.a] = true;
        const bool pb = refresh && !prepped[q.b];
        prepped[q.b] = true;
        pair_stage_bytes(P[i], (int)cs, elem_size(q.a->dtype), elem_size(q.b->dtype), pa, pb, b);
        for (int s = 0; s < ST_COUNT; ++s) p->stage_bytes[s] += b[s];
        // whole-path model: inputs once, row spectra, lag rows, tables written and read once
        double m[ST_COUNT];
        pair_stage_bytes(P[i], (int)cs, elem_size(q.a->dtype), elem_size(q.b->dtype), true, true, m);
        p->model_bytes += m[ST_PREP] - (double)elem_size(q.a->dtype) * P[i].w1 * P[i].h1 -
                          (double)elem_size(q.b->dtype) * P[i].w2 * P[i].h2 + m[ST_R2C] + m[ST_COL] + m[ST_C2R];
    }

    // ---- pinned results + graph ----
    XC_HIP(ctx, hipHostMalloc((void **)&p->h_res, sizeof(xc_result) * n, hipHostMallocDefault));
    XC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    st = enqueue(ctx, p->ex);
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

extern "C" int xc_plan_create(xc_ctx *ctx, const xc_pair *pairs, int n, int flags, xc_plan **out)
{
    if (!ctx || !pairs || n <= 0 || !out) return XC_E_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    xc_plan *p = new xc_plan;
    p->ctx = ctx;
    p->n = n;
    int st = plan_build(ctx, pairs, n, flags, p);
    if (st) {
        xc_plan_destroy(p);
        return st;
    }
    *out = p;
    return XC_OK;
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
