// xc_plain.hip -- plain FP64 form: one pair per call, on the device, with
// the reference's CorrPatches / CorrImagesR / CorrImagesS semantics.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "xc_exec.hpp"
#include "xc_fmode.hpp"
#include "xc_host.hpp"
#include "xc_kernels.hpp"

using namespace xc;

namespace {

// The staging copy of both lists into pinned memory (24 bytes per point, with the on-grid test of the
// splat), split over a few host threads for large lists: one thread copies about 10 GB/s, which would
// otherwise bound a call on two 2048 x 256 lists.
void stage_lists(xc_point *dp, double *dv, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
                 const double *iv2, int n2, bool *og1, bool *og2)
{
    struct OffGrid {
        bool a = false, b = false;
    };
    const long long total = (long long)n1 + n2;
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = total < (1 << 17) ? 1 : std::max(1, std::min(8, hw));
    // points [lo, hi) of the two lists taken as one
    auto part = [&](long long lo, long long hi, OffGrid *o) {
        if (lo < n1) {
            const int a = (int)lo, b = (int)std::min<long long>(hi, n1);
            o->a = copy_points_off_grid(dp + a, ip1 + a, b - a);
            memcpy(dv + a, iv1 + a, sizeof(double) * (b - a));
        }
        if (hi > n1) {
            const int a = (int)(std::max<long long>(lo, n1) - n1), b = (int)(hi - n1);
            o->b = copy_points_off_grid(dp + n1 + a, ip2 + a, b - a);
            memcpy(dv + n1 + a, iv2 + a, sizeof(double) * (b - a));
        }
    };
    std::vector<OffGrid> o(nt);
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    const long long step = (total + nt - 1) / nt;
    for (int t = 1; t < nt; ++t) {
        const long long lo = std::min(total, t * step), hi = std::min(total, lo + step);
        if (lo >= hi) continue;
        try {
            th.emplace_back([&part, &o, t, lo, hi] { part(lo, hi, &o[t]); });
        } catch (...) {  // no thread to be had: that part on this one
            part(lo, hi, &o[t]);
        }
    }
    part(0, std::min(total, step), &o[0]);
    for (auto &x : th) x.join();
    for (const OffGrid &x : o) {
        *og1 = *og1 || x.a;
        *og2 = *og2 || x.b;
    }
}

}  // namespace

namespace xc {

// Run one FP64 correlation of point lists (A = ip1/iv1, B = ip2/iv2).
// fmode (with mode PM_R): CorrImagesF -- the R image is kept and the peak
// is chosen by MakeF / OrderF / FPeak with nbmaxht.
// fixed (may be null): a device workspace the call must fit in; then the call
// allocates nothing (no device arena, no pinned staging: the lists are copied
// from the caller's memory, the descriptors from host memory of this call).
int run_plain(xc_ctx *ctx, int mode, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
              const double *iv2, int n2, int Ox, int Oy, int radius, int Rx, int Ry, double mincor,
              const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, xc_result *out, bool fmode, double nbmaxht,
              const FixedWs *fixed)
{
    if (!ctx) return XC_E_ARG;
    if (!ip1 || !iv1 || !ip2 || !iv2 || n1 <= 0 || n2 <= 0)
        return set_err(ctx, XC_E_ARG, "empty or null point list");
    if (rgn && rgn->enabled && rgn->table && (rgn->table_w <= 0 || rgn->table_h <= 0))
        return set_err(ctx, XC_E_ARG, "legality table without extent");
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));

    IBox B1, B2;
    bbox_from_points(B1, ip1, n1);
    bbox_from_points(B2, ip2, n2);
    const int w1 = B1.R - B1.L + 1, h1 = B1.T - B1.B + 1;
    const int w2 = B2.R - B2.L + 1, h2 = B2.T - B2.B + 1;

    PairDev P;
    memset(&P, 0, sizeof P);
    setup_pair_geometry(P, w1, h1, w2, h2, mode);
    P.Ox = Ox;
    P.Oy = Oy;
    P.radius = radius;
    P.Rx = Rx;
    P.Ry = Ry;
    P.offx = B2.L - B1.L;
    P.offy = B2.B - B1.B;
    P.mincor = mincor;

    if (mode != PM_NCC) {
        // CCorImg::SetDims (Correlation.cpp:1949-2003)
        const int mindelta = 11 / 2 + 1;
        if (P.nnegx < mindelta || P.nposx < mindelta || P.nnegy < mindelta || P.nposy < mindelta) {
            memset(out, 0, sizeof *out);
            out->status = XC_TOO_SMALL;
            out->nx = P.Nx_ref;
            out->ny = P.Ny_ref;
            return XC_TOO_SMALL;
        }
        if (mode == PM_S && mincor > 0.0)
            return set_err(ctx, XC_E_UNSUPPORTED, "S mode with mincor > 0 needs the R image too");
    }
    const int NH = P.Nx / 2, NY = P.Ny, H = NH + 1;
    if (NH > max_nh<double>() || NY > max_ny<double>())
        return set_err(ctx, XC_E_SHAPE, "padded size %d x %d exceeds the FP64 limit", P.Nx, P.Ny);

    // ---- workspace layout ----
    const size_t n = (size_t)n1 + n2;
    const size_t tbl = (rgn && rgn->enabled && rgn->table) ? (size_t)rgn->table_w * rgn->table_h : 0;
    const int G = row_group<double>(NH), C = col_group<double>(NY), GR = r2c_rows<double>(NH);
    const int nr2c = (h1 + GR - 1) / GR + (h2 + GR - 1) / GR;
    const int ncol = (H + C - 1) / C;
    const int nc2r = (P.hR + G - 1) / G;
    const TileSizes t1 = tile_sizes(w1, h1, false), t2 = tile_sizes(w2, h2, false);
    Layout lay;
    size_t i_pts = lay.add(sizeof(xc_point) * n), i_vals = lay.add(sizeof(double) * n);
    size_t i_iA = lay.add(sizeof(double) * w1 * h1), i_iB = lay.add(sizeof(double) * w2 * h2);
    size_t i_SA = lay.add(t1.SQ), i_ZA = lay.add(t1.Z);
    size_t i_SB = lay.add(t2.SQ), i_ZB = lay.add(t2.Z);
    size_t i_pA = lay.add(t1.part), i_mA = lay.add(t1.msd), i_bA = lay.add(t1.band);
    size_t i_pB = lay.add(t2.part), i_mB = lay.add(t2.msd), i_bB = lay.add(t2.band);
    size_t i_spA = lay.add(sizeof(cpx<double>) * h1 * H), i_spB = lay.add(sizeof(cpx<double>) * h2 * H);
    size_t i_lag = lay.add(sizeof(cpx<double>) * (size_t)P.hR * H);
    size_t i_rows = lay.add(sizeof(RowRec) * P.hR);
    size_t i_res = lay.add(sizeof(xc_result));
    // uploaded block 2: pair, tiles, work lists, table
    size_t i_pair = lay.add(sizeof(PairDev));
    size_t i_tiles = lay.add(sizeof(TileDev) * 2);
    size_t i_wr = lay.add(sizeof(int4) * nr2c), i_wc = lay.add(sizeof(int2) * ncol), i_wc2 = lay.add(sizeof(int2) * nc2r);
    size_t i_slist = lay.add(sizeof(int) * 2);
    size_t i_tbl = lay.add(tbl ? tbl : 1);
    const long long nR = (long long)P.wR * P.hR;
    size_t i_f = lay.add(fmode ? fmode_bytes(nR) : 1);
    // scratch for an ordered splat, in case a list is off the grid (found while staging it; with a fixed
    // workspace the lists are tested first and the scratch is reserved only for one that is off it)
    bool og1 = false, og2 = false;
    if (fixed) {
        og1 = points_off_grid(ip1, n1);
        og2 = points_off_grid(ip2, n2);
    }
    const size_t spl = std::max(fixed && !og1 ? 0 : splat_scratch_bytes(n1, (long long)w1 * h1),
                                fixed && !og2 ? 0 : splat_scratch_bytes(n2, (long long)w2 * h2));
    size_t i_spl = lay.add(spl ? spl : 1);
    int st;
    char *base;
    if (fixed) {
        if (lay.total > fixed->bytes)
            return set_err(ctx, XC_E_SHAPE, "the pair needs %zu bytes of device workspace; the fixed scratch holds %zu",
                           lay.total, fixed->bytes);
        base = (char *)fixed->dev;
    } else {
        if ((st = ensure_dev(ctx, ctx->arena, lay.total))) return st;
        base = (char *)ctx->arena.p;
    }
    auto at = [&](size_t i) { return base + lay.off[i]; };

    TileDev T[2];
    memset(T, 0, sizeof T);
    for (int k = 0; k < 2; ++k) {
        TileDev &D = T[k];
        D.v.data = at(k ? i_iB : i_iA);
        D.v.dtype = XC_F64;
        D.v.w = k ? w2 : w1;
        D.v.h = k ? h2 : h1;
        D.v.stride = D.v.w;
        D.v.mstride = D.v.w;
        D.msd = (double *)at(k ? i_mB : i_mA);
        D.v.msd = D.msd;
        D.SQ = at(k ? i_SB : i_SA);
        D.Z = (int32_t *)at(k ? i_ZB : i_ZA);
        D.part = (double *)at(k ? i_pB : i_pA);
        D.bandp = (double *)at(k ? i_bB : i_bA);
        D.nbands = sat_bands(D.v.h);  // a lone pair: its two images swept in row bands
        D.normalize = 0;
    }
    P.A = T[0].v;
    P.B = T[1].v;
    P.SQA = T[0].SQ;
    P.ZA = T[0].Z;
    P.SQB = T[1].SQ;
    P.ZB = T[1].Z;
    P.specA = at(i_spA);
    P.specB = at(i_spB);
    P.lag = at(i_lag);
    P.rows = (RowRec *)at(i_rows);
    P.out = at(i_res);
    P.rule = make_rule(rgn, cnt, (const uint8_t *)at(i_tbl));
    P.need_cnt = P.rule.cnt_enabled;
    FWork fw;
    memset(&fw, 0, sizeof fw);
    if (fmode) {
        fmode_carve(at(i_f), nR, fw);
        P.rimg = fw.R;
        P.cimg = fw.cand;
    }
    narrow_rows(P);
    std::vector<int> c2r_starts;
    live_row_groups(P, G, c2r_starts);

    // ---- host staging: points + values, then descriptors ----
    const size_t s1_lo = lay.off[i_pts], s1_hi = lay.off[i_vals] + sizeof(double) * n;
    const size_t s2_lo = lay.off[i_pair], s2_hi = lay.off[i_tbl] + (tbl ? tbl : 1);
    char *hs = nullptr, *hs2;
    xc_result *hres;
    std::vector<char> hdesc;  // fixed workspace: the descriptor block in this call's host memory
    xc_result hres_local;
    if (fixed) {
        hdesc.resize(s2_hi - s2_lo);
        hs2 = hdesc.data();
        hres = &hres_local;
    } else {
        if ((st = ensure_pinned(ctx, (s1_hi - s1_lo) + (s2_hi - s2_lo) + sizeof(xc_result) + 64))) return st;
        hs = (char *)ctx->pinned;
        hs2 = hs + (s1_hi - s1_lo);
        hres = (xc_result *)(((uintptr_t)(hs2 + (s2_hi - s2_lo)) + 63) & ~(uintptr_t)63);
    }
    auto h2p = [&](size_t i) { return hs2 + (lay.off[i] - s2_lo); };
    if (!fixed) {
        auto h1p = [&](size_t i) { return hs + (lay.off[i] - s1_lo); };
        stage_lists((xc_point *)h1p(i_pts), (double *)h1p(i_vals), ip1, iv1, n1, ip2, iv2, n2, &og1, &og2);
    }
    memcpy(h2p(i_pair), &P, sizeof P);
    memcpy(h2p(i_tiles), T, sizeof T);
    int4 *wr = (int4 *)h2p(i_wr);
    int k = 0;
    for (int y = 0; y < h1; y += GR) wr[k++] = make_int4(0, 0, y, 0);
    for (int y = 0; y < h2; y += GR) wr[k++] = make_int4(0, 1, y, 0);
    int2 *wc = (int2 *)h2p(i_wc);
    for (int i = 0; i < ncol; ++i) wc[i] = make_int2(0, i * C);
    int2 *wc2 = (int2 *)h2p(i_wc2);
    for (size_t i = 0; i < c2r_starts.size(); ++i) wc2[i] = make_int2(0, c2r_starts[i]);
    if (tbl) memcpy(h2p(i_tbl), rgn->table, tbl);
    std::vector<SatGroup> sgroups;
    std::vector<int> sflat, soff;
    make_sat_groups({w1, w2}, {}, {}, sgroups, sflat, soff);
    for (SatGroup &g : sgroups) g.nb = std::max(sat_bands(h1), sat_bands(h2));
    memcpy(h2p(i_slist), sflat.data(), sizeof(int) * sflat.size());
    for (size_t gi = 0; gi < sgroups.size(); ++gi) sgroups[gi].list = (const int *)at(i_slist) + soff[gi];

    hipStream_t s = ctx->stream;
    if (fixed) {
        // straight from the caller's lists (the host memory stays valid until the call has synchronised)
        char *dp = base + lay.off[i_pts], *dv = base + lay.off[i_vals];
        XC_HIP(ctx, hipMemcpyAsync(dp, ip1, sizeof(xc_point) * n1, hipMemcpyHostToDevice, s));
        XC_HIP(ctx, hipMemcpyAsync(dp + sizeof(xc_point) * n1, ip2, sizeof(xc_point) * n2, hipMemcpyHostToDevice, s));
        XC_HIP(ctx, hipMemcpyAsync(dv, iv1, sizeof(double) * n1, hipMemcpyHostToDevice, s));
        XC_HIP(ctx, hipMemcpyAsync(dv + sizeof(double) * n1, iv2, sizeof(double) * n2, hipMemcpyHostToDevice, s));
    } else {
        XC_HIP(ctx, hipMemcpyAsync(base + s1_lo, hs, s1_hi - s1_lo, hipMemcpyHostToDevice, s));
    }
    XC_HIP(ctx, hipMemcpyAsync(base + s2_lo, hs2, s2_hi - s2_lo, hipMemcpyHostToDevice, s));

    // ---- rasters: zero then splat (ImageFromValuesAndPoints) ----
    {
        const long long nA = (long long)w1 * h1, nB = (long long)w2 * h2;
        const int fb = (int)std::min<long long>(1024, (nA + nB + 255) / 256);
        hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, s, (double *)at(i_iA), nA, 0.0);
        hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, s, (double *)at(i_iB), nB, 0.0);
        xc_point *dp = (xc_point *)at(i_pts);
        double *dv = (double *)at(i_vals);
        if ((st = splat_points(ctx, s, dp, dv, n1, (double)B1.L, (double)B1.B, (double *)at(i_iA), w1, h1, og1,
                               at(i_spl), spl)))
            return st;
        if ((st = splat_points(ctx, s, dp + n1, dv + n1, n2, (double)B2.L, (double)B2.B, (double *)at(i_iB), w2, h2,
                               og2, at(i_spl), spl)))
            return st;
    }
    // ---- tables, transforms, evaluation ----
    Exec ex;
    ex.fp64 = true;
    ex.ntiles = 2;
    ex.d_tiles = (const TileDev *)at(i_tiles);
    ex.sat_groups = sgroups;
    ex.d_pairs = (const PairDev *)at(i_pair);
    Bucket b;
    b.NH = NH;
    b.NY = NY;
    b.whiten = mode == PM_S;
    b.first = 0;
    b.count = 1;
    b.d_r2c = (const int4 *)at(i_wr);
    b.n_r2c = nr2c;
    b.d_col = (const int2 *)at(i_wc);
    b.n_col = ncol;
    b.d_c2r = (const int2 *)at(i_wc2);
    b.n_c2r = (int)c2r_starts.size();
    ex.buckets.push_back(b);
    if ((st = enqueue(ctx, ex))) return st;
    if (fmode && (st = fmode_select(ctx, s, ex.d_pairs, P.wR, P.hR, nbmaxht, fw))) return st;
    XC_HIP(ctx, hipMemcpyAsync(hres, at(i_res), sizeof(xc_result), hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    *out = *hres;
    return out->status;
}

}  // namespace xc

extern "C" int xc_corr_patches_f64(xc_ctx *ctx, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
                                   const double *iv2, int n2, int Ox, int Oy, int radius, const xc_rgn_rule *rgn,
                                   const xc_cnt_rule *cnt, double *dx, double *dy, double *r, xc_result *res)
{
    xc_result R;
    memset(&R, 0, sizeof R);
    int st = run_plain(ctx, PM_NCC, ip1, iv1, n1, ip2, iv2, n2, Ox, Oy, radius, 0, 0, 0.0, rgn, cnt, &R, false, 0.0,
                       nullptr);
    if (st == XC_OK) {
        if (dx) *dx = R.dx;
        if (dy) *dy = R.dy;
        if (r) *r = R.r;
    } else if (st == XC_NO_PEAK) {
        // CorrPatches returns 0.0 and leaves dx, dy untouched (Correlation.cpp:1074-1087)
        if (r) *r = 0.0;
    }
    if (res) *res = R;
    return st;
}

extern "C" int xc_corr_images_f64(xc_ctx *ctx, int mode, const xc_point *ip1, const double *iv1, int n1,
                                  const xc_point *ip2, const double *iv2, int n2, const xc_rgn_rule *rgn,
                                  const xc_cnt_rule *cnt, double mincor, double nbmaxht, int Ox, int Oy, int Rx,
                                  int Ry, double *dx, double *dy, double *r, xc_result *res)
{
    if (mode != XC_MODE_R && mode != XC_MODE_S && mode != XC_MODE_F)
        return set_err(ctx, XC_E_ARG, "mode must be XC_MODE_R, XC_MODE_S or XC_MODE_F");
    xc_result R;
    memset(&R, 0, sizeof R);
    int st = run_plain(ctx, mode == XC_MODE_S ? PM_S : PM_R, ip1, iv1, n1, ip2, iv2, n2, Ox, Oy, 0, Rx, Ry, mincor,
                       rgn, cnt, &R, mode == XC_MODE_F, nbmaxht, nullptr);
    if (st == XC_OK) {
        if (dx) *dx = R.dx;
        if (dy) *dy = R.dy;
        if (r) *r = R.r;
    } else if (st == XC_NO_PEAK || st == XC_TOO_SMALL) {
        // CorrImages* set dx = dy = 0 and return 0 (Correlation.cpp:3098-3119, 3175-3194, 3254-3273)
        if (dx) *dx = 0.0;
        if (dy) *dy = 0.0;
        if (r) *r = 0.0;
    }
    if (res) *res = R;
    return st;
}
