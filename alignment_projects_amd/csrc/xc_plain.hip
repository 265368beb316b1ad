// xc_plain.hip -- plain FP64 form: one pair per call, on the device, with
// the reference's CorrPatches / CorrImagesR / CorrImagesS semantics.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "xc_host.hpp"
#include "xc_kernels.hpp"

using namespace xc;

namespace {

struct PlainBufs {
    xc_point *pts;
    double *vals;
    double *imgA, *imgB;
    double *SA, *QA, *SB, *QB;
    int32_t *ZA, *ZB;
    cpx<double> *specA, *specB, *lag;
    RowRec *rows;
    PairDev *pair;
    int4 *w_r2c;
    int2 *w_col, *w_c2r;
    xc_result *res;
    uint8_t *table;
};

// Run one FP64 correlation of point lists (A = ip1/iv1, B = ip2/iv2).
int run_plain(xc_ctx *ctx, int mode, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
              const double *iv2, int n2, int Ox, int Oy, int radius, int Rx, int Ry, double mincor,
              const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, xc_result *out)
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
    const int G = row_group<double>(NH), C = col_group<double>(NY);
    const int nr2c = (h1 + G - 1) / G + (h2 + G - 1) / G;
    const int ncol = (H + C - 1) / C;
    const int nc2r = (P.hR + G - 1) / G;
    Layout lay;
    size_t i_pts = lay.add(sizeof(xc_point) * n), i_vals = lay.add(sizeof(double) * n);
    size_t i_iA = lay.add(sizeof(double) * w1 * h1), i_iB = lay.add(sizeof(double) * w2 * h2);
    size_t i_SA = lay.add(sizeof(double) * w1 * h1), i_QA = lay.add(sizeof(double) * w1 * h1);
    size_t i_SB = lay.add(sizeof(double) * w2 * h2), i_QB = lay.add(sizeof(double) * w2 * h2);
    size_t i_ZA = lay.add(sizeof(int32_t) * w1 * h1), i_ZB = lay.add(sizeof(int32_t) * w2 * h2);
    size_t i_spA = lay.add(sizeof(cpx<double>) * h1 * H), i_spB = lay.add(sizeof(cpx<double>) * h2 * H);
    size_t i_lag = lay.add(sizeof(cpx<double>) * (size_t)P.hR * H);
    size_t i_rows = lay.add(sizeof(RowRec) * P.hR);
    // staged (uploaded) part: pair, work lists, table
    size_t i_pair = lay.add(sizeof(PairDev));
    size_t i_wr = lay.add(sizeof(int4) * nr2c), i_wc = lay.add(sizeof(int2) * ncol), i_wc2 = lay.add(sizeof(int2) * nc2r);
    size_t i_tbl = lay.add(tbl ? tbl : 1);
    size_t i_res = lay.add(sizeof(xc_result));
    int st = ensure_dev(ctx, ctx->arena, lay.total);
    if (st) return st;
    char *base = (char *)ctx->arena.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };

    P.imgA = at(i_iA);
    P.imgB = at(i_iB);
    P.SA = (double *)at(i_SA);
    P.QA = (double *)at(i_QA);
    P.SB = (double *)at(i_SB);
    P.QB = (double *)at(i_QB);
    P.ZA = (int32_t *)at(i_ZA);
    P.ZB = (int32_t *)at(i_ZB);
    P.specA = at(i_spA);
    P.specB = at(i_spB);
    P.lag = at(i_lag);
    P.rows = (RowRec *)at(i_rows);
    P.out = at(i_res);
    P.rule = make_rule(rgn, cnt, (const uint8_t *)at(i_tbl));
    P.need_cnt = P.rule.cnt_enabled;

    // ---- host staging: points, values, pair, work lists, table ----
    const size_t stage_lo = lay.off[i_pts];
    const size_t stage_hi = lay.off[i_vals] + sizeof(double) * n;
    const size_t stage2_lo = lay.off[i_pair];
    const size_t stage2_hi = lay.off[i_tbl] + (tbl ? tbl : 1);
    st = ensure_pinned(ctx, (stage_hi - stage_lo) + (stage2_hi - stage2_lo) + sizeof(xc_result));
    if (st) return st;
    char *hs = (char *)ctx->pinned;
    char *hs2 = hs + (stage_hi - stage_lo);
    xc_result *hres = (xc_result *)(hs2 + (stage2_hi - stage2_lo));
    memcpy(hs + (lay.off[i_pts] - stage_lo), ip1, sizeof(xc_point) * n1);
    memcpy(hs + (lay.off[i_pts] - stage_lo) + sizeof(xc_point) * n1, ip2, sizeof(xc_point) * n2);
    memcpy(hs + (lay.off[i_vals] - stage_lo), iv1, sizeof(double) * n1);
    memcpy(hs + (lay.off[i_vals] - stage_lo) + sizeof(double) * n1, iv2, sizeof(double) * n2);
    memcpy(hs2 + (lay.off[i_pair] - stage2_lo), &P, sizeof P);
    int4 *wr = (int4 *)(hs2 + (lay.off[i_wr] - stage2_lo));
    int k = 0;
    for (int y = 0; y < h1; y += G) wr[k++] = make_int4(0, 0, y, 0);
    for (int y = 0; y < h2; y += G) wr[k++] = make_int4(0, 1, y, 0);
    int2 *wc = (int2 *)(hs2 + (lay.off[i_wc] - stage2_lo));
    for (int i = 0; i < ncol; ++i) wc[i] = make_int2(0, i * C);
    int2 *wc2 = (int2 *)(hs2 + (lay.off[i_wc2] - stage2_lo));
    for (int i = 0; i < nc2r; ++i) wc2[i] = make_int2(0, i * G);
    if (tbl) memcpy(hs2 + (lay.off[i_tbl] - stage2_lo), rgn->table, tbl);

    hipStream_t s = ctx->stream;
    XC_HIP(ctx, hipMemcpyAsync(base + stage_lo, hs, stage_hi - stage_lo, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(base + stage2_lo, hs2, stage2_hi - stage2_lo, hipMemcpyHostToDevice, s));

    // ---- rasters: zero then splat (ImageFromValuesAndPoints) ----
    {
        const long long nA = (long long)w1 * h1, nB = (long long)w2 * h2;
        const int fb = (int)std::min<long long>(1024, (nA + nB + 255) / 256);
        hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, s, (double *)P.imgA, nA, 0.0);
        hipLaunchKernelGGL(k_fill<double>, dim3(fb), dim3(256), 0, s, (double *)P.imgB, nB, 0.0);
        xc_point *dp = (xc_point *)at(i_pts);
        double *dv = (double *)at(i_vals);
        hipLaunchKernelGGL(k_splat_points, dim3((n1 + 255) / 256), dim3(256), 0, s, dp, dv, n1, (double)B1.L,
                           (double)B1.B, (double *)P.imgA, w1, h1);
        hipLaunchKernelGGL(k_splat_points, dim3((n2 + 255) / 256), dim3(256), 0, s, dp + n1, dv + n1, n2,
                           (double)B2.L, (double)B2.B, (double *)P.imgB, w2, h2);
        XC_HIP(ctx, hipGetLastError());
    }
    // ---- summed-area tables ----
    hipLaunchKernelGGL(k_sat_rows<double>, dim3(h1), dim3(256), 0, s, (const double *)P.imgA, w1, h1,
                       (double *)P.SA, (double *)P.QA, (int32_t *)P.ZA);
    hipLaunchKernelGGL(k_sat_rows<double>, dim3(h2), dim3(256), 0, s, (const double *)P.imgB, w2, h2,
                       (double *)P.SB, (double *)P.QB, (int32_t *)P.ZB);
    hipLaunchKernelGGL(k_sat_cols, dim3((w1 + 63) / 64), dim3(256), 0, s, w1, h1, (double *)P.SA, (double *)P.QA,
                       (int32_t *)P.ZA);
    hipLaunchKernelGGL(k_sat_cols, dim3((w2 + 63) / 64), dim3(256), 0, s, w2, h2, (double *)P.SB, (double *)P.QB,
                       (int32_t *)P.ZB);
    XC_HIP(ctx, hipGetLastError());

    // ---- transforms and evaluation ----
    const PairDev *dP = (const PairDev *)at(i_pair);
    if ((st = launch_row_r2c<double>(ctx, NH, nr2c, dP, (const int4 *)at(i_wr)))) return st;
    if ((st = launch_col<double>(ctx, NY, mode == PM_S, ncol, dP, (const int2 *)at(i_wc)))) return st;
    if ((st = launch_row_c2r<double>(ctx, NH, nc2r, dP, (const int2 *)at(i_wc2)))) return st;
    if ((st = launch_finalize<double>(ctx, 1, dP))) return st;
    XC_HIP(ctx, hipMemcpyAsync(hres, at(i_res), sizeof(xc_result), hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    *out = *hres;
    return out->status;
}

}  // namespace

extern "C" int xc_corr_patches_f64(xc_ctx *ctx, const xc_point *ip1, const double *iv1, int n1, const xc_point *ip2,
                                   const double *iv2, int n2, int Ox, int Oy, int radius, const xc_rgn_rule *rgn,
                                   const xc_cnt_rule *cnt, double *dx, double *dy, double *r, xc_result *res)
{
    xc_result R;
    memset(&R, 0, sizeof R);
    int st = run_plain(ctx, PM_NCC, ip1, iv1, n1, ip2, iv2, n2, Ox, Oy, radius, 0, 0, 0.0, rgn, cnt, &R);
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
    (void)nbmaxht;
    if (mode != XC_MODE_R && mode != XC_MODE_S) return set_err(ctx, XC_E_ARG, "mode must be XC_MODE_R or XC_MODE_S");
    xc_result R;
    memset(&R, 0, sizeof R);
    int st = run_plain(ctx, mode == XC_MODE_R ? PM_R : PM_S, ip1, iv1, n1, ip2, iv2, n2, Ox, Oy, 0, Rx, Ry, mincor,
                       rgn, cnt, &R);
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
