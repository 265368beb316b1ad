// xc_launch.hip -- context management, workspace and kernel dispatch of
// libxcorr (everything the C-ABI entry points share).
#include <math.h>

#include <algorithm>
#include <vector>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

int set_err(xc_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int ensure_dev(xc_ctx *ctx, DevBuf &b, size_t bytes, bool slack)
{
    if (bytes <= b.bytes) return XC_OK;
    if (b.p) {
        XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        XC_HIP(ctx, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    const size_t want = slack ? bytes + bytes / 4 : bytes;
    if (hipMalloc(&b.p, want) != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return set_err(ctx, XC_E_NOMEM, "hipMalloc(%zu) failed", want);
    }
    b.bytes = want;
    return XC_OK;
}

int ensure_pinned(xc_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->pinned_bytes) return XC_OK;
    if (ctx->pinned) {
        XC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        XC_HIP(ctx, hipHostFree(ctx->pinned));
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    if (hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault) != hipSuccess) {
        ctx->pinned = nullptr;
        (void)hipGetLastError();
        return set_err(ctx, XC_E_NOMEM, "hipHostMalloc(%zu) failed", want);
    }
    ctx->pinned_bytes = want;
    return XC_OK;
}

void bbox_from_points(IBox &B, const xc_point *p, int n)
{
    double xlo, xhi, ylo, yhi;
    xlo = xhi = p[0].x;
    ylo = yhi = p[0].y;
    for (int i = 1; i < n; ++i) {
        double t;
        if ((t = p[i].x) < xlo) xlo = t;
        else if (t > xhi) xhi = t;
        if ((t = p[i].y) < ylo) ylo = t;
        else if (t > yhi) yhi = t;
    }
    B.L = (int)floor(xlo);
    B.R = (int)ceil(xhi);
    B.B = (int)floor(ylo);
    B.T = (int)ceil(yhi);
}

int ceil_pow2(int n)
{
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

RuleDev make_rule(const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, const uint8_t *dev_table)
{
    RuleDev r;
    memset(&r, 0, sizeof r);
    if (rgn && rgn->enabled) {
        r.rgn_enabled = 1;
        r.min_w = rgn->min_w;
        r.min_h = rgn->min_h;
        r.exact_w = rgn->exact_w;
        r.exact_h = rgn->exact_h;
        r.or_w_ge = rgn->or_w_ge;
        r.or_h_gt = rgn->or_h_gt;
        r.min_area = rgn->min_area;
        r.table_w = rgn->table_w;
        r.table_h = rgn->table_h;
        r.table = rgn->table ? dev_table : nullptr;
    }
    if (cnt && cnt->enabled) {
        r.cnt_enabled = 1;
        r.k1 = cnt->k1;
        r.k2 = cnt->k2;
    }
    return r;
}

int setup_pair_geometry(PairDev &P, int w1, int h1, int w2, int h2, int mode)
{
    P.w1 = w1;
    P.h1 = h1;
    P.w2 = w2;
    P.h2 = h2;
    P.Nx_ref = ceil_pow2(w1 + w2 - 1);
    P.Ny_ref = ceil_pow2(h1 + h2 - 1);
    P.Nx = P.Nx_ref < 2 ? 2 : P.Nx_ref;
    P.Ny = P.Ny_ref;
    P.mode = mode;
    P.lag_order = mode == PM_NCC ? LAG_RASTER : LAG_CENTRED;
    P.nnegx = w1 - 1;
    P.nposx = w2;
    P.nnegy = h1 - 1;
    P.nposy = h2;
    P.wR = P.nnegx + P.nposx;
    P.hR = P.nnegy + P.nposy;
    P.ylo = -0x7fffffff;  // every row (narrow_rows may shrink this)
    P.yhi = 0x7fffffff;
    return XC_OK;
}

// The scalar part of the LegalRgn descriptor (the device's rgn_ok without a table).
static bool rgn_scalar_ok(const RuleDev &g, int olw, int olh)
{
    if (olw < g.min_w || olh < g.min_h) return false;
    if (!((long long)olw * olh > g.min_area)) return false;
    if (g.exact_w > 0 && olw != g.exact_w) return false;
    if (g.exact_h > 0 && olh != g.exact_h) return false;
    if ((g.or_w_ge > 0 || g.or_h_gt > 0) && !(olw >= g.or_w_ge || olh > g.or_h_gt)) return false;
    return true;
}

void narrow_rows(PairDev &P)
{
    if (P.mode != PM_NCC) return;  // R/S images and the P2I surface need every row
    const RuleDev &g = P.rule;
    const int maxw = std::min(P.w1, P.w2);  // widest overlap any x lag reaches
    int lo = 0x7fffffff, hi = -0x7fffffff;
    for (int y = -P.nnegy; y < P.nposy; ++y) {
        const long long dy = (long long)y - P.Oy;
        if (dy * dy > (long long)P.radius * P.radius) continue;  // the disc excludes the whole row
        const int B = std::max(0, y);
        const int T = std::max(B, std::min(P.h2 - 1, y + P.h1 - 1));
        const int olh = T - B + 1;
        // every rule term is non-decreasing in olw, so the widest overlap decides (or the one exact width)
        int olw = maxw;
        if (g.rgn_enabled && !g.table) {
            if (g.exact_w > 0) {
                if (g.exact_w > maxw) continue;
                olw = g.exact_w;
            }
            if (!rgn_scalar_ok(g, olw, olh)) continue;
        }
        // present-pixel counts cannot exceed the overlap area
        if (g.cnt_enabled && !((long long)olw * olh > g.k1 && (long long)olw * olh > g.k2)) continue;
        lo = std::min(lo, y);
        hi = std::max(hi, y);
    }
    if (lo > hi) {
        P.ylo = 1;  // no live row: no legal lag anywhere
        P.yhi = 0;
    } else {
        P.ylo = lo;
        P.yhi = hi;
    }
}

void live_row_groups(const PairDev &P, int G, std::vector<int> &starts)
{
    starts.clear();
    for (int y0 = 0; y0 < P.hR; y0 += G) {
        bool any = false;
        for (int y = y0; y < y0 + G && y < P.hR && !any; ++y) any = row_live(P, y);
        if (any) starts.push_back(y0);
    }
}

int live_rows(const PairDev &P)
{
    int n = 0;
    for (int y = 0; y < P.hR; ++y) n += row_live(P, y);
    return n;
}

template <> const cpx<double> *twiddles<double>(xc_ctx *ctx) { return ctx->tw64; }
template <> const cpx<float> *twiddles<float>(xc_ctx *ctx) { return ctx->tw32; }

}  // namespace xc

using namespace xc;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

extern "C" const char *xc_version(void) { return "xcorr 0.1 (gfx950)"; }

extern "C" const char *xc_last_error(const xc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int xc_ctx_create(int device, xc_ctx **out)
{
    if (!out) return XC_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return XC_E_DEVICE;
    }
    if (device < 0 || device >= ndev) return XC_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return XC_E_DEVICE;
    xc_ctx *ctx = new xc_ctx;
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return XC_E_DEVICE;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        xc_ctx_destroy(ctx);
        return XC_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cu = prop.multiProcessorCount;
        ctx->lds_max = prop.sharedMemPerBlock;
    }
    if (const char *tf = getenv("XC_TUNE_FILE")) {
        // "col NH NY whiten c xcd us" and "r2c NH NY kernel rpt np us" lines written by earlier tuning runs
        // (kernel: 0 the general r2c, 1 the 8-bit one, 2 the 8-bit one's general rows)
        ctx->tune_file = tf;
        if (FILE *f = fopen(tf, "r")) {
            char tag[8];
            int nh, ny, wh, c, x;
            double us;
            while (fscanf(f, "%7s %d %d %d %d %d %lf", tag, &nh, &ny, &wh, &c, &x, &us) == 7)
                if (!strcmp(tag, "col") && (c == 0 || col_shape_ok(ctx, ny, c))) {
                    xc_col_tune t;
                    t.c = c;
                    t.xcd = x;
                    t.us = us;
                    ctx->tune[std::make_tuple(nh, ny, wh)] = t;
                } else if (!strcmp(tag, "r2c") && (c == 0 || c == r2c_default_rpt<float>(nh) ||
                                                   c == r2c_alt_rpt<float>(nh)) &&
                           (x == 0 || x == 1 || x == 2 || x == 4)) {
                    xc_col_tune t;
                    t.c = c;
                    t.xcd = x;
                    t.us = us;
                    ctx->tune_r2c[std::make_tuple(nh, ny, wh)] = t;
                }
            fclose(f);
        }
    }
    // e^{-2 pi i m / XC_TW_N}, from long-double exact-angle cos/sin
    std::vector<cpx<double>> t64(XC_TW_N);
    std::vector<cpx<float>> t32(XC_TW_N);
    for (int m = 0; m < XC_TW_N; ++m) {
        long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)XC_TW_N;
        t64[m].x = (double)cosl(a);
        t64[m].y = (double)-sinl(a);
        t32[m].x = (float)cosl(a);
        t32[m].y = (float)-sinl(a);
    }
    if (hipMalloc(&ctx->tw64, sizeof(cpx<double>) * XC_TW_N) != hipSuccess ||
        hipMalloc(&ctx->tw32, sizeof(cpx<float>) * XC_TW_N) != hipSuccess ||
        hipMemcpy(ctx->tw64, t64.data(), sizeof(cpx<double>) * XC_TW_N, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->tw32, t32.data(), sizeof(cpx<float>) * XC_TW_N, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        xc_ctx_destroy(ctx);
        return XC_E_DEVICE;
    }
    // the allocation-free entry point's fixed scratch on this device (looked up once per device)
    if (noalloc_init(ctx) != XC_OK) {
        xc_ctx_destroy(ctx);
        return XC_E_DEVICE;
    }
    *out = ctx;
    return XC_OK;
}

extern "C" int xc_ctx_destroy(xc_ctx *ctx)
{
    if (!ctx) return XC_E_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->arena.p) (void)hipFree(ctx->arena.p);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->tw64) (void)hipFree(ctx->tw64);
    if (ctx->tw32) (void)hipFree(ctx->tw32);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return XC_OK;
}

extern "C" int xc_fft_size(int n1, int n2) { return ceil_pow2(n1 + n2 - 1); }

extern "C" int xc_fft_size_sqr(int w1, int h1, int w2, int h2)
{
    int a = w1 + w2, b = h1 + h2;
    return ceil_pow2((a > b ? a : b) - 1);
}
