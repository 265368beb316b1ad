// xc_noalloc.hip -- xc_corr_patches_noalloc_f64: the plain FP64 CorrPatches
// correlation for callers that cannot allocate (code under graph capture,
// serving loops, harnesses that forbid hipMalloc).  Its device workspace is a
// fixed file-scope array instead of memory the library allocates.
//
// In host code `scratch` names a host-side stand-in, not a device address: its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current, at context creation) and only that address is handed to the
// kernels and copies.  Each GPU holds its own copy of the array until the
// process ends.
#include <string.h>

#include <mutex>

#include "xc_host.hpp"

using namespace xc;

namespace {

constexpr int MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[XC_NOALLOC_SCRATCH_BYTES];

std::mutex g_init_mu;
void *g_scratch[MAX_DEVICES];            // device address of `scratch` per device (null: not looked up yet)
std::mutex g_use_mu[MAX_DEVICES];        // calls on one device share its array: one at a time

}  // namespace

namespace xc {

// Called by xc_ctx_create with ctx->device current.
int noalloc_init(xc_ctx *ctx)
{
    const int d = ctx->device;
    if (d < 0 || d >= MAX_DEVICES) return set_err(ctx, XC_E_ARG, "device %d beyond the fixed-scratch table", d);
    std::lock_guard<std::mutex> lock(g_init_mu);
    if (g_scratch[d]) return XC_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return set_err(ctx, XC_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    }
    g_scratch[d] = p;
    return XC_OK;
}

}  // namespace xc

extern "C" int xc_corr_patches_noalloc_f64(xc_ctx *ctx, const xc_point *ip1, const double *iv1, int n1,
                                           const xc_point *ip2, const double *iv2, int n2, int Ox, int Oy,
                                           int radius, const xc_rgn_rule *rgn, const xc_cnt_rule *cnt, double *dx,
                                           double *dy, double *r, xc_result *res)
{
    if (!ctx) return XC_E_ARG;
    const int d = ctx->device;
    if (d < 0 || d >= MAX_DEVICES) return XC_E_ARG;
    void *p;
    {
        std::lock_guard<std::mutex> lock(g_init_mu);
        p = g_scratch[d];
    }
    if (!p) return set_err(ctx, XC_E_DEVICE, "no fixed scratch on device %d", d);
    FixedWs ws;
    ws.dev = p;
    ws.bytes = XC_NOALLOC_SCRATCH_BYTES;
    xc_result R;
    memset(&R, 0, sizeof R);
    int st;
    {
        // the call synchronises its stream before it returns, so the array is free again on unlock
        std::lock_guard<std::mutex> lock(g_use_mu[d]);
        st = run_plain(ctx, PM_NCC, ip1, iv1, n1, ip2, iv2, n2, Ox, Oy, radius, 0, 0, 0.0, rgn, cnt, &R, false, 0.0,
                       &ws);
    }
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
