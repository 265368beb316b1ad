// xc_thumbs.hip -- thumbnail preparation on the device: DecimateVector
// (Maths.cpp:1315-1364), Normalize (Maths.cpp:192-212) and
// CThmUtil::MakeThumbs (1_DMesh/CThmUtil.cpp:470-523).
//
// DecimateVector sums each point's value into its lbin x lbin cell in point
// order and counts the points in an 8-bit counter (a cell of 256 points wraps
// to 0 and is dropped, as in the reference), then lists the cells with a
// non-zero count in cell order with their means.  Point-list form: every
// point emits (cell, value); a stable radix sort by cell keeps point order
// inside a cell, so each run sums exactly as the reference does; a scan over
// the occupied cells places the outputs.  Raster form (a device picture, every
// pixel a point in raster order): each cell sums its rows in order, which is
// the same point order; cells of at most 15 x 15 pixels (no counter wrap).
//
// Before them, CThmUtil::Crop / Crop_NoCR (1_DMesh/CThmUtil.cpp:352-464) cut
// the likely overlap out of the two resident pictures: the box arithmetic and
// the selection of a connected region's points run on the host (the points
// are host data), the values are gathered from the device pictures (one
// 4-byte index up, one 8-byte value down per point); without connected
// regions the box is one strided copy.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "xc_host.hpp"

namespace xc {

constexpr int TH_NT = 256;

#pragma clang fp contract(off)

__global__ void k_dec_emit(const xc_point *__restrict__ p, const double *__restrict__ v, int n, int lbin, int wd,
                           int hd, uint32_t *__restrict__ key, double *__restrict__ val, int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)p[i].x / lbin, y = (int)p[i].y / lbin;
    const bool in = p[i].x >= 0.0 && p[i].y >= 0.0 && x < wd && y < hd;
    key[i] = in ? (uint32_t)x + (uint32_t)wd * (uint32_t)y : (uint32_t)wd * (uint32_t)hd;
    val[i] = v[i];
    if (!in) *bad = 1;
}

// Run heads: the cell's in-order sum and its 8-bit count.
__global__ void k_dec_runs(const uint32_t *__restrict__ key, const double *__restrict__ val, int m, uint32_t nr,
                           double *__restrict__ csum, int *__restrict__ ccnt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t k = key[j];
    if (k >= nr || (j > 0 && key[j - 1] == k)) return;
    double s = 0.0;
    int c = 0;
    for (int t = j; t < m && key[t] == k; ++t) {
        s += val[t];
        ++c;
    }
    csum[k] = s;
    ccnt[k] = c & 255;  // the reference's uint8 counter
}

__global__ void k_dec_flags(const int *__restrict__ ccnt, int nr, int *__restrict__ flag)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nr) flag[c] = ccnt[c] != 0;
}

__global__ void k_dec_out(const double *__restrict__ csum, const int *__restrict__ ccnt, const int *__restrict__ idx,
                          int nr, int wd, xc_point *__restrict__ op, double *__restrict__ ov, int *__restrict__ kout)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nr) return;
    const int cnt = ccnt[c];
    if (c == nr - 1) *kout = idx[c] + (cnt != 0);
    if (!cnt) return;
    const int y = c / wd, x = c - wd * y;
    op[idx[c]] = xc_point{(double)x, (double)y};
    ov[idx[c]] = csum[c] / cnt;
}

// Raster form: cell (cx, cy) = mean of its pixels, rows in order.
template <typename P>
__global__ void k_dec_raster(const P *__restrict__ ras, int w, int h, int stride, int lbin, int wd, int hd,
                             double *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= wd * hd) return;
    const int cy = c / wd, cx = c - wd * cy;
    const int x0 = cx * lbin, y0 = cy * lbin;
    const int x1 = min(w, x0 + lbin), y1 = min(h, y0 + lbin);
    double s = 0.0;
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) s += (double)ras[(long long)y * stride + x];
    out[c] = s / ((x1 - x0) * (y1 - y0));
}

// Sums (v, v^2) of n values, partial per workgroup; then Normalize in place.
__global__ void __launch_bounds__(TH_NT) k_th_stats(const double *__restrict__ v, int n, double *__restrict__ part)
{
    __shared__ double sa[TH_NT], sb[TH_NT];
    double a = 0.0, b = 0.0;
    for (int i = blockIdx.x * TH_NT + threadIdx.x; i < n; i += gridDim.x * TH_NT) {
        a += v[i];
        b += v[i] * v[i];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = TH_NT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sa[threadIdx.x] += sa[threadIdx.x + off];
            sb[threadIdx.x] += sb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

// MeanStd::Stats (Maths.cpp:111-124), n-1 deviation; msd = (avg, sd).
__global__ void k_th_fin(const double *__restrict__ part, int np, int n, double *__restrict__ msd)
{
    double s = 0.0, q = 0.0;
    for (int i = 0; i < np; ++i) {
        s += part[2 * i];
        q += part[2 * i + 1];
    }
    double avg = s, sd = 0.0;
    if (n > 1) {
        avg /= n;
        sd = sqrt(fmax(q - n * avg * avg, 0.0) / (n - 1.0));
    }
    msd[0] = avg;
    msd[1] = sd;
}

__global__ void k_th_norm(double *__restrict__ v, int n, const double *__restrict__ msd)
{
    const double avg = msd[0], sd = msd[1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = (v[i] - avg) / sd;
}

// SubI_ThesePoints' value lookup (CThmUtil.cpp:179): out[i] = v[idx[i]].
__global__ void k_crop_gather(const double *__restrict__ v, const uint32_t *__restrict__ idx, int n,
                              double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v[idx[i]];
}

#pragma clang fp contract(on)

namespace {

constexpr int TH_PARTS = 64;

// Normalize the device array v[n] in place; *sd (host) receives the natural sd.
int normalize_dev(xc_ctx *ctx, hipStream_t s, double *v, int n, double *part, double *msd, double *sd)
{
    if (n <= 0) {
        if (sd) *sd = 0.0;
        return XC_OK;
    }
    k_th_stats<<<TH_PARTS, TH_NT, 0, s>>>(v, n, part);
    k_th_fin<<<1, 1, 0, s>>>(part, TH_PARTS, n, msd);
    k_th_norm<<<std::min((n + TH_NT - 1) / TH_NT, 4096), TH_NT, 0, s>>>(v, n, msd);
    XC_HIP(ctx, hipGetLastError());
    if (sd) {
        double h[2];
        XC_HIP(ctx, hipMemcpyAsync(h, msd, sizeof h, hipMemcpyDeviceToHost, s));
        XC_HIP(ctx, hipStreamSynchronize(s));
        *sd = h[1];
    }
    return XC_OK;
}

// DecimateVector of n host points into the device arrays op / ov; *k = cells kept.
int decimate_points(xc_ctx *ctx, const xc_point *p, const double *v, int n, int w, int h, int lbin, DevBuf &buf,
                    xc_point **dop, double **dov, int *k)
{
    hipStream_t s = ctx->stream;
    const int wd = (w + lbin - 1) / lbin, hd = (h + lbin - 1) / lbin;
    const long long nrl = (long long)wd * hd;
    if (nrl >= 0x7fffffffLL) return set_err(ctx, XC_E_SHAPE, "DecimateVector: %lld cells", nrl);
    const int nr = (int)nrl;
    int bits = 1;
    while (bits < 32 && ((uint32_t)nr >> bits) != 0) ++bits;
    size_t sort_tmp = 0, scan_tmp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const double *)nullptr, (double *)nullptr, n, 0, bits);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_tmp, (const int *)nullptr, (int *)nullptr, nr);
    Layout lay;
    const size_t i_p = lay.add(sizeof(xc_point) * n), i_v = lay.add(sizeof(double) * n);
    const size_t i_k0 = lay.add(sizeof(uint32_t) * n), i_k1 = lay.add(sizeof(uint32_t) * n);
    const size_t i_v0 = lay.add(sizeof(double) * n), i_v1 = lay.add(sizeof(double) * n);
    const size_t i_cs = lay.add(sizeof(double) * nr), i_cc = lay.add(sizeof(int) * nr);
    const size_t i_fl = lay.add(sizeof(int) * nr), i_ix = lay.add(sizeof(int) * nr);
    const size_t i_op = lay.add(sizeof(xc_point) * nr), i_ov = lay.add(sizeof(double) * nr);
    const size_t i_tmp = lay.add(std::max(sort_tmp, scan_tmp) + 16), i_sc = lay.add(sizeof(int) * 4);
    int st = ensure_dev(ctx, buf, lay.total);
    if (st) return st;
    char *base = (char *)buf.p;
    auto at = [&](size_t i) { return base + lay.off[i]; };
    XC_HIP(ctx, hipMemcpyAsync(at(i_p), p, sizeof(xc_point) * n, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemcpyAsync(at(i_v), v, sizeof(double) * n, hipMemcpyHostToDevice, s));
    XC_HIP(ctx, hipMemsetAsync(at(i_cc), 0, sizeof(int) * nr, s));
    XC_HIP(ctx, hipMemsetAsync(at(i_sc), 0, sizeof(int) * 4, s));
    int *sc = (int *)at(i_sc);  // [0] out-of-range flag, [1] cells kept
    const int nb = (n + TH_NT - 1) / TH_NT, cb = (nr + TH_NT - 1) / TH_NT;
    k_dec_emit<<<nb, TH_NT, 0, s>>>((const xc_point *)at(i_p), (const double *)at(i_v), n, lbin, wd, hd,
                                    (uint32_t *)at(i_k0), (double *)at(i_v0), sc);
    XC_HIP(ctx, hipGetLastError());
    size_t tb = std::max(sort_tmp, scan_tmp);
    XC_HIP(ctx, hipcub::DeviceRadixSort::SortPairs((void *)at(i_tmp), tb, (const uint32_t *)at(i_k0),
                                                   (uint32_t *)at(i_k1), (const double *)at(i_v0), (double *)at(i_v1),
                                                   n, 0, bits, s));
    k_dec_runs<<<nb, TH_NT, 0, s>>>((const uint32_t *)at(i_k1), (const double *)at(i_v1), n, (uint32_t)nr,
                                    (double *)at(i_cs), (int *)at(i_cc));
    k_dec_flags<<<cb, TH_NT, 0, s>>>((const int *)at(i_cc), nr, (int *)at(i_fl));
    XC_HIP(ctx, hipGetLastError());
    tb = std::max(sort_tmp, scan_tmp);
    XC_HIP(ctx, hipcub::DeviceScan::ExclusiveSum((void *)at(i_tmp), tb, (const int *)at(i_fl), (int *)at(i_ix), nr, s));
    k_dec_out<<<cb, TH_NT, 0, s>>>((const double *)at(i_cs), (const int *)at(i_cc), (const int *)at(i_ix), nr, wd,
                                   (xc_point *)at(i_op), (double *)at(i_ov), sc + 1);
    XC_HIP(ctx, hipGetLastError());
    int h2[2];
    XC_HIP(ctx, hipMemcpyAsync(h2, sc, sizeof h2, hipMemcpyDeviceToHost, s));
    XC_HIP(ctx, hipStreamSynchronize(s));
    if (h2[0]) return set_err(ctx, XC_E_ARG, "DecimateVector: a point lies outside the %d x %d box", w, h);
    *dop = (xc_point *)at(i_op);
    *dov = (double *)at(i_ov);
    *k = h2[1];
    return XC_OK;
}

// GetOlapBoxes (1_DMesh/CThmUtil.cpp:352-373): a's origin in b's system scaled by xyconf / scl and
// truncated, then BoxesFromShifts (Geometry.cpp:150-176) of two ws x hs pictures.  Returns 1 when Crop uses
// the whole pictures instead (xyconf 0, or a box below max(olap1d, 8): CThmUtil.cpp:393-411), -1 for an
// offset that is not a finite number below 2^30.
int crop_boxes(IBox &Ba, IBox &Bb, int ws, int hs, double tx, double ty, double xyconf, double scl, int olap1d)
{
    if (xyconf == 0.0) return 1;
    const double fx = xyconf * tx / scl, fy = xyconf * ty / scl;
    if (!(fabs(fx) < 1073741824.0) || !(fabs(fy) < 1073741824.0)) return -1;
    const int x = (int)fx, y = (int)fy;
    Bb.L = std::max(0, x);
    Bb.R = std::max(Bb.L, std::min(ws - 1, x + ws - 1));
    Bb.B = std::max(0, y);
    Bb.T = std::max(Bb.B, std::min(hs - 1, y + hs - 1));
    Ba = IBox{Bb.L - x, Bb.R - x, Bb.B - y, Bb.T - y};
    const int min1d = std::max(olap1d, 8);
    return (Ba.R - Ba.L + 1 < min1d || Ba.T - Ba.B + 1 < min1d) ? 1 : 0;
}

// One side of Crop: SubI_ThisBox (CThmUtil.cpp:195-235) when box is set, else SubI_ThesePoints
// (:163-189).  The kept points (list order) go to op relative to their bounding box's corner, their values
// are gathered from the device picture dv into ov.  idx / dout: device room for n indexes and values; the
// indexes and values pass through the context's pinned staging (room for both, values at voff).
int crop_side(xc_ctx *ctx, const double *dv, int ws, int hs, const xc_point *p, int n, const IBox *box,
              int64_t olap2d, xc_point *op, double *ov, xc_subimage *S, uint32_t *idx, double *dout, size_t voff)
{
    uint32_t *hidx = (uint32_t *)ctx->pinned;
    double *hval = (double *)((char *)ctx->pinned + voff);
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const double x = p[i].x, y = p[i].y;
        if (box && !(x >= box->L && x <= box->R && y >= box->B && y <= box->T)) continue;
        // (int) truncates toward zero, as the reference's int(P.x): -0.5 is column 0
        const int ix = fabs(x) < 2147483647.0 ? (int)x : -1, iy = fabs(y) < 2147483647.0 ? (int)y : -1;
        if (ix < 0 || ix >= ws || iy < 0 || iy >= hs)
            return set_err(ctx, XC_E_ARG, "Crop: point %d (%g, %g) lies outside the %d x %d picture", i, x, y, ws, hs);
        op[k] = p[i];
        hidx[k++] = (uint32_t)ix + (uint32_t)ws * (uint32_t)iy;
    }
    if (box && (k <= olap2d || k == 0))
        return set_err(ctx, XC_TOO_SMALL, "Crop: small intersection %d (required %lld)", k, (long long)olap2d);
    hipStream_t s = ctx->stream;
    XC_HIP(ctx, hipMemcpyAsync(idx, hidx, sizeof(uint32_t) * k, hipMemcpyHostToDevice, s));
    k_crop_gather<<<(k + TH_NT - 1) / TH_NT, TH_NT, 0, s>>>(dv, idx, k, dout);
    XC_HIP(ctx, hipGetLastError());
    XC_HIP(ctx, hipMemcpyAsync(hval, dout, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    IBox B;  // while the gather runs
    bbox_from_points(B, op, k);
    for (int i = 0; i < k; ++i) {
        op[i].x -= B.L;
        op[i].y -= B.B;
    }
    S->Ox = B.L;
    S->Oy = B.B;
    S->w = B.R - B.L + 1;
    S->h = B.T - B.B + 1;
    S->n = k;
    XC_HIP(ctx, hipStreamSynchronize(s));
    memcpy(ov, hval, sizeof(double) * k);
    return XC_OK;
}

}  // namespace
}  // namespace xc

using namespace xc;

extern "C" int xc_crop_f64(xc_ctx *ctx, const double *av, const double *bv, int ws, int hs, const xc_point *apts,
                           int na, const xc_point *bpts, int nb, double tx, double ty, double xyconf, double scl,
                           int olap1d, int64_t olap2d, xc_point *ap_out, double *av_out, xc_subimage *a,
                           xc_point *bp_out, double *bv_out, xc_subimage *b, int *whole)
{
    if (!ctx) return XC_E_ARG;
    if (!av || !bv || ws < 1 || hs < 1 || (long long)ws * hs >= 0xffffffffLL || !apts || !bpts || na < 1 ||
        nb < 1 || !ap_out || !av_out || !a || !bp_out || !bv_out || !b || !whole)
        return set_err(ctx, XC_E_ARG, "Crop: bad arguments");
    IBox Ba{}, Bb{};
    const int wh = crop_boxes(Ba, Bb, ws, hs, tx, ty, xyconf, scl, olap1d);
    if (wh < 0) return set_err(ctx, XC_E_ARG, "Crop: offset (%g, %g) x %g / %g out of range", tx, ty, xyconf, scl);
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    const int n = std::max(na, nb);
    DevBuf buf;
    const size_t off = ((sizeof(uint32_t) * n + 255) / 256) * 256;
    int st = ensure_dev(ctx, buf, off + sizeof(double) * n);
    if (!st) st = ensure_pinned(ctx, off + sizeof(double) * n);
    if (st) {
        (void)hipFree(buf.p);
        return st;
    }
    uint32_t *idx = (uint32_t *)buf.p;
    double *dout = (double *)((char *)buf.p + off);
    *whole = wh;
    st = crop_side(ctx, av, ws, hs, apts, na, wh ? nullptr : &Ba, olap2d, ap_out, av_out, a, idx, dout, off);
    if (!st)
        st = crop_side(ctx, bv, ws, hs, bpts, nb, wh ? nullptr : &Bb, olap2d, bp_out, bv_out, b, idx, dout, off);
    (void)hipFree(buf.p);
    return st;
}

extern "C" int xc_crop_nocr_f64(xc_ctx *ctx, const double *av, const double *bv, int ws, int hs, double tx,
                                double ty, double xyconf, double scl, int olap1d, double *av_out, double *bv_out,
                                xc_point *p_out, xc_subimage *a, xc_subimage *b, int *whole)
{
    if (!ctx) return XC_E_ARG;
    if (!av || !bv || ws < 1 || hs < 1 || (long long)ws * hs >= 0x7fffffffLL || !av_out || !bv_out || !a || !b ||
        !whole)
        return set_err(ctx, XC_E_ARG, "Crop_NoCR: bad arguments");
    IBox Ba{}, Bb{};
    const int wh = crop_boxes(Ba, Bb, ws, hs, tx, ty, xyconf, scl, olap1d);
    if (wh < 0) return set_err(ctx, XC_E_ARG, "Crop_NoCR: offset (%g, %g) x %g / %g out of range", tx, ty, xyconf, scl);
    if (wh) Ba = Bb = IBox{0, ws - 1, 0, hs - 1};
    const int ow = Ba.R - Ba.L + 1, oh = Ba.T - Ba.B + 1;
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // the two boxes packed on the device, one contiguous copy into pinned staging, then into the outputs
    const size_t row = sizeof(double) * ws, wb = sizeof(double) * ow, box = wb * oh;
    DevBuf buf;
    int st = ensure_dev(ctx, buf, 2 * box);
    if (!st) st = ensure_pinned(ctx, 2 * box);
    if (!st && (hipMemcpy2DAsync(buf.p, wb, av + Ba.L + (size_t)ws * Ba.B, row, wb, oh, hipMemcpyDeviceToDevice,
                                 s) != hipSuccess ||
                hipMemcpy2DAsync((char *)buf.p + box, wb, bv + Bb.L + (size_t)ws * Bb.B, row, wb, oh,
                                 hipMemcpyDeviceToDevice, s) != hipSuccess ||
                hipMemcpyAsync(ctx->pinned, buf.p, 2 * box, hipMemcpyDeviceToHost, s) != hipSuccess))
        st = set_err(ctx, XC_E_DEVICE, "Crop_NoCR: copy failed");
    if (!st && p_out)  // MakeZeroBasedPoints (CPoint.cpp:82-95) while the copies run
        for (int y = 0; y < oh; ++y)
            for (int x = 0; x < ow; ++x) p_out[(size_t)ow * y + x] = xc_point{(double)x, (double)y};
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = set_err(ctx, XC_E_DEVICE, "Crop_NoCR: copy failed");
    (void)hipFree(buf.p);
    if (st) return st;
    memcpy(av_out, ctx->pinned, box);
    memcpy(bv_out, (char *)ctx->pinned + box, box);
    a->Ox = Ba.L;
    a->Oy = Ba.B;
    b->Ox = Bb.L;
    b->Oy = Bb.B;
    a->w = b->w = ow;
    a->h = b->h = oh;
    a->n = b->n = (int64_t)ow * oh;
    *whole = wh;
    return XC_OK;
}

extern "C" int xc_decimate_vector_f64(xc_ctx *ctx, xc_point *p, double *v, int n, int w, int h, int lbin, int *n_out)
{
    if (!ctx) return XC_E_ARG;
    if (n < 0 || (n > 0 && (!p || !v)) || w < 1 || h < 1 || lbin < 1)
        return set_err(ctx, XC_E_ARG, "DecimateVector: bad arguments");
    if (n == 0) {
        if (n_out) *n_out = 0;
        return XC_OK;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    DevBuf buf;
    xc_point *dop;
    double *dov;
    int k = 0;
    int st = decimate_points(ctx, p, v, n, w, h, lbin, buf, &dop, &dov, &k);
    if (!st && k > 0) {
        if (hipMemcpy(p, dop, sizeof(xc_point) * k, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(v, dov, sizeof(double) * k, hipMemcpyDeviceToHost) != hipSuccess)
            st = set_err(ctx, XC_E_DEVICE, "DecimateVector: download failed");
    }
    (void)hipFree(buf.p);
    if (!st && n_out) *n_out = k;
    return st;
}

extern "C" int xc_make_thumbs_f64(xc_ctx *ctx, xc_point *ap, double *av, int na, int aw, int ah, xc_point *bp,
                                  double *bv, int nb, int bw, int bh, int decfactor, int *olap1D, int64_t *req_area,
                                  int *na_out, int *nb_out)
{
    if (!ctx || !olap1D || !req_area || !na_out || !nb_out) return XC_E_ARG;
    if (na < 1 || nb < 1 || !ap || !av || !bp || !bv || decfactor < 1)
        return set_err(ctx, XC_E_ARG, "MakeThumbs: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    xc_point *P[2] = {ap, bp};
    double *V[2] = {av, bv};
    int k[2] = {na, nb};
    DevBuf buf[2], nbuf;
    double *dv[2] = {nullptr, nullptr};
    int st = XC_OK;
    auto finish = [&](int code) -> int {
        (void)hipStreamSynchronize(s);
        for (DevBuf &b : buf) (void)hipFree(b.p);
        (void)hipFree(nbuf.p);
        return code;
    };
    // values of side `side` on the device back into the caller's list
    auto down = [&](int side) -> int {
        if (k[side] > 0 && (hipMemcpyAsync(V[side], dv[side], sizeof(double) * k[side], hipMemcpyDeviceToHost, s) !=
                                hipSuccess ||
                            hipStreamSynchronize(s) != hipSuccess))
            return set_err(ctx, XC_E_DEVICE, "MakeThumbs: download failed");
        return XC_OK;
    };
    // both sides decimated first (CThmUtil.cpp:483-494); their point lists are final from here on
    for (int side = 0; side < 2 && !st; ++side) {
        const int n = side ? nb : na, w = side ? bw : aw, h = side ? bh : ah;
        if (decfactor > 1) {
            xc_point *dp;
            st = decimate_points(ctx, P[side], V[side], n, w, h, decfactor, buf[side], &dp, &dv[side], &k[side]);
            if (!st && k[side] > 0 &&
                hipMemcpyAsync(P[side], dp, sizeof(xc_point) * k[side], hipMemcpyDeviceToHost, s) != hipSuccess)
                st = set_err(ctx, XC_E_DEVICE, "MakeThumbs: download failed");
        } else if (!(st = ensure_dev(ctx, buf[side], sizeof(double) * n))) {
            dv[side] = (double *)buf[side].p;
            if (hipMemcpyAsync(dv[side], V[side], sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess)
                st = set_err(ctx, XC_E_DEVICE, "MakeThumbs: upload failed");
        }
    }
    if (st) return finish(st);
    int64_t req = *req_area;
    int olap = *olap1D;
    if (decfactor > 1) {
        req /= (int64_t)decfactor * decfactor;
        olap /= decfactor;
    }
    *req_area = req;
    *olap1D = olap;
    *na_out = k[0];
    *nb_out = k[1];
    // the size test comes before any Normalize: the lists stay decimated, not normalised (:496-503)
    if (k[0] < req) {
        if (decfactor > 1 && ((st = down(0)) || (st = down(1)))) return finish(st);
        return finish(set_err(ctx, XC_TOO_SMALL, "MakeThumbs: small intersection %d (required %lld)", k[0],
                              (long long)req));
    }
    if ((st = ensure_dev(ctx, nbuf, sizeof(double) * (2 * TH_PARTS + 2)))) return finish(st);
    // Normalize A, then B; a stdev of 0 stops there, leaving the other side as it was (:505-519)
    for (int side = 0; side < 2; ++side) {
        double sd = 0.0;
        if ((st = normalize_dev(ctx, s, dv[side], k[side], (double *)nbuf.p, (double *)nbuf.p + 2 * TH_PARTS, &sd)) ||
            (st = down(side)))
            return finish(st);
        if (!(sd != 0.0)) {
            if (side == 0 && decfactor > 1 && (st = down(1))) return finish(st);
            return finish(set_err(ctx, XC_NO_PEAK, "MakeThumbs: image %c intersection region: stdev = 0",
                                  side ? 'B' : 'A'));
        }
    }
    return finish(XC_OK);
}

extern "C" int xc_make_thumb_raster_dev_f64(xc_ctx *ctx, const void *ras, int dtype, int w, int h, int stride,
                                            int lbin, double *out, double *sd)
{
    if (!ctx) return XC_E_ARG;
    if (!ras || !out || w < 1 || h < 1 || lbin < 1 || lbin > 15 || stride < w)
        return set_err(ctx, XC_E_ARG, "thumbnail raster: bad arguments (bins of 1..15)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    XC_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int wd = (w + lbin - 1) / lbin, hd = (h + lbin - 1) / lbin, nc = wd * hd, cb = (nc + TH_NT - 1) / TH_NT;
    switch (dtype) {
    case XC_U8: k_dec_raster<uint8_t><<<cb, TH_NT, 0, s>>>((const uint8_t *)ras, w, h, stride, lbin, wd, hd, out); break;
    case XC_U16:
        k_dec_raster<uint16_t><<<cb, TH_NT, 0, s>>>((const uint16_t *)ras, w, h, stride, lbin, wd, hd, out);
        break;
    case XC_F32: k_dec_raster<float><<<cb, TH_NT, 0, s>>>((const float *)ras, w, h, stride, lbin, wd, hd, out); break;
    case XC_F64: k_dec_raster<double><<<cb, TH_NT, 0, s>>>((const double *)ras, w, h, stride, lbin, wd, hd, out); break;
    default: return set_err(ctx, XC_E_ARG, "thumbnail raster: unknown pixel type %d", dtype);
    }
    XC_HIP(ctx, hipGetLastError());
    DevBuf nbuf;
    int st = ensure_dev(ctx, nbuf, sizeof(double) * (2 * TH_PARTS + 2));
    if (!st) st = normalize_dev(ctx, s, out, nc, (double *)nbuf.p, (double *)nbuf.p + 2 * TH_PARTS, sd);
    if (!st && !sd && hipStreamSynchronize(s) != hipSuccess) st = set_err(ctx, XC_E_DEVICE, "synchronise failed");
    (void)hipFree(nbuf.p);
    return st;
}
