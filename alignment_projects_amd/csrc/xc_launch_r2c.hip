// xc_launch_r2c.hip -- row r2c launches, workgroup shapes and finalize
// (one file per kernel family so that they compile in parallel).
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <typename T> int row_group(int NH)
{
    switch (NH) {
#define F(n) \
    case n: return RowCfg<n>::G;
        XC_POW2_CASES(F)
#undef F
    }
    return 1;
}

template <typename T> int r2c_rows(int NH)
{
    switch (NH) {
#define F(n) \
    case n: return R2cCfg<T, n>::ROWS;
        XC_POW2_CASES(F)
#undef F
    }
    return 1;
}

template <typename T> int col_group(int) { return COL_STEP; }

template <typename T, int NH, int RPT> static int r2c_go(xc_ctx *ctx, int nb, const PairDev *pd, const int4 *w)
{
    const size_t smem = sizeof(cpx<T>) * R2cCfg<T, NH>::ROWS * lds_len(NH);
    constexpr int threads = R2cCfg<T, NH>::threads(RPT);
    hipLaunchKernelGGL((k_row_r2c<T, NH, RPT>), dim3(nb), dim3(threads), smem, ctx->stream, pd, w, twiddles<T>(ctx));
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

// rpt: rows per thread (0: the default shape; else the default or R2cCfg's ALT)
template <typename T, int NH> static int r2c_n(xc_ctx *ctx, int nb, const PairDev *pd, const int4 *w, int rpt)
{
    if constexpr (NH > max_nh<T>()) {
        return set_err(ctx, XC_E_SHAPE, "row transform %d too long", 2 * NH);
    } else {
        constexpr int DEF = R2cCfg<T, NH>::RPT, ALT = R2cCfg<T, NH>::ALT;
        if (rpt == 0 || rpt == DEF) return r2c_go<T, NH, DEF>(ctx, nb, pd, w);
        if constexpr (ALT != 0) {
            if (rpt == ALT) return r2c_go<T, NH, ALT>(ctx, nb, pd, w);
        }
        return set_err(ctx, XC_E_ARG, "no row r2c with %d rows per thread for %d-point rows", rpt, 2 * NH);
    }
}

template <typename T> int r2c_alt_rpt(int NH)
{
    switch (NH) {
#define F(n) \
    case n: return R2cCfg<T, n>::ALT;
        XC_POW2_CASES(F)
#undef F
    }
    return 0;
}

template <typename T> int r2c_default_rpt(int NH)
{
    switch (NH) {
#define F(n) \
    case n: return R2cCfg<T, n>::RPT;
        XC_POW2_CASES(F)
#undef F
    }
    return 1;
}

template <typename T> int launch_row_r2c(xc_ctx *ctx, int NH, int nb, const PairDev *pd, const int4 *w, int rpt)
{
    if (nb <= 0) return XC_OK;
    switch (NH) {
#define F(n) \
    case n: return r2c_n<T, n>(ctx, nb, pd, w, rpt);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported row length %d", 2 * NH);
}

template <typename T> int launch_finalize(xc_ctx *ctx, int npairs, const PairDev *pd, int tabk)
{
    if (npairs <= 0) return XC_OK;
    if constexpr (sizeof(T) == 8) {
        if (tabk != TK_F64) return set_err(ctx, XC_E_ARG, "FP32 or exact tables with FP64 transforms");
        hipLaunchKernelGGL((k_finalize<T, double>), dim3(npairs), dim3(256), 0, ctx->stream, pd, twiddles<T>(ctx));
    } else if (tabk == TK_F32) {
        hipLaunchKernelGGL((k_finalize<T, float>), dim3(npairs), dim3(256), 0, ctx->stream, pd, twiddles<T>(ctx));
    } else if (tabk == TK_X) {
        hipLaunchKernelGGL((k_finalize<T, TabX>), dim3(npairs), dim3(256), 0, ctx->stream, pd, twiddles<T>(ctx));
    } else {
        hipLaunchKernelGGL((k_finalize<T, double>), dim3(npairs), dim3(256), 0, ctx->stream, pd, twiddles<T>(ctx));
    }
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

template int launch_row_r2c<double>(xc_ctx *, int, int, const PairDev *, const int4 *, int);
template int launch_row_r2c<float>(xc_ctx *, int, int, const PairDev *, const int4 *, int);
template int r2c_alt_rpt<float>(int);
template int r2c_alt_rpt<double>(int);
template int r2c_default_rpt<float>(int);
template int r2c_default_rpt<double>(int);
template int launch_finalize<double>(xc_ctx *, int, const PairDev *, int);
template int launch_finalize<float>(xc_ctx *, int, const PairDev *, int);
template int row_group<double>(int);
template int row_group<float>(int);
template int r2c_rows<double>(int);
template int r2c_rows<float>(int);
template int col_group<double>(int);
template int col_group<float>(int);

}  // namespace xc
