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

template <typename T, int NH> static int r2c_n(xc_ctx *ctx, int nb, const PairDev *pd, const int4 *w)
{
    if constexpr (NH > max_nh<T>()) {
        return set_err(ctx, XC_E_SHAPE, "row transform %d too long", 2 * NH);
    } else {
        const size_t smem = sizeof(cpx<T>) * R2cCfg<T, NH>::ROWS * lds_len(NH);
        hipLaunchKernelGGL((k_row_r2c<T, NH>), dim3(nb), dim3(RowCfg<NH>::THREADS), smem, ctx->stream, pd, w,
                           twiddles<T>(ctx));
        XC_HIP(ctx, hipGetLastError());
        return XC_OK;
    }
}

template <typename T> int launch_row_r2c(xc_ctx *ctx, int NH, int nb, const PairDev *pd, const int4 *w)
{
    if (nb <= 0) return XC_OK;
    switch (NH) {
#define F(n) \
    case n: return r2c_n<T, n>(ctx, nb, pd, w);
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

template int launch_row_r2c<double>(xc_ctx *, int, int, const PairDev *, const int4 *);
template int launch_row_r2c<float>(xc_ctx *, int, int, const PairDev *, const int4 *);
template int launch_finalize<double>(xc_ctx *, int, const PairDev *, int);
template int launch_finalize<float>(xc_ctx *, int, const PairDev *, int);
template int row_group<double>(int);
template int row_group<float>(int);
template int r2c_rows<double>(int);
template int r2c_rows<float>(int);
template int col_group<double>(int);
template int col_group<float>(int);

}  // namespace xc
