// xc_launch_col.hip -- fused column-pass launches
// (one file per kernel family so that they compile in parallel).
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <typename T, int NY, int C>
static int col_nc(xc_ctx *ctx, bool whiten, int xcd, int nitems, const PairDev *pd, const int2 *w)
{
    if constexpr (NY > max_ny<T>() || !ColCfg<T, NY>::fits(C)) {
        return set_err(ctx, XC_E_SHAPE, "no column pass of length %d with %d columns per workgroup", NY, C);
    } else {
        const size_t smem = (size_t)ColCfg<T, NY>::lds_bytes(C);
        if (smem > ctx->lds_max) return set_err(ctx, XC_E_SHAPE, "column pass needs %zu bytes of LDS", smem);
        const dim3 grid(nitems * (COL_STEP / C)), block(Plan<NY>::TPF * C);
        if (whiten) {
            if (smem > 65536)
                (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_col_fused<T, NY, true, C>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
            hipLaunchKernelGGL((k_col_fused<T, NY, true, C>), grid, block, smem, ctx->stream, pd, w, twiddles<T>(ctx),
                               xcd);
        } else {
            if (smem > 65536)
                (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_col_fused<T, NY, false, C>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
            hipLaunchKernelGGL((k_col_fused<T, NY, false, C>), grid, block, smem, ctx->stream, pd, w, twiddles<T>(ctx),
                               xcd);
        }
        XC_HIP(ctx, hipGetLastError());
        return XC_OK;
    }
}

template <typename T, int NY>
static int col_n(xc_ctx *ctx, bool whiten, int cwg, int xcd, int nitems, const PairDev *pd, const int2 *w)
{
    if (cwg <= 0) cwg = ColCfg<T, NY>::C;
    if constexpr (sizeof(T) == 8) {
        // FP64 (the plain path): the default shape only
        if (cwg != ColCfg<T, NY>::C) return set_err(ctx, XC_E_SHAPE, "FP64 column pass has one shape");
        return col_nc<T, NY, ColCfg<T, NY>::C>(ctx, whiten, xcd, nitems, pd, w);
    } else {
        switch (cwg) {
        case 1: return col_nc<T, NY, 1>(ctx, whiten, xcd, nitems, pd, w);
        case 2: return col_nc<T, NY, 2>(ctx, whiten, xcd, nitems, pd, w);
        case 4: return col_nc<T, NY, 4>(ctx, whiten, xcd, nitems, pd, w);
        case 8: return col_nc<T, NY, 8>(ctx, whiten, xcd, nitems, pd, w);
        case 16: return col_nc<T, NY, 16>(ctx, whiten, xcd, nitems, pd, w);
        }
        return set_err(ctx, XC_E_SHAPE, "bad columns per workgroup %d", cwg);
    }
}

template <typename T>
int launch_col(xc_ctx *ctx, int NY, bool whiten, int cwg, int xcd, int nitems, const PairDev *pd, const int2 *w)
{
    if (nitems <= 0) return XC_OK;
    switch (NY) {
#define F(n) \
    case n: return col_n<T, n>(ctx, whiten, cwg, xcd, nitems, pd, w);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported column length %d", NY);
}

bool col_shape_ok(const xc_ctx *ctx, int NY, int cwg)
{
    const int TPF = NY >= 8 ? NY / 8 : 1;
    const long long lds = (long long)cwg * ColCfg<float, 8>::AB * lds_len(NY) * (long long)sizeof(cpx<float>);
    return cwg >= 1 && COL_STEP % cwg == 0 && TPF * cwg <= 1024 && lds <= (long long)ctx->lds_max && lds <= 163840;
}

template int launch_col<double>(xc_ctx *, int, bool, int, int, int, const PairDev *, const int2 *);
template int launch_col<float>(xc_ctx *, int, bool, int, int, int, const PairDev *, const int2 *);

}  // namespace xc
