// xc_launch_col.hip -- fused column-pass launches
// (one file per kernel family so that they compile in parallel).
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <typename T, int NY> static int col_n(xc_ctx *ctx, bool whiten, int nb, const PairDev *pd, const int2 *w)
{
    if constexpr (NY > max_ny<T>()) {
        return set_err(ctx, XC_E_SHAPE, "column transform %d too long", NY);
    } else {
        using CC = ColCfg<T, NY>;
        const size_t smem = sizeof(cpx<T>) * CC::C * lds_len(NY);
        if (whiten)
            hipLaunchKernelGGL((k_col_fused<T, NY, true>), dim3(nb), dim3(CC::THREADS), smem, ctx->stream, pd, w,
                               twiddles<T>(ctx));
        else
            hipLaunchKernelGGL((k_col_fused<T, NY, false>), dim3(nb), dim3(CC::THREADS), smem, ctx->stream, pd, w,
                               twiddles<T>(ctx));
        XC_HIP(ctx, hipGetLastError());
        return XC_OK;
    }
}

template <typename T> int launch_col(xc_ctx *ctx, int NY, bool whiten, int nb, const PairDev *pd, const int2 *w)
{
    if (nb <= 0) return XC_OK;
    switch (NY) {
#define F(n) \
    case n: return col_n<T, n>(ctx, whiten, nb, pd, w);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported column length %d", NY);
}

template int launch_col<double>(xc_ctx *, int, bool, int, const PairDev *, const int2 *);
template int launch_col<float>(xc_ctx *, int, bool, int, const PairDev *, const int2 *);

}  // namespace xc
