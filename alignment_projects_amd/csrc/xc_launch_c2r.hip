// xc_launch_c2r.hip -- row c2r + evaluation launches
// (one file per kernel family so that they compile in parallel).
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <typename T, int NH, bool CNT, typename TB>
static int c2r_go(xc_ctx *ctx, int nb, const PairDev *pd, const int2 *w)
{
    const size_t smem = c2r_smem<T, NH>();
    hipLaunchKernelGGL((k_row_c2r_eval<T, NH, CNT, TB>), dim3(nb), dim3(RowCfg<NH>::THREADS), smem, ctx->stream, pd,
                       w, twiddles<T>(ctx));
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

template <typename T, int NH> static int c2r_n(xc_ctx *ctx, int nb, const PairDev *pd, const int2 *w, bool cnt, bool tab32)
{
    if constexpr (NH > max_nh<T>()) {
        return set_err(ctx, XC_E_SHAPE, "row transform %d too long", 2 * NH);
    } else if constexpr (sizeof(T) == 8) {
        if (tab32) return set_err(ctx, XC_E_ARG, "FP32 tables with FP64 transforms");
        return c2r_go<T, NH, false, double>(ctx, nb, pd, w);
    } else {
        // the count-table scan holds more registers: a separate variant keeps the others' occupancy
        if (tab32) return cnt ? c2r_go<T, NH, true, float>(ctx, nb, pd, w) : c2r_go<T, NH, false, float>(ctx, nb, pd, w);
        return cnt ? c2r_go<T, NH, true, double>(ctx, nb, pd, w) : c2r_go<T, NH, false, double>(ctx, nb, pd, w);
    }
}

template <typename T>
int launch_row_c2r(xc_ctx *ctx, int NH, int nb, const PairDev *pd, const int2 *w, bool cnt, bool tab32)
{
    if (nb <= 0) return XC_OK;
    switch (NH) {
#define F(n) \
    case n: return c2r_n<T, n>(ctx, nb, pd, w, cnt, tab32);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported row length %d", 2 * NH);
}

template int launch_row_c2r<double>(xc_ctx *, int, int, const PairDev *, const int2 *, bool, bool);
template int launch_row_c2r<float>(xc_ctx *, int, int, const PairDev *, const int2 *, bool, bool);

}  // namespace xc
