// xc_launch_c2r.hpp -- row c2r + evaluation launches, shared by the
// translation units that instantiate them per table form.
#pragma once

#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <typename T, int NH, bool CNT, typename TB>
int c2r_go(xc_ctx *ctx, int nb, const PairDev *pd, const int2 *w)
{
    const size_t smem = c2r_smem<T, NH>();
    hipLaunchKernelGGL((k_row_c2r_eval<T, NH, CNT, TB>), dim3(nb), dim3(RowCfg<NH>::THREADS), smem, ctx->stream, pd,
                       w, twiddles<T>(ctx));
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

// The count-table scan holds more registers: a separate variant keeps the others' occupancy.
template <typename T, int NH, typename TB>
int c2r_cnt(xc_ctx *ctx, int nb, const PairDev *pd, const int2 *w, bool cnt)
{
    return cnt ? c2r_go<T, NH, true, TB>(ctx, nb, pd, w) : c2r_go<T, NH, false, TB>(ctx, nb, pd, w);
}

}  // namespace xc
