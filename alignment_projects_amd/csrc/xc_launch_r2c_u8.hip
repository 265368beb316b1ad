// xc_launch_r2c_u8.hip -- launches of the 8-bit row r2c (unmasked u8 rows,
// FP32 transforms): rows per thread x work items per workgroup x row form
// (unmasked rows at even addresses, or any rows).
#include "xc_host.hpp"
#include "xc_kernels.hpp"

namespace xc {

template <int NH, int RPT, int NP, bool GEN> static int r2c_u8_go(xc_ctx *ctx, int nitems, const R2cU8 *items)
{
    const size_t smem = sizeof(cpx<float>) * R2cCfg<float, NH>::ROWS * lds_len(NH);
    constexpr int threads = R2cCfg<float, NH>::threads(RPT);
    hipLaunchKernelGGL((k_row_r2c_u8<NH, RPT, NP, GEN>), dim3((nitems + NP - 1) / NP), dim3(threads), smem, ctx->stream,
                       items, nitems, twiddles<float>(ctx));
    XC_HIP(ctx, hipGetLastError());
    return XC_OK;
}

template <int NH, int RPT, bool GEN> static int r2c_u8_np(xc_ctx *ctx, int nitems, const R2cU8 *items, int np)
{
    switch (np) {
    case 1: return r2c_u8_go<NH, RPT, 1, GEN>(ctx, nitems, items);
    case 2: return r2c_u8_go<NH, RPT, 2, GEN>(ctx, nitems, items);
    case 4: return r2c_u8_go<NH, RPT, 4, GEN>(ctx, nitems, items);
    }
    return set_err(ctx, XC_E_SHAPE, "no 8-bit row r2c with %d work items per workgroup", np);
}

template <int NH> static int r2c_u8_n(xc_ctx *ctx, int nitems, const R2cU8 *items, int rpt, int np, bool gen)
{
    constexpr int DEF = R2cCfg<float, NH>::RPT, ALT = R2cCfg<float, NH>::ALT;
    if (rpt == 0 || rpt == DEF)
        return gen ? r2c_u8_np<NH, DEF, true>(ctx, nitems, items, np) : r2c_u8_np<NH, DEF, false>(ctx, nitems, items, np);
    if constexpr (ALT != 0) {
        if (rpt == ALT)
            return gen ? r2c_u8_np<NH, ALT, true>(ctx, nitems, items, np)
                       : r2c_u8_np<NH, ALT, false>(ctx, nitems, items, np);
    }
    return set_err(ctx, XC_E_ARG, "no row r2c with %d rows per thread for %d-point rows", rpt, 2 * NH);
}

bool r2c_u8_length(int NH) { return NH >= 256 && NH <= 4096 && (NH & (NH - 1)) == 0; }

int launch_row_r2c_u8(xc_ctx *ctx, int NH, int nitems, const R2cU8 *items, int rpt, int np, bool gen)
{
    if (nitems <= 0) return XC_OK;
    if (np <= 0) np = R2C_U8_NP;
    switch (NH) {
    case 256: return r2c_u8_n<256>(ctx, nitems, items, rpt, np, gen);
    case 512: return r2c_u8_n<512>(ctx, nitems, items, rpt, np, gen);
    case 1024: return r2c_u8_n<1024>(ctx, nitems, items, rpt, np, gen);
    case 2048: return r2c_u8_n<2048>(ctx, nitems, items, rpt, np, gen);
    case 4096: return r2c_u8_n<4096>(ctx, nitems, items, rpt, np, gen);
    }
    return set_err(ctx, XC_E_SHAPE, "no 8-bit row r2c for %d-point rows", 2 * NH);
}

}  // namespace xc
