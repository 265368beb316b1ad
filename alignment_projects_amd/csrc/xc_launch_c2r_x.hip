// xc_launch_c2r_x.hip -- row c2r + evaluation launches over exact integer
// tables (FP32 transforms).
#include "xc_launch_c2r.hpp"

namespace xc {

int launch_row_c2r_x(xc_ctx *ctx, int NH, int nb, const PairDev *pd, const int2 *w, bool cnt)
{
    if (nb <= 0) return XC_OK;
    switch (NH) {
#define F(n) \
    case n: return c2r_cnt<float, n, TabX>(ctx, nb, pd, w, cnt);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported row length %d", 2 * NH);
}

}  // namespace xc
