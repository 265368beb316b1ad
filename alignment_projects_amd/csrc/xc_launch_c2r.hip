// xc_launch_c2r.hip -- row c2r + evaluation launches over FP tables
// (one file per kernel family so that they compile in parallel).
#include "xc_launch_c2r.hpp"

namespace xc {

template <typename T, int NH> static int c2r_n(xc_ctx *ctx, int nb, const PairDev *pd, const int2 *w, bool cnt, int tabk)
{
    if constexpr (NH > max_nh<T>()) {
        return set_err(ctx, XC_E_SHAPE, "row transform %d too long", 2 * NH);
    } else if constexpr (sizeof(T) == 8) {
        if (tabk != TK_F64) return set_err(ctx, XC_E_ARG, "FP32 or exact tables with FP64 transforms");
        return c2r_go<T, NH, false, double>(ctx, nb, pd, w);
    } else {
        if (tabk == TK_F32) return c2r_cnt<T, NH, float>(ctx, nb, pd, w, cnt);
        return c2r_cnt<T, NH, double>(ctx, nb, pd, w, cnt);
    }
}

template <typename T>
int launch_row_c2r(xc_ctx *ctx, int NH, int nb, const PairDev *pd, const int2 *w, bool cnt, int tabk)
{
    if (nb <= 0) return XC_OK;
    if (tabk == TK_X) {
        if (sizeof(T) == 8) return set_err(ctx, XC_E_ARG, "FP32 or exact tables with FP64 transforms");
        return launch_row_c2r_x(ctx, NH, nb, pd, w, cnt);
    }
    switch (NH) {
#define F(n) \
    case n: return c2r_n<T, n>(ctx, nb, pd, w, cnt, tabk);
        XC_POW2_CASES(F)
#undef F
    }
    return set_err(ctx, XC_E_SHAPE, "unsupported row length %d", 2 * NH);
}

template int launch_row_c2r<double>(xc_ctx *, int, int, const PairDev *, const int2 *, bool, int);
template int launch_row_c2r<float>(xc_ctx *, int, int, const PairDev *, const int2 *, bool, int);

}  // namespace xc
