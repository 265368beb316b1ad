// xc_launch_sat.hpp -- table-sweep launches (statistics tables of the
// image preparation), one translation unit per (value, table) precision so
// that the sweep's many shape instantiations compile in parallel.
#pragma once

#include "xc_exec.hpp"
#include "xc_kernels.hpp"

namespace xc {

// Workgroups of a sweep launch of n items: one per item, or at most cap (> 0).
inline int sat_grid(int n, int cap) { return cap > 0 && cap < n ? cap : n; }

template <typename T, typename TB>
int launch_sats(xc_ctx *ctx, hipStream_t s, const TileDev *d_tiles, const std::vector<SatGroup> &groups)
{
    for (const auto &g : groups) {
        {
            const int key = g.cpt * 10000 + g.threads;
            switch (key) {
#define XC_SAT(c, n)                                                                                                \
    case c * 10000 + n:                                                                                             \
        if (g.nb > 1)                                                                                               \
            hipLaunchKernelGGL((k_tp_colsum<T, c, n>), dim3(g.count * g.nb), dim3(n), 0, s, d_tiles, g.list, g.nb); \
        hipLaunchKernelGGL((k_tp_sat<T, TB, c, n>), dim3(sat_grid(g.count * g.nb, g.grid)), dim3(n),               \
                           g.stage ? SATX_BLK : 0, s, d_tiles, g.list, g.nb, g.stage ? 1 : 0, g.count * g.nb);       \
        break;
                XC_SAT(2, 64) XC_SAT(2, 128) XC_SAT(2, 256) XC_SAT(2, 512) XC_SAT(2, 1024) XC_SAT(4, 1024)
                XC_SAT(8, 1024)
#undef XC_SAT
            default: return set_err(ctx, XC_E_SHAPE, "no table sweep for %d threads x %d columns", g.threads, g.cpt);
            }
        }
    }
    return XC_OK;
}

}  // namespace xc
