// xc_exec.hip -- enqueue of the prepared stages (image preparation, row
// r2c, fused column pass, row c2r + evaluation, finalize).
#include <algorithm>
#include <map>
#include <tuple>

#include "xc_exec.hpp"
#include "xc_kernels.hpp"

namespace xc {

void make_sat_groups(const std::vector<int> &widths, const std::vector<char> &tabk, const std::vector<char> &u8,
                     std::vector<SatGroup> &groups, std::vector<int> &flat, std::vector<int> &offsets)
{
    groups.clear();
    flat.clear();
    offsets.clear();
    std::map<std::tuple<int, int, int, int>, std::vector<int>> m;
    for (int i = 0; i < (int)widths.size(); ++i) {
        const int k = i < (int)tabk.size() ? tabk[i] : TK_F64;
        const int b = i < (int)u8.size() && u8[i] ? 1 : 0;
        m[std::make_tuple(sat_cpt(widths[i]), sat_threads(widths[i]), k, b)].push_back(i);
    }
    for (auto &kv : m) {
        SatGroup g;
        g.cpt = std::get<0>(kv.first);
        g.threads = std::get<1>(kv.first);
        g.tabk = std::get<2>(kv.first);
        g.stage = std::get<3>(kv.first) != 0;
        g.count = (int)kv.second.size();
        offsets.push_back((int)flat.size());
        for (int i : kv.second) flat.push_back(i);
        groups.push_back(g);
    }
}

TileSizes tile_sizes(int w, int h, int tabk)
{
    TileSizes t;
    const size_t n = (size_t)w * h;
    t.SQ = (tabk == TK_F32 ? sizeof(float2) : tabk == TK_X ? sizeof(unsigned long long) : sizeof(double2)) * n;
    t.Z = sizeof(int32_t) * n;
    t.part = sizeof(double) * TP_PARTS * 3;
    t.msd = sizeof(double) * 3;
    t.band = (sizeof(double) * 2 + sizeof(int)) * SAT_NB * (size_t)w;
    return t;
}

void make_buckets(const std::vector<PairDev> &pairs, int, std::vector<int> &order, std::vector<Bucket> &buckets)
{
    std::map<std::tuple<int, int, int, int>, std::vector<int>> m;
    for (int i = 0; i < (int)pairs.size(); ++i) {
        const PairDev &P = pairs[i];
        m[std::make_tuple(P.Nx / 2, P.Ny, P.mode == PM_S ? 1 : 0, (int)P.tabk)].push_back(i);
    }
    order.clear();
    buckets.clear();
    for (auto &kv : m) {
        Bucket b;
        b.NH = std::get<0>(kv.first);
        b.NY = std::get<1>(kv.first);
        b.whiten = std::get<2>(kv.first) != 0;
        b.tabk = std::get<3>(kv.first);
        b.first = (int)order.size();
        b.count = (int)kv.second.size();
        for (int i : kv.second) {
            order.push_back(i);
            b.cnt = b.cnt || (pairs[i].mode == PM_NCC && pairs[i].rule.cnt_enabled);
        }
        buckets.push_back(b);
    }
}

void pair_stage_bytes(const PairDev &P, int s, int ba, int bb, bool prep_a, bool prep_b, double bytes[ST_COUNT])
{
    const double H = P.Nx / 2 + 1;
    const double n1 = (double)P.w1 * P.h1, n2 = (double)P.w2 * P.h2;
    const double zb = P.need_cnt ? 4.0 : 0.0;
    const double te = P.tabk == TK_F64 ? 16.0 : 8.0;    // one (S, S^2) table entry
    const double sat = (P.mode == PM_S) ? zb : te + zb;  // S mode: count table only
    bytes[ST_PREP] = (prep_a ? ba * n1 + sat * n1 : 0.0) + (prep_b ? bb * n2 + sat * n2 : 0.0);
    bytes[ST_R2C] = ba * n1 + bb * n2 + s * H * (P.h1 + P.h2);
    const double rows = live_rows(P);  // lag rows written by the column pass and read by c2r
    bytes[ST_COL] = s * H * (P.h1 + P.h2) + s * H * rows;
    bytes[ST_C2R] = s * H * rows + sat * (n1 + n2);
    bytes[ST_FIN] = sizeof(RowRec) * rows;
}

// Table sweeps of every group, by table form: FP64 (S, S^2) entries; FP32
// entries and exact integer tables (transforms in FP32 only).
// side_waves > 0: the sweeps run beside the row transforms, each launch
// holding at most side_waves wavefronts per CU.
template <typename T>
static int launch_sats_any(xc_ctx *ctx, hipStream_t s, const TileDev *d_tiles, const std::vector<SatGroup> &groups,
                           int side_waves)
{
    std::vector<SatGroup> g64, g32, gx;
    for (SatGroup g : groups) {
        const int waves = g.threads >= 64 ? g.threads / 64 : 1;
        g.grid = side_waves > 0 && ctx->num_cu > 0 ? std::max(1, ctx->num_cu * side_waves / waves) : 0;
        (g.tabk == TK_F32 ? g32 : g.tabk == TK_X ? gx : g64).push_back(g);
    }
    int st = launch_sats<T, double>(ctx, s, d_tiles, g64);
    if (st) return st;
    if (g32.empty() && gx.empty()) return XC_OK;
    if constexpr (sizeof(T) == 4) {
        if ((st = launch_sats<T, float>(ctx, s, d_tiles, g32))) return st;
        return launch_sats<T, TabX>(ctx, s, d_tiles, gx);
    }
    return set_err(ctx, XC_E_ARG, "FP32 or exact tables with FP64 transforms");
}

template <typename T> static int enqueue_t(xc_ctx *ctx, const Exec &ex, int s0, int s1)
{
    hipStream_t s = ctx->stream;
    int st;
    const bool prep = s0 <= ST_PREP && ST_PREP < s1;
    // Only the c2r evaluation and the finalize read the tables, so when the
    // whole pipeline is enqueued the sweeps (barrier-bound, one row at a time)
    // run on the side stream while the row and column transforms run here.
    const bool fork = prep && ST_C2R < s1 && ctx->side && (!ex.sat_groups.empty() || !ex.dsat_groups.empty());
    if (prep && ex.ntiles > 0) {
        hipLaunchKernelGGL(k_tp_stats, dim3(ex.ntiles * TP_PARTS), dim3(256), 0, s, ex.d_tiles);
        hipLaunchKernelGGL(k_tp_stats_fin, dim3((ex.ntiles + 63) / 64), dim3(64), 0, s, ex.d_tiles, ex.ntiles);
        XC_HIP(ctx, hipGetLastError());
    }
    if (prep && ex.n_rot_work > 0) {
        hipLaunchKernelGGL(k_rot_norm, dim3(ex.n_rnorm_work), dim3(256), 0, s, ex.d_rnorm, ex.d_rnorm_work);
        hipLaunchKernelGGL(k_rot_splat<T>, dim3(ex.n_rot_work), dim3(256), 0, s, ex.d_rot, ex.d_rot_work);
        XC_HIP(ctx, hipGetLastError());
    }
    if (prep) {
        hipStream_t ss = s;
        if (fork) {
            XC_HIP(ctx, hipEventRecord(ctx->fork_ev, s));
            XC_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
            ss = ctx->side;
        }
        const int sw = fork ? ex.sat_waves : 0;
        if (ex.ntiles > 0 && (st = launch_sats_any<T>(ctx, ss, ex.d_tiles, ex.sat_groups, sw))) return st;
        if (!ex.dsat_groups.empty() && (st = launch_sats_any<T>(ctx, ss, ex.d_dtiles, ex.dsat_groups, sw)))
            return st;
        XC_HIP(ctx, hipGetLastError());
        if (fork) XC_HIP(ctx, hipEventRecord(ctx->join_ev, ss));
    }
    // stage-major: a row spectrum shared by pairs of several buckets is
    // produced before any bucket's column pass reads it
    if (s0 <= ST_R2C && ST_R2C < s1)
        for (const Bucket &b : ex.buckets) {
            if constexpr (sizeof(T) == 4) {
                if (b.d_r2c_u8) {
                    if ((st = launch_row_r2c_u8(ctx, b.NH, b.n_r2c, b.d_r2c_u8, b.r2c_rpt, b.r2c_np, b.r2c_gen)))
                        return st;
                    continue;
                }
            }
            if ((st = launch_row_r2c<T>(ctx, b.NH, b.n_r2c, ex.d_pairs, b.d_r2c, b.r2c_rpt))) return st;
        }
    if (s0 <= ST_COL && ST_COL < s1)
        for (const Bucket &b : ex.buckets)
            if ((st = launch_col<T>(ctx, b.NY, b.whiten, b.col_c, b.col_xcd, b.n_col, ex.d_pairs, b.d_col))) return st;
    if (fork) XC_HIP(ctx, hipStreamWaitEvent(s, ctx->join_ev, 0));
    if (s0 <= ST_C2R && ST_C2R < s1)
        for (const Bucket &b : ex.buckets)
            if ((st = launch_row_c2r<T>(ctx, b.NH, b.n_c2r, ex.d_pairs, b.d_c2r, b.cnt, b.tabk))) return st;
    if (s0 <= ST_FIN && ST_FIN < s1)
        for (const Bucket &b : ex.buckets)
            if ((st = launch_finalize<T>(ctx, b.count, ex.d_pairs + b.first, b.tabk))) return st;
    return XC_OK;
}

bool enqueue_forks(const xc_ctx *ctx, const Exec &ex)
{
    return ctx->side && (!ex.sat_groups.empty() || !ex.dsat_groups.empty());
}

int enqueue(xc_ctx *ctx, const Exec &ex, int s0, int s1)
{
    return ex.fp64 ? enqueue_t<double>(ctx, ex, s0, s1) : enqueue_t<float>(ctx, ex, s0, s1);
}

}  // namespace xc
