// xc_exec.hpp -- the shared executor: a laid-out set of image preparations
// and pair correlations, enqueued stage by stage on the context's stream.
#pragma once

#include <vector>

#include "xc_host.hpp"

namespace xc {

// ST_PREP covers the tiles' statistics and tables and, in sweeps, the
// rotated copies and their tables.
enum Stage { ST_PREP = 0, ST_R2C = 1, ST_COL = 2, ST_C2R = 3, ST_FIN = 4, ST_COUNT = 5 };

struct Bucket {
    int NH = 0, NY = 0;
    bool whiten = false;
    int first = 0, count = 0;  // pair range in the PairDev array
    const int4 *d_r2c = nullptr;
    int n_r2c = 0;
    const int2 *d_col = nullptr;
    int n_col = 0;     // column work items (COL_STEP columns each)
    int col_c = 0;     // columns per workgroup of the column pass (0: default)
    int col_xcd = 1;   // XCD-aware workgroup order in the column pass
    double col_us = 0; // tuned time of the column pass (0: not tuned)
    int r2c_rpt = 0;   // rows per thread of the row r2c (0: default shape)
    const R2cU8 *d_r2c_u8 = nullptr;  // the same n_r2c items for the 8-bit row r2c (null: the general kernel)
    int r2c_np = 0;    // work items per workgroup of the 8-bit row r2c (0: the default)
    bool r2c_gen = false;  // the 8-bit row r2c's general rows (masked, or starting at odd addresses)
    double r2c_us = 0; // tuned time of the row r2c (0: not tuned)
    const int2 *d_c2r = nullptr;
    int n_c2r = 0;
    bool cnt = false;  // some NCC pair of the bucket has a count rule (c2r variant with the count-table scan)
    int tabk = TK_F64;  // TabKind of the bucket's pairs' tables
};

// Device-resident execution description.  All pointers are device pointers
// into memory owned by the caller (a plan or the per-call arena).
// Images of one table-sweep launch (same columns-per-thread).
struct SatGroup {
    int cpt = 0;
    int threads = 0;
    int count = 0;
    int nb = 1;  // row bands per image (the largest of the group's images)
    int tabk = TK_F64;  // TabKind of the group's tables
    bool stage = false;  // some image is u8: the launch carries the LDS block that stages its rows
    int grid = 0;        // workgroups of the launch (0: one per item)
    const int *list = nullptr;  // device: tile indices
};

// Wavefronts per CU that the table sweeps running beside the row transforms
// may hold by default (of 32; 0: one workgroup per item, no limit).
constexpr int SAT_SIDE_WAVES = 16;

// Tile indices grouped by sweep shape (sat_cpt(w), sat_threads(w)), table
// form (tabk[i], a TabKind; missing entries: TK_F64) and whether the image is
// u8 (u8[i]; missing: no): fills `groups` and the
// concatenated index list `flat` (upload it, then point each group's list
// into the device copy at its offset).
void make_sat_groups(const std::vector<int> &widths, const std::vector<char> &tabk, const std::vector<char> &u8,
                     std::vector<SatGroup> &groups,
                     std::vector<int> &flat, std::vector<int> &offsets);

struct Exec {
    int ntiles = 0;
    const TileDev *d_tiles = nullptr;
    std::vector<SatGroup> sat_groups;
    // rotated copies (sweeps): splat, then their tables (no statistics); the
    // same sweep launches build the plan's derived FP64 tables of exact tiles
    const RotDev *d_rot = nullptr;
    const int2 *d_rot_work = nullptr;
    int n_rot_work = 0;
    const RotNorm *d_rnorm = nullptr;  // the copies' sources, normalised before the splat
    const int2 *d_rnorm_work = nullptr;
    int n_rnorm_work = 0;
    const TileDev *d_dtiles = nullptr;
    std::vector<SatGroup> dsat_groups;
    const PairDev *d_pairs = nullptr;
    std::vector<Bucket> buckets;
    bool fp64 = false;
    int sat_waves = SAT_SIDE_WAVES;  // wavefronts per CU of the sweeps beside the transforms (0: no limit)
};

// Table sweeps of `groups` on stream s: values rounded to T, (S, S^2) entries
// stored as TB (double / float pairs, or TabX exact sums; xc_launch_sat_*.hip
// instantiate the four combinations).
template <typename T, typename TB>
int launch_sats(xc_ctx *ctx, hipStream_t s, const TileDev *d_tiles, const std::vector<SatGroup> &groups);

// Enqueue stages [s0, s1) of `ex` on ctx->stream.
int enqueue(xc_ctx *ctx, const Exec &ex, int s0 = 0, int s1 = ST_COUNT);
// Whether a whole-pipeline enqueue of ex runs its table sweeps on the side stream.
bool enqueue_forks(const xc_ctx *ctx, const Exec &ex);

// Host-side helpers that size and fill descriptors.

// Bytes of device scratch/table memory one prepared image needs.
struct TileSizes {
    size_t SQ, Z, part, msd, band;
};
TileSizes tile_sizes(int w, int h, int tabk);  // tabk: the TabKind of the tables

// Group pairs into (NH, NY, whiten, table form) buckets; returns the
// order of pair indices (bucket-major) and the bucket list (first/count filled).
void make_buckets(const std::vector<PairDev> &pairs, int rows_g64, std::vector<int> &order,
                  std::vector<Bucket> &buckets);

// Algorithmic HBM bytes of each stage for one pair (the compulsory traffic
// model: inputs once, each intermediate written once and read once).
void pair_stage_bytes(const PairDev &P, int elem_bytes, int in_bytes_a, int in_bytes_b, bool prep_a, bool prep_b,
                      double bytes[ST_COUNT]);

}  // namespace xc
