// grouped_plan.h -- the host half of a grouped plan: problem and work-item
// descriptors, the tuning-file format and the planner that assigns tiles,
// split-K and queues (hbs_grouped_plan_create's host work, and all of
// hbs_grouped_plan_describe).  Host-only C++: the device side
// (gemm_grouped.hip) reads the descriptors, and the planner also builds
// without HIP for the sanitizer build.
#pragma once

#include <map>
#include <tuple>
#include <vector>

#include "../../include/hbsgemm.h"
#include "hbs_host.h"

namespace hbs {
namespace grouped {

struct Prob {
    const float *A;
    const float *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
};

struct Item {
    int prob;
    int tm, tn;
    int k0, k1;
    int cfg;       // tile shape id
    int nsplit;    // splits of this tile
    int split;     // this item's split index
    int slot;      // counter index of the tile (split tiles only)
    int64_t poff;  // float offset of the tile's partial slots (split tiles only)
};

// Tile configurations by tile id (the HBS_TILE_* ids of include/hbsgemm.h).
// `big` < 0: the 256-thread launch (4 waves, two workgroups per CU, split-K
// allowed); else the large-tile slot of a 512-thread launch (8 waves, one
// workgroup per CU, whole problems, tiles chained through one pipeline).
// bk_*: K stage depth per precision, 0 where the configuration does not exist.
struct TileCfg {
    int bm, bn;          // tile rows x columns
    int wm, wn;          // waves as wm x wn
    int bk_f32, bk_b3;   // stage depth: FP32, BF16x3
    int big;             // large-tile slot, or -1
};
constexpr int kTileIds = 10;
constexpr TileCfg kTileCfg[kTileIds] = {
    {128, 128, 2, 2, 16, 32, -1},  // 0: HBS_TILE_128x128
    {64, 64, 2, 2, 16, 32, -1},    // 1: HBS_TILE_64x64
    {128, 64, 2, 2, 16, 32, -1},   // 2: HBS_TILE_128x64
    {64, 128, 2, 2, 16, 32, -1},   // 3: HBS_TILE_64x128
    {256, 256, 4, 2, 32, 32, 0},   // 4: HBS_TILE_256x256 (the default)
    {128, 128, 4, 1, 16, 0, -1},   // 5: HBS_TILE_128x128_4x1, waves of 32 x 128 (FP32 only)
    {128, 128, 2, 2, 32, 0, -1},   // 6: HBS_TILE_128x128_BK32 (FP32 only)
    {128, 128, 4, 2, 32, 32, 1},   // 7: HBS_TILE_128x128_8W, 8 waves of 32 x 64
    {256, 128, 4, 2, 32, 32, 2},   // 8: HBS_TILE_256x128, 8 waves of 64 x 64
    {128, 256, 2, 4, 32, 32, 3},   // 9: HBS_TILE_128x256, 8 waves of 64 x 64
};
constexpr int kTileBig = 4;  // the default tile
constexpr int kBigCfgs = 4;  // large-tile slots: tile ids 4, 7, 8, 9
constexpr int kBigCfgTile[kBigCfgs] = {4, 7, 8, 9};
// Large-tile launches: 2 slot + 0 for problems whose K is a multiple of the
// 32-deep stage (clamped loads), 2 slot + 1 for any other K (zero fill past
// K), and kEdgeLaunch for problems whose operands allow no 16-byte loads (the
// 256 x 256 tile with per-element predicated loads).
constexpr int kBigLaunches = 2 * kBigCfgs + 1;
constexpr int kEdgeLaunch = 2 * kBigCfgs;
enum BigMode { kBigClamp = 0, kBigClampK = 1, kBigEdge = 2 };
inline int big_launch_tile(int g) { return g == kEdgeLaunch ? kTileBig : kBigCfgTile[g / 2]; }
inline int big_launch_mode(int g) { return g == kEdgeLaunch ? kBigEdge : g % 2; }
constexpr int kXcdQueues = 8;  // one item queue per XCD on each large-tile launch

struct Choice {
    int tile;
    int split;
    double ns;  // measured time of one full tile (0 = unknown)
};

using TuneKey = std::tuple<int, int, int, int>;  // prec, M class, N class, K class

// Size class of a dimension: the power of two nearest to it in log scale,
// clamped to [64, 2048] (the autotuner measures at these sizes).
int size_class(int x);
// Reads a tuning file into `table`; false if it cannot be opened.
bool load_tuning(const char *path, std::map<TuneKey, Choice> &table);
bool tile_supported(int precision, int tile);
int validate_all(const hbs_gemm_problem *problems, int count, int precision);
bool vec_loads(const hbs_gemm_problem &q);

// The host half of a plan: descriptors, work items in queue order and the
// launch sizes for a GPU of `cus` compute units.  No HIP call.
struct HostPlan {
    std::vector<Prob> probs;
    std::vector<Item> sorted;         // 256-thread list
    std::vector<Item> sorted_big[kBigLaunches];  // large-tile launches, in queue order
    std::vector<int> qoff_big[kBigLaunches];
    int nq_big[kBigLaunches] = {}, nblocks = 0, nblocks_big[kBigLaunches] = {};
    int64_t slots = 0, poff = 0, tiles_total = 0;
    double imb = 1.0, imb_big[kBigLaunches] = {};
    double est_imbalance() const {
        double e = nblocks > 0 ? imb : 1.0;
        for (int g = 0; g < kBigLaunches; ++g)
            if (nblocks_big[g] > 0) e = std::max(e, imb_big[g]);
        return e;
    }
};

// Plans `count` problems for a GPU of `cus` compute units: a forced tile and
// split-K for every problem, else the tuning table's choice per size class,
// else the large tile.  No HIP call.
int plan_host(const hbs_gemm_problem *problems, int count, int precision, const Choice *forced,
              const std::map<TuneKey, Choice> *table, int cus, HostPlan &hp);

}  // namespace grouped
}  // namespace hbs
