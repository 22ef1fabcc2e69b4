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

// Tile shapes: id -> (BM, BN).  Ids 0..3 run on the 256-thread launch as
// 2 x 2 waves; id 4 (kTileBig) is the 512-thread large-tile launch.
constexpr int kTileBM[5] = {128, 64, 128, 64, 256};
constexpr int kTileBN[5] = {128, 64, 64, 128, 256};
constexpr int kTileBig = 4, kBig = 256;
constexpr int kBigLaunches = 3;  // large-tile launches: clamped, zero fill past K, predicated
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
    int nq_big[kBigLaunches] = {1, 1, 1}, nblocks = 0, nblocks_big[kBigLaunches] = {0, 0, 0};
    int64_t slots = 0, poff = 0, tiles_total = 0;
    double imb = 1.0, imb_big[kBigLaunches] = {1.0, 1.0, 1.0};
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
