// grouped_plan.cpp -- the grouped planner on the host (no HIP): tuning table,
// validation, tile / split-K choice, work-item expansion and the queue order
// of every launch, and the dry run hbs_grouped_plan_describe.
#include "grouped_plan.h"

#include <algorithm>
#include <fstream>
#include <queue>
#include <sstream>
#include <string>

namespace hbs {
namespace grouped {

int size_class(int x) {
    int c = 64;
    while (c < 2048 && (int64_t)x * x >= (int64_t)2 * c * c) c *= 2;  // x >= c * sqrt(2)
    return c;
}

// Text format, one entry per line ('#' starts a comment):
//   <fp32|bf16x3> <M class> <N class> <K class> <tile id> <split> <ns per tile>
bool load_tuning(const char *path, std::map<TuneKey, Choice> &table) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string prec;
        int mc, nc, kc, tile, split;
        double ns = 0.0;
        if (!(ss >> prec >> mc >> nc >> kc >> tile >> split)) continue;
        ss >> ns;
        const int pr = prec == "fp32" ? HBS_PREC_FP32 : prec == "bf16x3" ? HBS_PREC_BF16X3 : -1;
        if (pr < 0 || !tile_supported(pr, tile) || split < 1 || split > 64) continue;
        table[TuneKey(pr, mc, nc, kc)] = Choice{tile, split, ns};
    }
    return true;
}

bool tile_supported(int precision, int tile) {
    if (tile < 0 || tile >= kTileIds) return false;
    return (precision == HBS_PREC_FP32 ? kTileCfg[tile].bk_f32 : kTileCfg[tile].bk_b3) > 0;
}

// By default every problem takes the large-tile path, whatever its M, N and K:
// the 256 x 256 tile runs at a higher per-CU rate than the 128 x 128 tiles of
// two workgroups per CU (for BF16x3 two to five times), its edge tiles skip the
// MFMAs of waves outside M x N and its per-XCD queues even out their shorter
// times, and one launch leaves no second tail.  Any K
// (a multiple of 4, as the grouped API requires) can take it: problems whose K
// is not a multiple of the 32-deep stage form a second launch in which the
// last stage zero-fills past K (kTileClampK).  On such a group that launch
// beat the 256-thread one by 33% (BF16x3) and 2.4% (FP32, whose large tile
// spills 56 bytes in that mode).  Tuning files and forced plans can still
// select the 256-thread tiles per size class.

int validate(const hbs_gemm_problem &q) {
    if (q.M < 1 || q.N < 1 || q.K < 1) return HBS_ERR_ARG;
    if (q.lda < q.K || q.ldb < q.N || q.ldc < q.N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(q.M, q.N, q.K, q.lda, q.ldb, q.ldc)) return HBS_ERR_ARG;
    if (!q.A || !q.B || !q.C) return HBS_ERR_NULL;
    return HBS_SUCCESS;
}

// Whether the problem's operands allow the 16-byte loads of the clamped tile
// modes (A and B 16-byte aligned; N, K, lda and ldb multiples of 4).  Other
// problems run on the large-tile launch with per-element predicated loads.
bool vec_loads(const hbs_gemm_problem &q) {
    return hbs::aligned16(q.A) && hbs::aligned16(q.B) && q.N % 4 == 0 && q.K % 4 == 0 &&
           q.lda % 4 == 0 && q.ldb % 4 == 0;
}

int validate_all(const hbs_gemm_problem *problems, int count, int precision) {
    if (count < 1) return HBS_ERR_ARG;
    if (!problems) return HBS_ERR_NULL;
    if (precision != HBS_PREC_FP32 && precision != HBS_PREC_BF16X3) return HBS_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        const int v = validate(problems[i]);
        if (v) return v;
    }
    return HBS_SUCCESS;
}

int plan_host(const hbs_gemm_problem *problems, int count, int precision, const Choice *forced,
              const std::map<TuneKey, Choice> *table, int cus, HostPlan &hp) {
    const int v = validate_all(problems, count, precision);
    if (v) return v;
    if (cus < 1) return HBS_ERR_ARG;

    // Problems on the 256-thread launch, each with its tile shape and split-K.
    struct Small {
        int prob;
        Choice c;
    };
    std::vector<Prob> probs(count);
    std::vector<Small> small;
    std::vector<Item> items_big[kBigLaunches];
    std::vector<double> cost_big[kBigLaunches];
    // the forced choice, else the tuning table's for the size class, else the
    // large tile (see above)
    auto lookup = [&](int M, int N, int K) {
        if (forced) return *forced;
        if (table) {
            auto f = table->find(TuneKey(precision, size_class(M), size_class(N), size_class(K)));
            if (f != table->end() && tile_supported(precision, f->second.tile)) return f->second;
        }
        return Choice{kTileBig, 1, 0.0};
    };
    for (int i = 0; i < count; ++i) {
        const hbs_gemm_problem &q = problems[i];
        probs[i] = Prob{q.A, q.B, q.C, q.M, q.N, q.K, q.lda, q.ldb, q.ldc};
        Choice c = lookup(q.M, q.N, q.K);
        const bool vec = vec_loads(q);
        // operands that allow no 16-byte loads: the whole problem on the
        // 256 x 256 launch with per-element predicated loads, whatever the
        // tuning says
        if (!vec) c = Choice{kTileBig, 1, 0.0};
        const TileCfg &cfg = kTileCfg[c.tile];
        if (cfg.big < 0) {
            small.push_back(Small{i, c});
            continue;
        }
        // the whole problem in tiles of a large-tile launch; edge tiles clamp
        // their loads (zero fill past K in the last stage when K is not a
        // multiple of 32), or predicate them per element
        const int g = !vec ? kEdgeLaunch : 2 * cfg.big + (q.K % 32 == 0 ? 0 : 1);
        for (int tm = 0; tm < (q.M + cfg.bm - 1) / cfg.bm; ++tm)
            for (int tn = 0; tn < (q.N + cfg.bn - 1) / cfg.bn; ++tn) {
                items_big[g].push_back(Item{i, tm, tn, 0, q.K, c.tile, 1, 0, -1, 0});
                cost_big[g].push_back((double)((q.K + 31) / 32 * 32));
            }
    }

    std::vector<Item> items;
    std::vector<double> cost;
    int64_t slots = 0, poff = 0,
            tiles_total = 0;
    for (int g = 0; g < kBigLaunches; ++g) tiles_total += (int64_t)items_big[g].size();
    for (const Small &sp : small) {
        const Prob &q = probs[sp.prob];
        const Choice &c = sp.c;
        const TileCfg &cfg = kTileCfg[c.tile];
        const int BM = cfg.bm, BN = cfg.bn;
        const int BK = precision == HBS_PREC_FP32 ? cfg.bk_f32 : cfg.bk_b3;
        const int tm_n = (q.M + BM - 1) / BM, tn_n = (q.N + BN - 1) / BN;
        // K range per split, a whole number of stages (int64: 2K may exceed int)
        const int chunk = (int)(hbs::ceil_div(hbs::ceil_div(q.K, c.split), BK) * BK);
        const int ns = (int)hbs::ceil_div(q.K, chunk);
        // Work-item cost in ns of one workgroup slot: the autotuner's measured
        // time when the tuning file has this class, else a throughput model
        // (FP32: 157.3 TFLOP/s, BF16x3: 2516.6 / 3 TFLOP/s-equivalent, shared
        // by 2 x CUs slots) with a tile-shape efficiency.
        const double slots_all = 2.0 * cus;
        const double peak = precision == HBS_PREC_FP32 ? 157.3e12 : 2516.6e12 / 3.0;
        const double eff = BM * BN >= 128 * 128 ? 0.7 : (BM * BN <= 64 * 64 ? 0.35 : 0.5);
        for (int tm = 0; tm < tm_n; ++tm) {
            for (int tn = 0; tn < tn_n; ++tn) {
                ++tiles_total;
                for (int s = 0; s < ns; ++s) {
                    Item it;
                    it.prob = sp.prob;
                    it.tm = tm;
                    it.tn = tn;
                    it.k0 = s * chunk;
                    it.k1 = (int)std::min<int64_t>(q.K, (int64_t)(s + 1) * chunk);
                    it.cfg = c.tile;
                    it.nsplit = ns;
                    it.split = s;
                    it.slot = ns > 1 ? (int)slots : -1;
                    it.poff = ns > 1 ? poff : 0;
                    items.push_back(it);
                    const double klen = it.k1 - it.k0;
                    cost.push_back(c.ns > 0.0 ? c.ns * klen / chunk
                                              : 2.0 * BM * BN * klen / (peak / slots_all * eff) * 1e9 + 2000.0);
                }
                if (ns > 1) {
                    ++slots;
                    poff += (int64_t)ns * BM * BN;
                }
            }
        }
    }
    if (items.size() > 0x7fffffffULL) return HBS_ERR_ARG;

    for (int g = 0; g < kBigLaunches; ++g)
        if (items_big[g].size() > 0x7fffffffULL) return HBS_ERR_ARG;

    // 256-thread launch: items in decreasing cost, handed out in that order to
    // whichever workgroup is free first.  With the modelled costs this is
    // longest-processing-time-first list scheduling; the same simulation
    // gives the imbalance estimate.
    auto lpt = [](const std::vector<Item> &its, const std::vector<double> &cst, int nb,
                  std::vector<Item> &sorted, double &imbalance) {
        std::vector<int64_t> order(its.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int64_t)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return cst[x] > cst[y]; });
        using Load = std::pair<double, int>;
        std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
        for (int b = 0; b < nb; ++b) heap.push(Load(0.0, b));
        for (int64_t idx : order) {
            Load l = heap.top();
            heap.pop();
            l.first += cst[idx];
            heap.push(l);
        }
        double maxl = 0.0, suml = 0.0;
        while (!heap.empty()) {
            maxl = std::max(maxl, heap.top().first);
            suml += heap.top().first;
            heap.pop();
        }
        imbalance = suml > 0.0 ? maxl / (suml / nb) : 1.0;
        sorted.clear();
        sorted.reserve(its.size());
        for (int64_t idx : order) sorted.push_back(its[idx]);
    };
    // Large-tile launch: one workgroup per CU, workgroup b on XCD b % 8, and
    // one item queue per XCD.  The tiles are ordered by problem K (longest
    // first), then by problem, then in a raster of 4 tile rows, and dealt to
    // the queues in runs of 32 (one per workgroup of an XCD): the tiles an XCD
    // runs at the same time are a compact block of one problem and share A and
    // B panels in that XCD's L2.  Within an XCD the workgroups take tiles in
    // queue order as they finish, so edge tiles that run short (their idle
    // waves skip the MFMAs) and tiles that run long even out.
    auto xcd_queues = [&](const std::vector<Item> &its, const std::vector<double> &cst, int nb,
                         std::vector<Item> &sorted, std::vector<int> &qoff, int &nq,
                         double &imbalance) {
        std::vector<int64_t> order(its.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int64_t)i;
        auto key = [&](int64_t x) {
            const Item &t = its[x];
            const int bm = kTileCfg[t.cfg].bm, bn = kTileCfg[t.cfg].bn;
            const int gm = 4, rows = (probs[t.prob].M + bm - 1) / bm;
            const int g = t.tm / gm, first = g * gm, h = std::min(gm, rows - first);
            const int64_t pos = ((int64_t)g * ((probs[t.prob].N + bn - 1) / bn) + t.tn) * h +
                                (t.tm - first);
            return std::make_tuple(-cst[x], t.prob, pos, t.split);
        };
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return key(x) < key(y); });
        nq = nb % kXcdQueues == 0 ? kXcdQueues : 1;
        const int per = nb / nq;  // workgroups per queue
        std::vector<std::vector<int64_t>> lists(nq);
        for (size_t t = 0; t < order.size(); ++t) lists[(t % nb) / per].push_back(order[t]);
        // imbalance estimate: each queue's list scheduled in order onto its
        // `per` workgroups with the modelled costs
        double maxl = 0.0, suml = 0.0;
        for (int q = 0; q < nq; ++q) {
            using Load = std::pair<double, int>;
            std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
            for (int w = 0; w < per; ++w) heap.push(Load(0.0, w));
            for (int64_t idx : lists[q]) {
                Load l = heap.top();
                heap.pop();
                l.first += cst[idx];
                suml += cst[idx];
                heap.push(l);
            }
            while (!heap.empty()) {
                maxl = std::max(maxl, heap.top().first);
                heap.pop();
            }
        }
        imbalance = suml > 0.0 ? maxl / (suml / nb) : 1.0;
        sorted.clear();
        qoff.assign(nq + 1, 0);
        for (int q = 0; q < nq; ++q) {
            qoff[q] = (int)sorted.size();
            for (int64_t idx : lists[q]) sorted.push_back(its[idx]);
        }
        qoff[nq] = (int)sorted.size();
    };
    // resident workgroups per CU on the 256-thread launch: two (up to 256
    // VGPRs each; 38 KB LDS for FP32, 74 KB for BF16x3); the large-tile launch
    // holds one per CU
    const int per_cu = 2;
    const int nblocks = (int)std::min<int64_t>((int64_t)per_cu * cus, (int64_t)items.size());
    std::vector<Item> sorted;
    double imb = 1.0;
    lpt(items, cost, nblocks, sorted, imb);
    for (int g = 0; g < kBigLaunches; ++g) {
        hp.nblocks_big[g] = (int)std::min<int64_t>(cus, (int64_t)items_big[g].size());
        xcd_queues(items_big[g], cost_big[g], hp.nblocks_big[g], hp.sorted_big[g],
                   hp.qoff_big[g], hp.nq_big[g], hp.imb_big[g]);
    }

    hp.probs = std::move(probs);
    hp.sorted = std::move(sorted);
    hp.nblocks = nblocks;
    hp.slots = slots;
    hp.poff = poff;
    hp.tiles_total = tiles_total;
    hp.imb = imb;
    return HBS_SUCCESS;
}

}  // namespace grouped
}  // namespace hbs

using namespace hbs::grouped;

extern "C" int hbs_grouped_plan_describe(const hbs_gemm_problem *problems, int count,
                                         int precision, const char *tuning_file, int tile,
                                         int split_k, int cus, hbs_grouped_info *info) {
    if (!info) return HBS_ERR_NULL;
    const bool tuned = tuning_file && tuning_file[0];
    std::map<TuneKey, Choice> table;
    const Choice c{tile, split_k, 0.0};
    if (tile >= 0) {
        if (tuned || !tile_supported(precision, tile) || split_k < 1 || split_k > 64)
            return HBS_ERR_ARG;
    } else if (tile != -1 || (tuned && !load_tuning(tuning_file, table))) {
        return HBS_ERR_ARG;
    }
    HostPlan hp;
    const int st = plan_host(problems, count, precision, tile >= 0 ? &c : nullptr,
                             tuned ? &table : nullptr, cus, hp);
    if (st) return st;
    info->problems = count;
    info->workgroups = hp.nblocks;
    info->work_items = (int64_t)hp.sorted.size();
    for (int g = 0; g < kBigLaunches; ++g) {
        info->workgroups += hp.nblocks_big[g];
        info->work_items += (int64_t)hp.sorted_big[g].size();
    }
    info->tiles = hp.tiles_total;
    info->split_tiles = hp.slots;
    info->est_imbalance = hp.est_imbalance();
    return HBS_SUCCESS;
}
