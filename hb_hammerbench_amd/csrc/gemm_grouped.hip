// gemm_grouped.hip -- grouped persistent GEMM: many independent FP32 problems
// of uneven M/N/K in one launch, FP32-MFMA or BF16x3 arithmetic.
//
// No reference counterpart: the batching pattern built around apps/sgemm's
// C = A.B.  The reference runs its NITER equal-size slices in one launch
// (apps/sgemm/kernel.cpp:293-297); this generalises that to uneven problems.
//
// Design:
//  * a plan (host side, built once) chooses a tile shape and split-K factor per
//    problem -- from the autotuner's tuning file or a forced choice, else the
//    256 x 256 large tile for every problem -- and expands every problem into
//    work items (one output tile times one K range);
//  * two persistent launches, captured together in a hipGraph, each with a
//    fixed number of workgroups walking their own item lists (no host work or
//    memset between replays):
//      - the large-tile launch, one 512-thread workgroup per CU, runs the
//        square kernels' 256 x 256 tile (tile_f32.h, tile_bf16x3.h) over whole
//        problems, edge tiles with clamped loads; its tiles are dealt in
//        XCD-local raster blocks to one item queue per XCD;
//      - the 256-thread launch, two workgroups per CU, runs 128 x 128 to
//        64 x 64 tiles with the same main loops over each item's K range; its
//        workgroups take items from one list in decreasing cost order through
//        an atomic queue head (greedy list scheduling, which adapts to the
//        measured rather than the modelled item times), and the last
//        workgroup to leave resets the queue for the next run;
//  * split-K tiles write FP32 partials; the last of a tile's splits to finish
//    (an agent-scope counter) adds them in split order -- deterministic -- and
//    resets the counter to zero for the next run;
//  * BF16x3 tiles read the float32 operands directly and split them into bf16
//    hi/lo planes in registers on the way to LDS: no pre-pass and no operand
//    copies.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <map>
#include <mutex>
#include <queue>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "tile_bf16x3.h"
#include "tile_f32.h"

constexpr int kBigLaunches = 3;  // large-tile launches: clamped, zero fill past K, predicated

struct hbs_grouped_plan {
    int device;
    int precision;
    int count;
    int nblocks;
    int64_t items;
    int64_t tiles;
    int64_t split_tiles;
    double est_imbalance;
    void *dmem;  // problems | items | counters + queues | partials | large-tile items, offsets
    const void *d_probs;
    const void *d_items;  // 256-thread launch: one list, decreasing cost
    int *d_counters;      // split-K counters
    int *d_queue;         // queue head and finished-workgroup count (zero at rest)
    float *d_partials;
    // large-tile launches (whole problems in 256 x 256 tiles): [0] problems
    // whose K is a multiple of the 32-deep stage (kTileClamp), [1] problems
    // with any other K (kTileClampK), [2] problems whose operands do not allow
    // 16-byte loads (kTileEdge)
    int nblocks_big[kBigLaunches];
    int64_t items_big[kBigLaunches];
    const void *d_items_big[kBigLaunches];
    const int *d_qoff_big[kBigLaunches];  // item range of each of its nq_big queues
    int nq_big[kBigLaunches];
    int *d_queue_big[kBigLaunches];       // nq_big heads and a finished-workgroup count
    int graphs;                // executable graphs made from this plan and not yet destroyed
};

namespace {

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
constexpr int kThreads = 256;
constexpr int kTileBig = 4, kBig = 256, kBigThreads = 512;

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
// Epilogue for a 2x2-wave tile of FMxFN 32x32 accumulator blocks per wave.
// Direct tiles store C; split tiles store a partial, and the last split to
// arrive sums all partials in split order and stores C.
template <int BM, int BN>
__device__ __forceinline__ void finish_tile(floatx16 (&acc)[BM / 64][BN / 64], const Prob &p,
                                            const Item &it, float *partials, int *counters,
                                            int *s_flag) {
    constexpr int FM = BM / 64, FN = BN / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l32 = lane & 31, h = lane >> 5;
    const int m0 = it.tm * BM, n0 = it.tn * BN;

    if (it.nsplit > 1) {
        float *own = partials + it.poff + (int64_t)it.split * BM * BN + wave * (FM * FN * 16 * 64);
#pragma unroll
        for (int r = 0; r < FM; ++r)
#pragma unroll
            for (int c = 0; c < FN; ++c)
#pragma unroll
                for (int q = 0; q < 16; ++q) own[((r * FN + c) * 16 + q) * 64 + lane] = acc[r][c][q];
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            const int old = __hip_atomic_fetch_add(&counters[it.slot], 1, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            *s_flag = (old == it.nsplit - 1);
        }
        __syncthreads();
        if (!*s_flag) return;
        __threadfence();
        // last split: sum the partials in split order straight into C, one
        // 32x32 block at a time (the accumulators are no longer needed)
#pragma unroll 1
        for (int b = 0; b < FM * FN; ++b) {
            const int r = b / FN, c = b % FN;
            const int row0 = m0 + wm * (BM / 2) + r * 32 + 4 * h;
            const int col = n0 + wn * (BN / 2) + c * 32 + l32;
            float *cp = p.C + (int64_t)row0 * p.ldc + col;
            const float *src0 = partials + it.poff + wave * (FM * FN * 16 * 64) + b * 16 * 64 + lane;
            float v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = 0.0f;
#pragma unroll 1
            for (int sp = 0; sp < it.nsplit; ++sp) {
                const float *src = src0 + (int64_t)sp * BM * BN;
#pragma unroll
                for (int q = 0; q < 16; ++q) v[q] += __builtin_nontemporal_load(&src[q * 64]);
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int dr = 8 * (q >> 2) + (q & 3);
                if (col < p.N && row0 + dr < p.M) hbs::stg(cp + (int64_t)dr * p.ldc, v[q]);
            }
        }
        if (tid == 0)
            __hip_atomic_store(&counters[it.slot], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    // one base pointer per 32x32 block; the per-register row offsets are
    // wave-uniform multiples of ldc
#pragma unroll
    for (int r = 0; r < FM; ++r) {
        const int row0 = m0 + wm * (BM / 2) + r * 32 + 4 * h;
#pragma unroll
        for (int c = 0; c < FN; ++c) {
            const int col = n0 + wn * (BN / 2) + c * 32 + l32;
            float *cp = p.C + (int64_t)row0 * p.ldc + col;
            const bool colok = col < p.N;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int dr = 8 * (q >> 2) + (q & 3);
                if (colok && row0 + dr < p.M) hbs::stg(cp + (int64_t)dr * p.ldc, acc[r][c][q]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Tiles of the 256-thread launch (2 x 2 waves): the square kernels' main loops
// (tile_f32.h, BK = 16; tile_bf16x3.h, BK = 32) over the item's K range
// [k0, k1), then finish_tile, with clamped loads (the API requires
// 16-byte-loadable A and B and N, K multiples of 4; split ranges start on a
// stage).
// ---------------------------------------------------------------------------
template <int BM, int BN>
__device__ void f32_tile(float *smem, const Prob &p, const Item &it, float *partials,
                         int *counters, int *s_flag) {
    using G = hbs::F32Cfg<BM, BN, 16, 2, 2>;
    floatx16 acc[G::kFM][G::kFN];
    const int m0 = it.tm * BM, n0 = it.tn * BN, K = it.k1 - it.k0;
    const float *Ag = p.A + (int64_t)m0 * p.lda + it.k0;
    const float *Bg = p.B + (int64_t)it.k0 * p.ldb + n0;
    const hbs::TileArgs t{Ag, Bg, p.M, p.N, K, p.lda, p.ldb, m0, n0};
    typename G::Regs regs;
    hbs::NoChainGlobal ch;  // operands from the problem table: global loads
    hbs::f32_block_mma<BM, BN, 16, 2, 2, hbs::kTileClampK>(smem, t, acc, regs, ch);
    finish_tile<BM, BN>(acc, p, it, partials, counters, s_flag);
}

template <int BM, int BN>
__device__ void b3_tile(uint16_t *smem, const Prob &p, const Item &it, float *partials,
                        int *counters, int *s_flag) {
    using G = hbs::B3Cfg<BM, BN, 32, 2, 2>;
    floatx16 acc[G::kFM][G::kFN];
    const int m0 = it.tm * BM, n0 = it.tn * BN, K = it.k1 - it.k0;
    const float *Ag = p.A + (int64_t)m0 * p.lda + it.k0;
    const float *Bg = p.B + (int64_t)it.k0 * p.ldb + n0;
    const hbs::TileArgs t{Ag, Bg, p.M, p.N, K, p.lda, p.ldb, m0, n0};
    typename G::Regs regs;
    hbs::NoChainGlobal ch;  // operands from the problem table: global loads
    hbs::bf16x3_block_mma<G, hbs::kTileClampK>(smem, t, acc, regs, ch);
    finish_tile<BM, BN>(acc, p, it, partials, counters, s_flag);
}

constexpr int kLdsF32 = hbs::F32Cfg<128, 128, 16, 2, 2>::kLdsFloats * 4;
constexpr int kLdsB3 = 2 * hbs::B3Cfg<128, 128, 32, 2, 2>::kStage * 2;

template <int PREC>
__global__ __launch_bounds__(kThreads, 2) void grouped_kernel(const Prob *__restrict__ probs,
                                                              const Item *__restrict__ items,
                                                              int nitems, float *partials,
                                                              int *counters, int *queue) {
    constexpr int LDS = PREC == HBS_PREC_FP32 ? kLdsF32 : kLdsB3;
    __shared__ __attribute__((aligned(16))) char smem[LDS];
    __shared__ int s_flag, s_next;
    // queue[0]: next item to hand out; queue[1]: workgroups that have left.
    // Thread 0 fetches the next index when an item starts and publishes it in
    // LDS when the item ends, so the fetch is hidden behind the item.
    if (threadIdx.x == 0)
        s_next = __hip_atomic_fetch_add(&queue[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    int i = s_next;
    while (i < nitems) {
        int next = 0;
        if (threadIdx.x == 0)
            next = __hip_atomic_fetch_add(&queue[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const Item it = items[i];
        const Prob p = probs[it.prob];
        if constexpr (PREC == HBS_PREC_FP32) {
            float *s = (float *)smem;
            switch (it.cfg) {
                case 0: f32_tile<128, 128>(s, p, it, partials, counters, &s_flag); break;
                case 1: f32_tile<64, 64>(s, p, it, partials, counters, &s_flag); break;
                case 2: f32_tile<128, 64>(s, p, it, partials, counters, &s_flag); break;
                default: f32_tile<64, 128>(s, p, it, partials, counters, &s_flag); break;
            }
        } else {
            uint16_t *s = (uint16_t *)smem;
            switch (it.cfg) {
                case 0: b3_tile<128, 128>(s, p, it, partials, counters, &s_flag); break;
                case 1: b3_tile<64, 64>(s, p, it, partials, counters, &s_flag); break;
                case 2: b3_tile<128, 64>(s, p, it, partials, counters, &s_flag); break;
                default: b3_tile<64, 128>(s, p, it, partials, counters, &s_flag); break;
            }
        }
        if (threadIdx.x == 0) s_next = next;
        __syncthreads();  // s_flag and LDS are reused by the next item; s_next is set
        i = s_next;
        __syncthreads();  // every thread has read s_next
    }
    // every fetch of this workgroup has returned; the last workgroup to get
    // here resets the queue (no fetch can follow: all others have left)
    if (threadIdx.x == 0 &&
        __hip_atomic_fetch_add(&queue[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
            (int)gridDim.x - 1) {
        __hip_atomic_store(&queue[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&queue[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Large-tile launch: one 512-thread workgroup per CU takes 256 x 256 tiles
// from its XCD's queue and runs them with the square kernels' tile code (waves
// whose sub-tile lies outside M x N skip their MFMAs).  A problem on this path
// is covered whole: tiles at its ragged bottom and right edges read clamped
// rows and columns (kTileClamp) and store only what lies inside M x N.  No
// split-K.  MODE kTileClamp takes problems whose K is a multiple of the
// 32-deep stage; kTileEdge (a launch of its own) problems whose A or B is not
// 16-byte aligned or whose N, K, lda or ldb is not a multiple of 4, with
// per-element predicated loads; kTileClampK (a launch of its own) the others,
// zero-filling
// past K in the last stage.  One kernel holding both modes ran out of
// registers and cost the first kind 5%.
// ---------------------------------------------------------------------------
using BigF32 = hbs::F32Cfg<256, 256, 32, 4, 2>;
constexpr int kXcdQueues = 8;  // one item queue per XCD on the large-tile launch
using BigB3 = hbs::B3Main;
static_assert(BigF32::kThreads == kBigThreads && BigB3::kThreads == kBigThreads, "big tiles");

template <int PREC, int MODE>
__global__ __launch_bounds__(kBigThreads) void grouped_big_kernel(const Prob *__restrict__ probs,
                                                                  const Item *__restrict__ items,
                                                                  const int *__restrict__ qoff,
                                                                  int nq, int *queue) {
    constexpr bool kF32 = PREC == HBS_PREC_FP32;
    static_assert(MODE == hbs::kTileClamp || MODE == hbs::kTileClampK || MODE == hbs::kTileEdge,
                  "mode");
    constexpr int LDS = kF32 ? BigF32::kLdsFloats * 4 : 2 * BigB3::kStage * 2;
    using Regs = std::conditional_t<kF32, BigF32::Regs, BigB3::Regs>;
    static_assert(BigF32::kFM == BigB3::kFM && BigF32::kFN == BigB3::kFN, "accumulators");
    __shared__ __attribute__((aligned(16))) char smem[LDS];
    __shared__ int s_next[2];
    // queue q = blockIdx.x % nq (the workgroup's XCD when nq = 8) hands out
    // items[qoff[q] .. qoff[q + 1]) in order; queue[q] is its head and
    // queue[kXcdQueues] counts the workgroups that have left
    const int q = blockIdx.x % nq, first = qoff[q], end = qoff[q + 1];
    auto fetch = [&]() {
        return first +
               __hip_atomic_fetch_add(&queue[q], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto tile = [&](int i, hbs::TileArgs &t, float *&Cg, int &ldc) {
        const Item it = items[i];
        const Prob p = probs[it.prob];
        const int m0 = it.tm * kBig, n0 = it.tn * kBig;
        t = hbs::TileArgs{p.A + (int64_t)m0 * p.lda, p.B + n0, p.M, p.N, p.K, p.lda, p.ldb, m0, n0};
        Cg = p.C;
        ldc = p.ldc;
    };
    if (threadIdx.x == 0) s_next[0] = fetch();
    __syncthreads();
    const int i0 = __builtin_amdgcn_readfirstlane(s_next[0]);
    if (i0 < end) {
        // Tiles run back to back through one pipeline: while a tile of three
        // or more K stages finishes, the next one's first two stages are
        // loaded (hbs::TileChain).  Thread 0 fetches the next item at the
        // start of each tile and publishes it in s_next[pw]; the two slots
        // alternate, so a slot is rewritten only after every thread has read it.
        hbs::TileArgs t;
        float *Cg, *nCg = nullptr;
        int ldc, nldc = 0, nxt = 0, pw = 1;
        tile(i0, t, Cg, ldc);
        auto publish = [&]() {
            if (threadIdx.x == 0) s_next[pw] = nxt;
        };
        auto peek = [&](hbs::TileArgs &nt) {
            const int j = __builtin_amdgcn_readfirstlane(s_next[pw]);
            if (j >= end) return false;
            tile(j, nt, nCg, nldc);
            return true;
        };
        auto ch = hbs::make_chain(publish, peek);
        Regs regs;
        floatx16 acc[BigF32::kFM][BigF32::kFN];
        for (;;) {
            if (threadIdx.x == 0) nxt = fetch();
            if constexpr (kF32) {
                hbs::f32_block_mma<256, 256, 32, 4, 2, MODE>((float *)smem, t, acc, regs, ch);
                hbs::f32_block_store<256, 256, 32, 4, 2, hbs::kTileClamp>(Cg, t.M, t.N, ldc, t.m0,
                                                                          t.n0, acc);
            } else {
                hbs::bf16x3_block_mma<BigB3, MODE>((uint16_t *)smem, t, acc, regs, ch);
                hbs::bf16x3_block_store<BigB3, hbs::kTileClamp>(Cg, t.M, t.N, ldc, t.m0, t.n0,
                                                                acc);
            }
            if (!ch.has_next) break;
            t = ch.next;
            Cg = nCg;
            ldc = nldc;
            pw ^= 1;
        }
    }
    // every fetch of this workgroup has returned (its value was published);
    // the last workgroup to get here resets the queues
    if (threadIdx.x == 0 &&
        __hip_atomic_fetch_add(&queue[kXcdQueues], 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1) {
        for (int j = 0; j <= kXcdQueues; ++j)
            __hip_atomic_store(&queue[j], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// host side: tuning table, plan construction
// ---------------------------------------------------------------------------
// Size class of a dimension: the power of two nearest to it in log scale,
// clamped to [64, 2048] (the autotuner measures at these sizes).
int size_class(int x) {
    int c = 64;
    while (c < 2048 && (int64_t)x * x >= (int64_t)2 * c * c) c *= 2;  // x >= c * sqrt(2)
    return c;
}

struct Choice {
    int tile;
    int split;
    double ns;  // measured time of one full tile (0 = unknown)
};

using TuneKey = std::tuple<int, int, int, int>;  // prec, M class, N class, K class

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
        if (pr < 0 || tile < 0 || tile > kTileBig || split < 1 || split > 64) continue;
        table[TuneKey(pr, mc, nc, kc)] = Choice{tile, split, ns};
    }
    return true;
}

bool tile_supported(int precision, int tile) {
    (void)precision;
    return tile >= 0 && tile <= kTileBig;
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
    // by launch: K a multiple of 32, other K, no 16-byte loads
    std::vector<Item> items_big[kBigLaunches];
    std::vector<double> cost_big[kBigLaunches];
    auto lookup = [&](int M, int N, int K, bool &big) {
        Choice c{0, 1, 0.0};  // the 256-thread tile, when one is chosen
        big = true;           // the default (see above)
        if (forced) {
            big = forced->tile == kTileBig;
            if (!big) c = *forced;
        } else if (table) {
            auto f = table->find(TuneKey(precision, size_class(M), size_class(N), size_class(K)));
            if (f != table->end() && tile_supported(precision, f->second.tile)) {
                big = f->second.tile == kTileBig;
                if (!big) c = f->second;
            }
        }
        return c;
    };
    for (int i = 0; i < count; ++i) {
        const hbs_gemm_problem &q = problems[i];
        probs[i] = Prob{q.A, q.B, q.C, q.M, q.N, q.K, q.lda, q.ldb, q.ldc};
        bool big = false;
        const Choice c = lookup(q.M, q.N, q.K, big);
        const bool vec = vec_loads(q);
        if (!big && vec) {
            small.push_back(Small{i, c});
            continue;
        }
        // the whole problem in 256 x 256 tiles; edge tiles clamp their loads,
        // or predicate them per element when the operands allow no 16-byte
        // loads (such problems take this launch whatever the tuning says)
        const int g = !vec ? 2 : q.K % 32 == 0 ? 0 : 1;
        for (int tm = 0; tm < (q.M + kBig - 1) / kBig; ++tm)
            for (int tn = 0; tn < (q.N + kBig - 1) / kBig; ++tn) {
                items_big[g].push_back(Item{i, tm, tn, 0, q.K, kTileBig, 1, 0, -1, 0});
                cost_big[g].push_back((double)((q.K + 31) / 32 * 32));
            }
    }

    std::vector<Item> items;
    std::vector<double> cost;
    int64_t slots = 0, poff = 0,
            tiles_total = (int64_t)(items_big[0].size() + items_big[1].size() +
                                    items_big[2].size());
    for (const Small &sp : small) {
        const Prob &q = probs[sp.prob];
        const Choice &c = sp.c;
        const int BM = kTileBM[c.tile], BN = kTileBN[c.tile];
        const int BK = precision == HBS_PREC_FP32 ? 16 : 32;
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
        const double eff = (c.tile == 0) ? 0.7 : (c.tile == 1 ? 0.35 : 0.5);
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
            const int bm = kTileBM[t.cfg], bn = kTileBN[t.cfg];
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

int build_plan(const hbs_gemm_problem *problems, int count, int precision, const Choice *forced,
               const std::map<TuneKey, Choice> *table, hbs_grouped_plan **out) {
    if (!out) return HBS_ERR_NULL;
    *out = nullptr;
    int st = validate_all(problems, count, precision);
    if (st) return st;
    st = hbs::check_device();
    if (st) return st;
    int dev = 0, cus = 256;
    HBS_HIP_TRY(hipGetDevice(&dev));
    HBS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HostPlan hp;
    st = plan_host(problems, count, precision, forced, table, cus, hp);
    if (st) return st;
    const std::vector<Prob> &probs = hp.probs;
    const std::vector<Item> &sorted = hp.sorted;
    const int nblocks = hp.nblocks;
    const int64_t slots = hp.slots, poff = hp.poff, tiles_total = hp.tiles_total;

    // one device allocation: problems | items | split counters + queues |
    // partials | per large-tile launch: items, queue offsets
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t b_probs = align(probs.size() * sizeof(Prob));
    const size_t b_items = align(std::max<size_t>(sorted.size(), 1) * sizeof(Item));
    // split counters, the 256-thread launch's queue (2), each large-tile
    // launch's queues (kXcdQueues heads + 1)
    const size_t b_cnt = align((slots + 2 + kBigLaunches * (kXcdQueues + 1)) * sizeof(int));
    const size_t b_part = align(std::max<int64_t>(poff, 1) * sizeof(float));
    const size_t o_items = b_probs, o_cnt = o_items + b_items;
    const size_t o_part = o_cnt + b_cnt;
    size_t o_items_big[kBigLaunches], o_qoff_big[kBigLaunches], total = o_part + b_part;
    for (int g = 0; g < kBigLaunches; ++g) {
        o_items_big[g] = total;
        total += align(std::max<size_t>(hp.sorted_big[g].size(), 1) * sizeof(Item));
        o_qoff_big[g] = total;
        total += align(hp.qoff_big[g].size() * sizeof(int));
    }
    char *dmem = nullptr;
    HBS_HIP_TRY(hipMalloc(&dmem, total));
    hbs_grouped_plan *plan = new hbs_grouped_plan();
    plan->device = dev;
    plan->precision = precision;
    plan->count = count;
    plan->nblocks = nblocks;
    plan->items = (int64_t)sorted.size();
    plan->tiles = tiles_total;
    plan->split_tiles = slots;
    plan->est_imbalance = hp.est_imbalance();
    plan->dmem = dmem;
    plan->d_probs = dmem;
    plan->d_items = dmem + o_items;
    plan->d_counters = (int *)(dmem + o_cnt);
    plan->d_queue = plan->d_counters + slots;
    plan->d_partials = (float *)(dmem + o_part);
    for (int g = 0; g < kBigLaunches; ++g) {
        plan->d_queue_big[g] = plan->d_queue + 2 + g * (kXcdQueues + 1);
        plan->nblocks_big[g] = hp.nblocks_big[g];
        plan->items_big[g] = (int64_t)hp.sorted_big[g].size();
        plan->d_items_big[g] = dmem + o_items_big[g];
        plan->d_qoff_big[g] = (const int *)(dmem + o_qoff_big[g]);
        plan->nq_big[g] = hp.nq_big[g];
    }
    hipError_t e = hipMemcpy(dmem, probs.data(), probs.size() * sizeof(Prob), hipMemcpyHostToDevice);
    if (e == hipSuccess && !sorted.empty())
        e = hipMemcpy(dmem + o_items, sorted.data(), sorted.size() * sizeof(Item),
                      hipMemcpyHostToDevice);
    for (int g = 0; g < kBigLaunches; ++g) {
        const std::vector<Item> &sb = hp.sorted_big[g];
        const std::vector<int> &qo = hp.qoff_big[g];
        if (e == hipSuccess && !sb.empty())
            e = hipMemcpy(dmem + o_items_big[g], sb.data(), sb.size() * sizeof(Item),
                          hipMemcpyHostToDevice);
        if (e == hipSuccess && !qo.empty())
            e = hipMemcpy(dmem + o_qoff_big[g], qo.data(), qo.size() * sizeof(int),
                          hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemset(plan->d_counters, 0, b_cnt);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(dmem);
        delete plan;
        return (int)e;
    }
    *out = plan;
    return HBS_SUCCESS;
}

}  // namespace

extern "C" int hbs_grouped_plan_create(const hbs_gemm_problem *problems, int count, int precision,
                                       const char *tuning_file, hbs_grouped_plan **plan) {
    std::map<TuneKey, Choice> table;
    const bool have = tuning_file && tuning_file[0] && load_tuning(tuning_file, table);
    if (tuning_file && tuning_file[0] && !have) {
        if (plan) *plan = nullptr;
        return HBS_ERR_ARG;
    }
    return build_plan(problems, count, precision, nullptr, have ? &table : nullptr, plan);
}

extern "C" int hbs_grouped_plan_create_forced(const hbs_gemm_problem *problems, int count,
                                              int precision, int tile, int split_k,
                                              hbs_grouped_plan **plan) {
    if (!tile_supported(precision, tile) || split_k < 1 || split_k > 64) {
        if (plan) *plan = nullptr;
        return HBS_ERR_ARG;
    }
    const Choice c{tile, split_k, 0.0};
    return build_plan(problems, count, precision, &c, nullptr, plan);
}

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
    info->workgroups = hp.nblocks + hp.nblocks_big[0] + hp.nblocks_big[1] + hp.nblocks_big[2];
    info->tiles = hp.tiles_total;
    info->work_items = (int64_t)(hp.sorted.size() + hp.sorted_big[0].size() +
                                 hp.sorted_big[1].size() + hp.sorted_big[2].size());
    info->split_tiles = hp.slots;
    info->est_imbalance = hp.est_imbalance();
    return HBS_SUCCESS;
}

extern "C" int hbs_grouped_plan_info(const hbs_grouped_plan *plan, hbs_grouped_info *info) {
    if (!plan || !info) return HBS_ERR_NULL;
    info->problems = plan->count;
    info->workgroups =
        plan->nblocks + plan->nblocks_big[0] + plan->nblocks_big[1] + plan->nblocks_big[2];
    info->tiles = plan->tiles;
    info->work_items = plan->items + plan->items_big[0] + plan->items_big[1] + plan->items_big[2];
    info->split_tiles = plan->split_tiles;
    info->est_imbalance = plan->est_imbalance;
    return HBS_SUCCESS;
}

extern "C" int hbs_grouped_run(const hbs_grouped_plan *plan, void *stream) {
    if (!plan) return HBS_ERR_NULL;
    int dev = 0;
    HBS_HIP_TRY(hipGetDevice(&dev));
    if (dev != plan->device) return HBS_ERR_DEVICE;
    hipStream_t s = (hipStream_t)stream;
    const Prob *pr = (const Prob *)plan->d_probs;
    const Item *it = (const Item *)plan->d_items;
    const bool fp32 = plan->precision == HBS_PREC_FP32;
    for (int g = 0; g < kBigLaunches; ++g) {
        if (plan->nblocks_big[g] == 0) continue;
        const Item *itb = (const Item *)plan->d_items_big[g];
        const dim3 grid(plan->nblocks_big[g]);
        auto launch = [&](auto kernel) {
            kernel<<<grid, kBigThreads, 0, s>>>(pr, itb, plan->d_qoff_big[g], plan->nq_big[g],
                                                plan->d_queue_big[g]);
        };
        if (fp32) {
            if (g == 0) launch(grouped_big_kernel<HBS_PREC_FP32, hbs::kTileClamp>);
            else if (g == 1) launch(grouped_big_kernel<HBS_PREC_FP32, hbs::kTileClampK>);
            else launch(grouped_big_kernel<HBS_PREC_FP32, hbs::kTileEdge>);
        } else {
            if (g == 0) launch(grouped_big_kernel<HBS_PREC_BF16X3, hbs::kTileClamp>);
            else if (g == 1) launch(grouped_big_kernel<HBS_PREC_BF16X3, hbs::kTileClampK>);
            else launch(grouped_big_kernel<HBS_PREC_BF16X3, hbs::kTileEdge>);
        }
        const int rc = hbs::launch_status();
        if (rc) return rc;
    }
    if (plan->nblocks > 0) {
        if (fp32)
            grouped_kernel<HBS_PREC_FP32><<<plan->nblocks, kThreads, 0, s>>>(
                pr, it, (int)plan->items, plan->d_partials, plan->d_counters, plan->d_queue);
        else
            grouped_kernel<HBS_PREC_BF16X3><<<plan->nblocks, kThreads, 0, s>>>(
                pr, it, (int)plan->items, plan->d_partials, plan->d_counters, plan->d_queue);
        return hbs::launch_status();
    }
    return HBS_SUCCESS;
}

// Executable graphs made by hbs_grouped_graph_create and the plan each one
// replays: a plan cannot be destroyed while one of them exists (its queues,
// counters and partials are baked into the graph's kernel arguments).
std::mutex graph_mu;
std::map<void *, hbs_grouped_plan *> graph_plans;

extern "C" int hbs_grouped_plan_destroy(hbs_grouped_plan *plan) {
    if (!plan) return HBS_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(graph_mu);
        if (plan->graphs > 0) return HBS_ERR_IN_USE;
    }
    hipError_t e = hipFree(plan->dmem);
    delete plan;
    return e == hipSuccess ? HBS_SUCCESS : (int)e;
}

// hipGraph helpers: capture one grouped run into an executable graph.
extern "C" int hbs_grouped_graph_create(const hbs_grouped_plan *plan, void **graph_exec) {
    if (!plan || !graph_exec) return HBS_ERR_NULL;
    *graph_exec = nullptr;
    hipStream_t s;
    HBS_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    int rc = HBS_SUCCESS;
    if (e == hipSuccess) {
        rc = hbs_grouped_run(plan, s);
        e = hipStreamEndCapture(s, &g);
    }
    hipGraphExec_t ex = nullptr;
    if (e == hipSuccess && rc == HBS_SUCCESS) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(s);
    if (rc != HBS_SUCCESS) return rc;
    if (e != hipSuccess) return (int)e;
    {
        std::lock_guard<std::mutex> lk(graph_mu);
        graph_plans[(void *)ex] = const_cast<hbs_grouped_plan *>(plan);
        ++const_cast<hbs_grouped_plan *>(plan)->graphs;
    }
    *graph_exec = (void *)ex;
    return HBS_SUCCESS;
}

extern "C" int hbs_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return HBS_ERR_NULL;
    HBS_HIP_TRY(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return HBS_SUCCESS;
}

extern "C" int hbs_graph_destroy(void *graph_exec) {
    if (!graph_exec) return HBS_SUCCESS;
    HBS_HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    std::lock_guard<std::mutex> lk(graph_mu);
    auto f = graph_plans.find(graph_exec);
    if (f != graph_plans.end()) {
        --f->second->graphs;
        graph_plans.erase(f);
    }
    return HBS_SUCCESS;
}
