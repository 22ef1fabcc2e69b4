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
#include <map>
#include <mutex>
#include <vector>

#include "grouped_plan.h"
#include "tile_bf16x3.h"
#include "tile_f32.h"

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
    int nblocks_big[hbs::grouped::kBigLaunches];
    int64_t items_big[hbs::grouped::kBigLaunches];
    const void *d_items_big[hbs::grouped::kBigLaunches];
    const int *d_qoff_big[hbs::grouped::kBigLaunches];  // item range of each of its nq_big queues
    int nq_big[hbs::grouped::kBigLaunches];
    int *d_queue_big[hbs::grouped::kBigLaunches];       // nq_big heads and a finished-workgroup count
    int graphs;                // executable graphs made from this plan and not yet destroyed
};

namespace {

using namespace hbs::grouped;

constexpr int kThreads = 256;
constexpr int kBigThreads = 512;

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------
// Epilogue for a tile of WM x WN waves, each with FM x FN 32x32 accumulator
// blocks (wave w at wm = w % WM, wn = w / WM, as the tile configs place it).
// Direct tiles store C; split tiles store a partial, and the last split to
// arrive sums all partials in split order and stores C.  Q16: the
// accumulator layout of store_acc (the tile configuration's kQ16 for BF16x3,
// false for FP32).
template <int BM, int BN, int WM, int WN, bool Q16>
__device__ __forceinline__ void finish_tile(floatx16 (&acc)[BM / WM / 32][BN / WN / 32],
                                            const Prob &p, const Item &it, float *partials,
                                            int *counters, int *s_flag) {
    constexpr int FM = BM / WM / 32, FN = BN / WN / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
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
            const int row0 = m0 + wm * (BM / WM) + r * 32 + hbs::acc_lane_row<Q16>(lane);
            const int col = n0 + wn * (BN / WN) + c * 32 + hbs::acc_lane_col<Q16>(lane);
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
                const int dr = hbs::acc_row<Q16>(q), dc = hbs::acc_col<Q16>(q);
                if (col + dc < p.N && row0 + dr < p.M) hbs::stg(cp + ((int64_t)dr * p.ldc + dc), v[q]);
            }
        }
        if (tid == 0)
            __hip_atomic_store(&counters[it.slot], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const int rb = m0 + wm * (BM / WM), cb = n0 + wn * (BN / WN);
    if (m0 + BM <= p.M && n0 + BN <= p.N)
        hbs::store_acc<FM, FN, false, Q16>(p.C, p.ldc, rb, cb, p.M, p.N, FN, acc);
    else
        hbs::store_acc<FM, FN, true, Q16>(p.C, p.ldc, rb, cb, p.M, p.N, FN, acc);
}

// ---------------------------------------------------------------------------
// Tiles of the 256-thread launch (2 x 2 waves): the square kernels' main loops
// (tile_f32.h, BK = 16; tile_bf16x3.h, BK = 32) over the item's K range
// [k0, k1), then finish_tile, with clamped loads (the API requires
// 16-byte-loadable A and B and N, K multiples of 4; split ranges start on a
// stage).
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WM, int WN>
__device__ void f32_tile(float *smem, const Prob &p, const Item &it, float *partials,
                         int *counters, int *s_flag) {
    using G = hbs::F32Cfg<BM, BN, BK, WM, WN>;
    floatx16 acc[G::kFM][G::kFN];
    const int m0 = it.tm * BM, n0 = it.tn * BN, K = it.k1 - it.k0;
    const float *Ag = p.A + (int64_t)m0 * p.lda + it.k0;
    const float *Bg = p.B + (int64_t)it.k0 * p.ldb + n0;
    const hbs::TileArgs t{Ag, Bg, p.M, p.N, K, p.lda, p.ldb, m0, n0};
    typename G::Regs regs;
    hbs::NoChainGlobal ch;  // operands from the problem table: global loads
    hbs::f32_block_mma<BM, BN, BK, WM, WN, hbs::kTileClampK>(smem, t, acc, regs, ch);
    finish_tile<BM, BN, WM, WN, false>(acc, p, it, partials, counters, s_flag);
}

template <int BM, int BN, int WM, int WN>
__device__ void b3_tile(uint16_t *smem, const Prob &p, const Item &it, float *partials,
                        int *counters, int *s_flag) {
    using G = hbs::B3Cfg<BM, BN, 32, WM, WN>;
    floatx16 acc[G::kFM][G::kFN];
    const int m0 = it.tm * BM, n0 = it.tn * BN, K = it.k1 - it.k0;
    const float *Ag = p.A + (int64_t)m0 * p.lda + it.k0;
    const float *Bg = p.B + (int64_t)it.k0 * p.ldb + n0;
    const hbs::TileArgs t{Ag, Bg, p.M, p.N, K, p.lda, p.ldb, m0, n0};
    typename G::Regs regs;
    hbs::NoChainGlobal ch;  // operands from the problem table: global loads
    hbs::bf16x3_block_mma<G, hbs::kTileClampK>(smem, t, acc, regs, ch);
    finish_tile<BM, BN, WM, WN, G::kQ16>(acc, p, it, partials, counters, s_flag);
}

// LDS of the largest configuration of each precision (FP32: 128 x 128 with
// BK = 32, still two workgroups per CU)
constexpr int kLdsF32 = hbs::F32Cfg<128, 128, 32, 2, 2>::kLdsFloats * 4;
constexpr int kLdsB3 = 2 * hbs::B3Cfg<128, 128, 32, 2, 2>::kStage * 2;
static_assert(2 * kLdsF32 <= 160 * 1024 && 2 * kLdsB3 <= 160 * 1024, "two workgroups per CU");

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
            // tile ids of kTileCfg (grouped_plan.h) on this launch
            switch (it.cfg) {
                case 0: f32_tile<128, 128, 16, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 1: f32_tile<64, 64, 16, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 2: f32_tile<128, 64, 16, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 3: f32_tile<64, 128, 16, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 5: f32_tile<128, 128, 16, 4, 1>(s, p, it, partials, counters, &s_flag); break;
                default: f32_tile<128, 128, 32, 2, 2>(s, p, it, partials, counters, &s_flag); break;
            }
        } else {
            uint16_t *s = (uint16_t *)smem;
            switch (it.cfg) {
                case 0: b3_tile<128, 128, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 1: b3_tile<64, 64, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                case 2: b3_tile<128, 64, 2, 2>(s, p, it, partials, counters, &s_flag); break;
                default: b3_tile<64, 128, 2, 2>(s, p, it, partials, counters, &s_flag); break;
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
// Large-tile launch: one 512-thread workgroup per CU takes BM x BN tiles (8
// waves as WM x WN) from its XCD's queue and runs them with the square
// kernels' tile code; the default 256 x 256 tile spreads the valid part of an
// edge tile over the SIMDs.  A problem on this path is covered whole: tiles
// at its ragged bottom and right edges read clamped rows and columns
// (kTileClamp) and store only what lies inside M x N.  No split-K.  MODE
// kTileClamp takes problems whose K is a multiple of the 32-deep stage;
// kTileClampK (a launch of its own) the others, zero-filling past K in the
// last stage; kTileEdge (a launch of its own, 256 x 256 only) problems whose A
// or B is not 16-byte aligned or whose N, K, lda or ldb is not a multiple of 4,
// with per-element predicated loads.  One kernel holding two modes ran out of
// registers and cost the first kind 5%.
// ---------------------------------------------------------------------------
template <int PREC, int MODE, int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(kBigThreads) void grouped_big_kernel(const Prob *__restrict__ probs,
                                                                  const Item *__restrict__ items,
                                                                  const int *__restrict__ qoff,
                                                                  int nq, int *queue) {
    constexpr bool kF32 = PREC == HBS_PREC_FP32;
    static_assert(MODE == hbs::kTileClamp || MODE == hbs::kTileClampK || MODE == hbs::kTileEdge,
                  "mode");
    using GF = hbs::F32Cfg<BM, BN, 32, WM, WN>;
    // BF16x3 on 16 x 16 MFMAs (B3Cfg: more flops per joule at the power limit)
    using GB = hbs::B3Cfg<BM, BN, 32, WM, WN, true>;
    static_assert(GF::kThreads == kBigThreads && GB::kThreads == kBigThreads, "8 waves");
    constexpr int LDS = kF32 ? GF::kLdsFloats * 4 : 2 * GB::kStage * 2;
    using Regs = std::conditional_t<kF32, typename GF::Regs, typename GB::Regs>;
    static_assert(GF::kFM == GB::kFM && GF::kFN == GB::kFN, "accumulators");
    __shared__ __attribute__((aligned(16))) char smem[LDS];
    __shared__ int s_next[2];
    // queue q = blockIdx.x % nq (the workgroup's XCD when nq = 8) hands out
    // items[qoff[q] .. qoff[q + 1]) in order; queue[q] is its head and
    // queue[kXcdQueues] counts the workgroups that have left.  A workgroup
    // whose own queue is empty takes the next item of the other queues in
    // turn (q + 1, q + 2, ...), so no XCD idles while another still has tiles
    // (per-workgroup end stamps showed XCDs finishing up to 0.6 ms apart
    // without it); `total` means no item is left anywhere.
    const int q = blockIdx.x % nq, total = qoff[nq];
    auto fetch = [&]() {
        for (int d = 0; d < nq; ++d) {
            const int qq = q + d < nq ? q + d : q + d - nq;
            const int i = qoff[qq] + __hip_atomic_fetch_add(&queue[qq], 1, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
            if (i < qoff[qq + 1]) return i;
        }
        return total;
    };
    auto tile = [&](int i, hbs::TileArgs &t, float *&Cg, int &ldc) {
        const Item it = items[i];
        const Prob p = probs[it.prob];
        const int m0 = it.tm * BM, n0 = it.tn * BN;
        t = hbs::TileArgs{p.A + (int64_t)m0 * p.lda, p.B + n0, p.M, p.N, p.K, p.lda, p.ldb, m0, n0};
        Cg = p.C;
        ldc = p.ldc;
    };
    if (threadIdx.x == 0) s_next[0] = fetch();
    __syncthreads();
    const int i0 = __builtin_amdgcn_readfirstlane(s_next[0]);
    if (i0 < total) {
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
            if (j >= total) return false;
            tile(j, nt, nCg, nldc);
            return true;
        };
        auto ch = hbs::make_chain(publish, peek);
        Regs regs;
        floatx16 acc[GF::kFM][GF::kFN];
        for (;;) {
            if (threadIdx.x == 0) nxt = fetch();
            if constexpr (kF32) {
                hbs::f32_block_mma<BM, BN, 32, WM, WN, MODE>((float *)smem, t, acc, regs, ch);
                hbs::f32_block_store<BM, BN, 32, WM, WN, hbs::kTileClamp>(Cg, t.M, t.N, ldc, t.m0,
                                                                          t.n0, acc);
            } else {
                hbs::bf16x3_block_mma<GB, MODE>((uint16_t *)smem, t, acc, regs, ch);
                hbs::bf16x3_block_store<GB, hbs::kTileClamp>(Cg, t.M, t.N, ldc, t.m0, t.n0, acc);
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

// Held while a graph is captured and while a plan's device memory is
// allocated, filled or freed.  A device-wide synchronisation (which hipFree
// and hipMalloc may do) in one host thread while another thread's stream is
// capturing fails ("operation not permitted when stream is capturing") and
// invalidates that capture; a concurrency test met exactly this with plans
// created in some threads while others captured their graphs.
std::mutex capture_mu;

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
    std::lock_guard<std::mutex> cap(capture_mu);
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
    // upload on a private stream and wait for that stream only (no device-
    // wide synchronisation; this stream is never captured)
    hipStream_t up = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dmem, probs.data(), probs.size() * sizeof(Prob),
                           hipMemcpyHostToDevice, up);
    if (e == hipSuccess && !sorted.empty())
        e = hipMemcpyAsync(dmem + o_items, sorted.data(), sorted.size() * sizeof(Item),
                           hipMemcpyHostToDevice, up);
    for (int g = 0; g < kBigLaunches; ++g) {
        const std::vector<Item> &sb = hp.sorted_big[g];
        const std::vector<int> &qo = hp.qoff_big[g];
        if (e == hipSuccess && !sb.empty())
            e = hipMemcpyAsync(dmem + o_items_big[g], sb.data(), sb.size() * sizeof(Item),
                               hipMemcpyHostToDevice, up);
        if (e == hipSuccess && !qo.empty())
            e = hipMemcpyAsync(dmem + o_qoff_big[g], qo.data(), qo.size() * sizeof(int),
                               hipMemcpyHostToDevice, up);
    }
    if (e == hipSuccess) e = hipMemsetAsync(plan->d_counters, 0, b_cnt, up);
    if (up) {
        const hipError_t es = hipStreamSynchronize(up);
        if (e == hipSuccess) e = es;
        (void)hipStreamDestroy(up);
    }
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

extern "C" int hbs_grouped_plan_info(const hbs_grouped_plan *plan, hbs_grouped_info *info) {
    if (!plan || !info) return HBS_ERR_NULL;
    info->problems = plan->count;
    info->workgroups = plan->nblocks;
    info->work_items = plan->items;
    for (int g = 0; g < kBigLaunches; ++g) {
        info->workgroups += plan->nblocks_big[g];
        info->work_items += plan->items_big[g];
    }
    info->tiles = plan->tiles;
    info->split_tiles = plan->split_tiles;
    info->est_imbalance = plan->est_imbalance;
    return HBS_SUCCESS;
}

template <int V>
using IC = std::integral_constant<int, V>;

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
        // (tile id, mode) of launch g: kTileCfg / big_launch_* in grouped_plan.h
        auto by_mode = [&](auto prec, auto bm, auto bn, auto wm, auto wn) {
            constexpr int P = decltype(prec)::value, BM = decltype(bm)::value,
                          BN = decltype(bn)::value, WM = decltype(wm)::value,
                          WN = decltype(wn)::value;
            const int mode = big_launch_mode(g);
            if (mode == kBigClamp) launch(grouped_big_kernel<P, hbs::kTileClamp, BM, BN, WM, WN>);
            else if (mode == kBigClampK)
                launch(grouped_big_kernel<P, hbs::kTileClampK, BM, BN, WM, WN>);
            else if constexpr (BM == 256 && BN == 256)
                launch(grouped_big_kernel<P, hbs::kTileEdge, BM, BN, WM, WN>);
        };
        auto by_tile = [&](auto prec) {
            switch (big_launch_tile(g)) {
                case 7: by_mode(prec, IC<128>{}, IC<128>{}, IC<4>{}, IC<2>{}); break;
                case 8: by_mode(prec, IC<256>{}, IC<128>{}, IC<4>{}, IC<2>{}); break;
                case 9: by_mode(prec, IC<128>{}, IC<256>{}, IC<2>{}, IC<4>{}); break;
                default: by_mode(prec, IC<256>{}, IC<256>{}, IC<4>{}, IC<2>{}); break;
            }
        };
        if (fp32)
            by_tile(IC<HBS_PREC_FP32>{});
        else
            by_tile(IC<HBS_PREC_BF16X3>{});
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
    std::lock_guard<std::mutex> cap(capture_mu);
    hipError_t e = hipFree(plan->dmem);
    delete plan;
    return e == hipSuccess ? HBS_SUCCESS : (int)e;
}

// hipGraph helpers: capture one grouped run into an executable graph.
extern "C" int hbs_grouped_graph_create(const hbs_grouped_plan *plan, void **graph_exec) {
    if (!plan || !graph_exec) return HBS_ERR_NULL;
    *graph_exec = nullptr;
    std::lock_guard<std::mutex> cap(capture_mu);
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
