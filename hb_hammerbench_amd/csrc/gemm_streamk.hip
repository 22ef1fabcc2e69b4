// gemm_streamk.hip -- stream-K FP32-MFMA / BF16x3 GEMM for grids of output
// tiles that do not fill the GPU.
//
// No reference counterpart: a scheduling variant of apps/sgemm's C = A.B
// (kernel.cpp:279-333) for products with fewer 256 x 256 tiles than CUs
// (M = N = 1536 has 36 on 256 CUs, 3072 has 144), where the
// one-tile-per-workgroup kernels leave CUs idle.  It pays when K is deep
// enough that the recovered CUs outweigh summing the split tiles.
//
// Design:
//  * G 512-thread workgroups, at most one per CU, over the flattened space
//    of U units (tile, 32-deep K stage), tiles in the square kernels'
//    grouped raster; workgroup w takes the contiguous units
//    [w U / G, (w + 1) U / G), so every workgroup gets the same work to
//    within one stage; G trades idle CUs against the cost of summing split
//    tiles (make_plan);
//  * a tile whose stages one workgroup covers is computed and stored as in
//    the square kernels (the same main loops and epilogues, tile_f32.h and
//    tile_bf16x3.h, over a K range);
//  * a tile split over several workgroups: each writes its partial sums to a
//    slot of the workspace and increments the tile's counter; the last to
//    arrive adds the partials in K order (so the result does not depend on
//    which arrives last) and stores C, then resets the counter.  Nothing ever
//    waits for another workgroup, so the launch cannot deadlock whatever else
//    shares the GPU.
//  * A workgroup's range starts inside at most one tile and ends inside at
//    most one, so two partial slots per workgroup suffice: slot 2w for a
//    segment that ends its tile, 2w + 1 for one that does not.
#include <algorithm>

#include "tile_bf16x3.h"
#include "tile_f32.h"

namespace {

constexpr int kBM = 256, kBN = 256, kBK = 32, kWM = 4, kWN = 2, kThreads = 512;
using GF = hbs::F32Cfg<kBM, kBN, kBK, kWM, kWN>;
using GB = hbs::B3Main16;  // BF16x3 on 16 x 16 MFMAs (more flops per joule at the power limit)
static_assert(GF::kThreads == kThreads && GB::kThreads == kThreads, "8 waves");
static_assert(GF::kFM == GB::kFM && GF::kFN == GB::kFN, "accumulators");
constexpr int kFM = GF::kFM, kFN = GF::kFN;
constexpr int64_t kTileFloats = (int64_t)kBM * kBN;
constexpr int kWaveFloats = kFM * kFN * 16 * 64;

// Workspace: the tile counters first (a fixed region, so that calls of any
// shape find them where the previous call left them, all zero), then the
// 2 G partial slots.
constexpr int64_t kMaxTiles = 65536;
// Partials the last workgroup of a split tile loads together (per 32 x 32
// block of accumulators) before adding them in K order.
constexpr int kMaxSplit = 4;
constexpr size_t kCounterBytes = kMaxTiles * sizeof(int);

struct Args {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    int tiles_m, tiles_n, iters;  // iters: K stages per tile
    int64_t units;                // tiles * iters
    int grid;                     // G
    int remap;                    // give each XCD a contiguous range of workgroups
    float *partials;              // 2 G slots of kTileFloats
    int *counters;                // one per tile, zero at rest
};

// The workgroup whose unit range holds unit x (ranges [w U / G, (w+1) U / G)).
__device__ __forceinline__ int unit_owner(int64_t x, int64_t units, int grid) {
    return (int)(((x + 1) * grid + units - 1) / units - 1);
}

// The calling wave's part of a 256 x 256 tile, as the square kernels' stores
// place it (edge tiles spread their valid part over the SIMDs).
template <int PREC, int MODE>
__device__ __forceinline__ hbs::WaveTile wave_tile_of(int M, int N, int m0, int n0) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if constexpr (PREC == HBS_PREC_FP32)
        return hbs::f32_wave_tile<kBM, kBN, kBK, kWM, kWN, MODE>(M, N, m0, n0, wave);
    else
        return hbs::bf16x3_wave_tile<GB, MODE>(M, N, m0, n0, wave);
}

template <int PREC, int MODE>
__global__ __launch_bounds__(kThreads) void streamk_kernel(Args p) {
    constexpr bool kF32 = PREC == HBS_PREC_FP32;
    constexpr bool kQ16 = !kF32 && GB::kQ16;  // accumulator layout (store_acc)
    constexpr int LDS = kF32 ? GF::kLdsFloats * 4 : 2 * GB::kStage * 2;
    __shared__ __attribute__((aligned(16))) char smem[LDS];
    __shared__ int s_last;
    // Workgroups hold runs in block order.  With few segments per tile each
    // XCD takes a contiguous range of them (xcd_remap), so a split tile's
    // partials are summed from this XCD's L2.  With 8 or more per tile the
    // blocks are left in the hardware's round-robin over the XCDs, which puts
    // on one XCD the workgroups that cover the same K range of neighbouring
    // tiles: they share A and B panels in its L2 (measured +4% FP32, +8-13%
    // BF16x3 at 1024 x 1024 x 32768; the remap is better by 4-15% at
    // M = N = K <= 2048).
    const int w = p.remap ? hbs::xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int64_t u1 = (int64_t)(w + 1) * p.units / p.grid;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t u = (int64_t)w * p.units / p.grid; u < u1;) {
        const int t = (int)(u / p.iters);
        const int kb = (int)(u - (int64_t)t * p.iters);
        const int ke = (int)min((int64_t)p.iters, kb + (u1 - u));
        u += ke - kb;
        int tm, tn;
        hbs::tile_coords(t, p.tiles_m, p.tiles_n, 8, tm, tn);
        const int m0 = tm * kBM, n0 = tn * kBN;
        const int k0 = kb * kBK, k1 = min(ke * kBK, p.K);
        const hbs::TileArgs ta{p.A + (int64_t)m0 * p.lda + k0, p.B + (int64_t)k0 * p.ldb + n0,
                               p.M, p.N, k1 - k0, p.lda, p.ldb, m0, n0};
        floatx16 acc[kFM][kFN];
        hbs::NoChain ch;
        if constexpr (kF32) {
            typename GF::Regs regs;
            hbs::f32_block_mma<kBM, kBN, kBK, kWM, kWN, MODE>((float *)smem, ta, acc, regs, ch);
        } else {
            typename GB::Regs regs;
            hbs::bf16x3_block_mma<GB, MODE>((uint16_t *)smem, ta, acc, regs, ch);
        }
        auto store = [&]() {
            if constexpr (kF32)
                hbs::f32_block_store<kBM, kBN, kBK, kWM, kWN, MODE>(p.C, p.M, p.N, p.ldc, m0, n0,
                                                                    acc);
            else
                hbs::bf16x3_block_store<GB, MODE>(p.C, p.M, p.N, p.ldc, m0, n0, acc);
        };
        if (kb == 0 && ke == p.iters) {
            store();
        } else {
            // the tile's segments are those of workgroups wa .. wb, in K order
            const int wa = unit_owner((int64_t)t * p.iters, p.units, p.grid);
            const int wb = unit_owner((int64_t)(t + 1) * p.iters - 1, p.units, p.grid);
            const int nseg = wb - wa + 1;
            auto slot_of = [&](int ws) { return 2 * ws + (ws == wb ? 0 : 1); };
            // Partials move through device-coherent accesses (agent-scope
            // atomic stores and loads: written through to memory, read past
            // this XCD's L2) ordered by this workgroup's own waits and
            // barriers, instead of agent-scope fences: on this GPU such a
            // fence writes back and invalidates the whole L2 of the XCD,
            // which costs every workgroup there its cached panels.
            float *own = p.partials + slot_of(w) * kTileFloats + wave * kWaveFloats;
#pragma unroll
            for (int r = 0; r < kFM; ++r)
#pragma unroll
                for (int c = 0; c < kFN; ++c)
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        __hip_atomic_store(&own[((r * kFN + c) * 16 + q) * 64 + lane], acc[r][c][q],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's partials are written
            __syncthreads();                     // ... and every thread's
            if (tid == 0)
                s_last = __hip_atomic_fetch_add(&p.counters[t], 1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT) == nseg - 1;
            __syncthreads();
            if (s_last) {
                // one 32 x 32 block at a time, the partials of kMaxSplit
                // segments in flight together, added in K order and stored
                // straight to C: the accumulators are dead here, which keeps
                // the register pressure (and the spills around this code) low
                const hbs::WaveTile wt = wave_tile_of<PREC, MODE>(p.M, p.N, m0, n0);
                const int ncols = wt.kind == hbs::kWaveHalf ? kFN / 2 : kFN;
                const int rb = m0 + wt.om + hbs::acc_lane_row<kQ16>(lane);
                const int cb = n0 + wt.on + hbs::acc_lane_col<kQ16>(lane);
                float *lanep = p.C + (int64_t)rb * p.ldc + cb;
                const int vm = p.M - rb, vn = p.N - cb;  // valid rows / columns from the lane's own
#pragma unroll
                for (int r = 0; r < kFM; ++r) {
#pragma unroll
                    for (int c = 0; c < kFN; ++c) {
                        floatx16 v = floatx16{0};
#pragma unroll 1
                        for (int j0 = 0; j0 < nseg; j0 += kMaxSplit) {
                            floatx16 part[kMaxSplit];
#pragma unroll
                            for (int j = 0; j < kMaxSplit; ++j) {
                                const int jj = min(j0 + j, nseg - 1);
                                const float *src = p.partials + slot_of(wa + jj) * kTileFloats +
                                                   wave * kWaveFloats + (r * kFN + c) * 16 * 64 +
                                                   lane;
#pragma unroll
                                for (int q = 0; q < 16; ++q)
                                    part[j][q] = __hip_atomic_load(&src[q * 64], __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT);
                            }
#pragma unroll
                            for (int j = 0; j < kMaxSplit; ++j)
                                if (j0 + j < nseg) v += part[j];
                        }
                        if (wt.kind == hbs::kWaveNone || c >= ncols) continue;
#pragma unroll
                        for (int q = 0; q < 16; ++q) {
                            const int dr = r * 32 + hbs::acc_row<kQ16>(q);
                            const int dc = c * 32 + hbs::acc_col<kQ16>(q);
                            if (MODE == hbs::kTileInterior || (dr < vm && dc < vn))
                                hbs::stg(lanep + ((int64_t)dr * p.ldc + dc), v[q]);
                        }
                    }
                }
                if (tid == 0)
                    __hip_atomic_store(&p.counters[t], 0, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();  // LDS and s_last are reused by the next segment
    }
}

struct Plan {
    int tiles_m, tiles_n, iters, grid;
    int64_t units;
};

// Fix-up cost in units (one 256 x 256 x 32 stage on one CU), measured on
// MI355X: the last workgroup of a split tile spends about 25 us per
// kMaxSplit partials it reads (the reads come right after every CU's
// partial stores and are latency-bound), against a stage of about 7.7 us in
// FP32 and 2.7 us in BF16x3.
constexpr int kFixupUnits[2] = {3, 9};

// G minimises  ceil(U / G) + fix-up  over 1 <= G <= min(CUs, U): the longest
// run of units plus, when any tile is split, the fix-up of the most-split
// tile (a tile meets at most ceil(iters / run) + 1 runs of floor(U / G)
// units, exactly iters / run when the runs are equal and divide iters).  Ties
// keep the smaller G (less workspace).
Plan make_plan(int precision, int M, int N, int K, int cus) {
    Plan pl;
    pl.tiles_m = (int)hbs::ceil_div(M, kBM);
    pl.tiles_n = (int)hbs::ceil_div(N, kBN);
    pl.iters = (int)hbs::ceil_div(K, kBK);
    pl.units = (int64_t)pl.tiles_m * pl.tiles_n * pl.iters;
    const int64_t fix = kFixupUnits[precision == HBS_PREC_FP32 ? 0 : 1];
    const int gmax = (int)std::min<int64_t>(cus, pl.units);
    int64_t best = -1;
    pl.grid = 1;
    for (int g = 1; g <= gmax; ++g) {
        const int64_t run = pl.units / g, longest = hbs::ceil_div(pl.units, (int64_t)g);
        const bool even = pl.units % g == 0;
        const bool split = !even || run % pl.iters != 0;
        const int64_t segs = even && pl.iters % run == 0 ? pl.iters / run
                                                         : hbs::ceil_div((int64_t)pl.iters, run) + 1;
        const int64_t cost = longest + (split ? fix * hbs::ceil_div(segs, (int64_t)kMaxSplit) : 0);
        if (best < 0 || cost < best) { best = cost; pl.grid = g; }
    }
    return pl;
}

size_t workspace_bytes(const Plan &pl) {
    return kCounterBytes + (size_t)2 * pl.grid * kTileFloats * sizeof(float);
}

template <int PREC>
int launch(const Args &a, int mode, hipStream_t s) {
    const dim3 grid(a.grid);
    switch (mode) {
        case hbs::kTileInterior:
            streamk_kernel<PREC, hbs::kTileInterior><<<grid, kThreads, 0, s>>>(a);
            break;
        case hbs::kTileClamp: streamk_kernel<PREC, hbs::kTileClamp><<<grid, kThreads, 0, s>>>(a); break;
        case hbs::kTileClampK:
            streamk_kernel<PREC, hbs::kTileClampK><<<grid, kThreads, 0, s>>>(a);
            break;
        default: streamk_kernel<PREC, hbs::kTileEdge><<<grid, kThreads, 0, s>>>(a); break;
    }
    return hbs::launch_status();
}

}  // namespace

extern "C" size_t hbs_gemm_streamk_workspace_size(int precision, int M, int N, int K) {
    if (precision != HBS_PREC_FP32 && precision != HBS_PREC_BF16X3) return 0;
    if (M < 1 || N < 1 || K < 1 || M > hbs::kMaxDim || N > hbs::kMaxDim || K > hbs::kMaxDim)
        return 0;
    if ((int64_t)hbs::ceil_div(M, kBM) * hbs::ceil_div(N, kBN) > kMaxTiles) return 0;
    const Plan pl = make_plan(precision, M, N, K, hbs::cu_count());
    return workspace_bytes(pl);
}

extern "C" int hbs_gemm_streamk(int precision, const float *A, const float *B, float *C, int M,
                                int N, int K, int lda, int ldb, int ldc, void *workspace,
                                size_t workspace_bytes_given, void *stream) {
    if (precision != HBS_PREC_FP32 && precision != HBS_PREC_BF16X3) return HBS_ERR_ARG;
    if (M < 1 || N < 1 || K < 1 || lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(M, N, K, lda, ldb, ldc)) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    if ((int64_t)hbs::ceil_div(M, kBM) * hbs::ceil_div(N, kBN) > kMaxTiles) return HBS_ERR_ARG;
    const Plan pl = make_plan(precision, M, N, K, hbs::cu_count());
    if (!workspace || workspace_bytes_given < workspace_bytes(pl)) return HBS_ERR_WORKSPACE;
    if (!hbs::aligned16(workspace)) return HBS_ERR_ALIGN;
    const int st = hbs::check_device();
    if (st) return st;
    // tile modes as in the square kernels: interior when every tile is whole
    // and the operands allow 16-byte loads, clamped loads for ragged M or N,
    // zero fill past K when K is not a multiple of the stage, else
    // per-element predicates (FP32 loads B in 16-byte rows too; BF16x3 loads
    // B by columns and needs only A aligned)
    const bool fp32 = precision == HBS_PREC_FP32;
    const bool vec = hbs::aligned16(A) && lda % 4 == 0 &&
                     (!fp32 || (hbs::aligned16(B) && ldb % 4 == 0 && N % 4 == 0));
    int mode = hbs::kTileEdge;
    if (vec && K % kBK == 0 && M % kBM == 0 && N % kBN == 0) mode = hbs::kTileInterior;
    else if (vec && K % kBK == 0) mode = hbs::kTileClamp;
    else if (vec && K % 4 == 0 && (!fp32 || N % 4 == 0)) mode = hbs::kTileClampK;
    char *ws = (char *)workspace;
    const int remap = pl.grid < 8 * pl.tiles_m * pl.tiles_n;
    const Args a{A, B, C, M, N, K, lda, ldb, ldc, pl.tiles_m, pl.tiles_n, pl.iters, pl.units,
                 pl.grid, remap, (float *)(ws + kCounterBytes), (int *)ws};
    return fp32 ? launch<HBS_PREC_FP32>(a, mode, (hipStream_t)stream)
                : launch<HBS_PREC_BF16X3>(a, mode, (hipStream_t)stream);
}
