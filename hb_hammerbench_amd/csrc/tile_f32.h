// tile_f32.h -- one FP32-MFMA output tile, shared by the square / strided
// SGEMM kernel (sgemm_f32.hip) and the grouped kernel's large-tile launch.
//
// A workgroup of WM x WN waves computes the BM x BN tile at (m0, n0) of
// C = A . B over the whole K:
//   * global -> registers -> LDS double buffer, 128-bit coalesced loads; each
//     K stage is loaded a full step ahead and stored to LDS in the middle of
//     the previous step;
//   * each wave owns a (BM/WM) x (BN/WN) sub-tile held as 32x32 accumulator
//     blocks of v_mfma_f32_32x32x2_f32 (FP32 in, FP32 accumulate);
//   * an A fragment for four consecutive MFMAs comes from one 128-bit LDS read
//     (lane half h takes k = 4h..4h+3 of each 8-wide k chunk).
#pragma once

#include <type_traits>

#include "hbs_internal.h"

namespace hbs {

template <int BM, int BN, int BK, int WM, int WN>
struct F32Cfg {
    static constexpr int kThreads = WM * WN * 64;
    static constexpr int kWTM = BM / WM;  // wave tile rows
    static constexpr int kWTN = BN / WN;  // wave tile cols
    static constexpr int kFM = kWTM / 32; // 32x32 blocks per wave (rows)
    static constexpr int kFN = kWTN / 32; // 32x32 blocks per wave (cols)
    static constexpr int kLdA = BK + 4;   // LDS row stride of the [BM][BK] A tile
    static constexpr int kLdB = BN + 8;   // LDS row stride of the [BK][BN] B tile
    static constexpr int kAF4 = (BM * BK / 4) / kThreads;
    static constexpr int kBF4 = (BK * BN / 4) / kThreads;
    static constexpr int kStageA = BM * kLdA;
    static constexpr int kStageB = BK * kLdB;
    static constexpr int kLdsFloats = 2 * (kStageA + kStageB);
    // one stage in registers between its global loads and its store to LDS
    struct Regs {
        floatx4 a[kAF4], b[kBF4];
    };
    static_assert(kWTM % 32 == 0 && kWTN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert((BM * BK / 4) % kThreads == 0, "A tile must split evenly over threads");
    static_assert((BK * BN / 4) % kThreads == 0, "B tile must split evenly over threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16 (two or more 8-k chunks)");
    // Sub-tile (wm, wn) of wave `wave`.  With 8 waves as 4 x 2, waves w and
    // w + 4 share a SIMD (read from HW_ID on MI355X: the four SIMDs take waves
    // 0..3 in some order, then 4..7 in the same order).  Rotating the rows of
    // the second wave column by half the tile gives those two waves different
    // row halves and different column halves, so that in an edge tile with
    // only its top rows or left columns inside M x N, where the MFMAs of
    // waves wholly outside are skipped, every SIMD keeps work.  Other shapes
    // use wm = wave % WM, wn = wave / WM.
    __device__ static void wave_pos(int wave, int &wm, int &wn) {
        if (WM == 4 && WN == 2) {
            wn = wave >> 2;
            wm = (wave + 2 * wn) & 3;
        } else {
            wn = wave / WM;
            wm = wave % WM;
        }
    }
};

// Main loop: acc (this lane's kFM x kFN 32x32 blocks of the wave's sub-tile)
// = A[m0.., 0..K) . B[0..K, n0..), with Ag = A + m0 * lda (first row of the
// tile) and Bg = B + n0 (first column); a K range [k0, k1) of a larger
// product is Ag + k0, Bg + k0 * ldb and K = k1 - k0.  M, N, K bound the
// predicates.  smem holds F32Cfg::kLdsFloats floats.  Every thread of the
// workgroup must call it.  MODE (TileMode): kTileInterior, kTileEdge (any
// shape or alignment), kTileClamp (ragged M or N; K a multiple of BK, N of 4,
// 16-byte loadable operands) or kTileClampK (the same with K a multiple of 4).
// regs is the stage in flight.  ch: NoChain, or a TileChain to pipeline the
// next tile's first stages into this tile's last steps.  The wave computes the
// sub-tile at (om, on) of the tile: kFM x kFN blocks (CK = kWaveFull),
// kFM x kFN / 2 blocks (kWaveHalf) or none (kWaveNone: it still loads and
// stores its share of each stage and meets every barrier).
template <int BM, int BN, int BK, int WM, int WN, int MODE, int CK, class Chain>
__device__ __forceinline__ void f32_block_mma_impl(
    float *smem, const float *__restrict__ Ag, const float *__restrict__ Bg, int M, int N, int K,
    int lda, int ldb, int m0, int n0, int om, int on,
    floatx16 (&acc)[F32Cfg<BM, BN, BK, WM, WN>::kFM][F32Cfg<BM, BN, BK, WM, WN>::kFN],
    typename F32Cfg<BM, BN, BK, WM, WN>::Regs &regs, Chain &ch) {
    using G = F32Cfg<BM, BN, BK, WM, WN>;
    constexpr int kFNc = CK == kWaveHalf ? G::kFN / 2 : G::kFN;  // column blocks computed
    float *const As = smem;
    float *const Bs = smem + 2 * G::kStageA;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int l32 = lane & 31, h = lane >> 5;

#pragma unroll
    for (int r = 0; r < G::kFM; ++r)
#pragma unroll
        for (int c = 0; c < G::kFN; ++c) acc[r][c] = floatx16{0};

    // the stage in flight: the caller's registers when chained (they carry
    // the next tile's stage 1 across the call), else two local arrays
    floatx4 own_a[G::kAF4], own_b[G::kBF4];
    auto &ra = *(Chain::kOn ? &regs.a : &own_a);
    auto &rb = *(Chain::kOn ? &regs.b : &own_b);

    // Operands of the global loads: this tile's, until a chained loop points
    // them at the next tile once this tile's last stage has been loaded.
    const float *lA = Ag, *lB = Bg;
    int lM = M, lN = N, lK = K, llda = lda, lldb = ldb, lm0 = m0, ln0 = n0;
    auto retarget = [&](const TileArgs &n) {
        lA = n.Ag;
        lB = n.Bg;
        lM = n.M;
        lN = n.N;
        lK = n.K;
        llda = n.lda;
        lldb = n.ldb;
        lm0 = n.m0;
        ln0 = n.n0;
    };
    constexpr bool kG = Chain::kGlobal;
    auto gload_m = [&](auto mode, int k0) {
        constexpr int LM = decltype(mode)::value;
#pragma unroll
        for (int i = 0; i < G::kAF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int row = f / (BK / 4), c4 = f % (BK / 4);
            if constexpr (LM != kTileEdge) {
                const int srow = LM >= kTileClamp ? min(row, lM - lm0 - 1) : row;
                const floatx4 zero = {0, 0, 0, 0};
                ra[i] = (LM != kTileClampK || k0 + c4 * 4 < lK)
                            ? ld<kG>((const floatx4 *)(lA + (int64_t)srow * llda + k0 + c4 * 4))
                            : zero;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gk = k0 + c4 * 4 + j;
                    ra[i][j] =
                        (lm0 + row < lM && gk < lK) ? ld<kG>(lA + (int64_t)row * llda + gk) : 0.0f;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < G::kBF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int kr = f / (BN / 4), c4 = f % (BN / 4);
            if constexpr (LM != kTileEdge) {
                const int sc = LM >= kTileClamp ? min(c4 * 4, lN - ln0 - 4) : c4 * 4;
                const floatx4 zero = {0, 0, 0, 0};
                rb[i] = (LM != kTileClampK || k0 + kr < lK)
                            ? ld<kG>((const floatx4 *)(lB + (int64_t)(k0 + kr) * lldb + sc))
                            : zero;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gn = ln0 + c4 * 4 + j;
                    rb[i][j] = (k0 + kr < lK && gn < lN)
                                   ? ld<kG>(lB + (int64_t)(k0 + kr) * lldb + c4 * 4 + j)
                                   : 0.0f;
                }
            }
        }
    };
    // With kTileClampK only the stage that runs past K needs K predicates;
    // every full stage takes the clamped loads (the choice is uniform).
    auto gload = [&](int k0) {
        if (MODE == kTileClampK && k0 + BK > lK)
            gload_m(std::integral_constant<int, kTileClampK>{}, k0);
        else
            gload_m(std::integral_constant<int, MODE == kTileClampK ? kTileClamp : MODE>{}, k0);
    };

    auto sstore = [&](int buf) {
        float *as = As + buf * G::kStageA;
        float *bs = Bs + buf * G::kStageB;
#pragma unroll
        for (int i = 0; i < G::kAF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int row = f / (BK / 4), c4 = f % (BK / 4);
            *(floatx4 *)(as + row * G::kLdA + c4 * 4) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < G::kBF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int kr = f / (BN / 4), c4 = f % (BN / 4);
            *(floatx4 *)(bs + kr * G::kLdB + c4 * 4) = rb[i];
        }
    };

    // Fragments for one 8-wide k chunk: A from one 128-bit read per 32-row
    // block, B as four 32-bit reads per 32-col block.  The reads for chunk
    // kc + 1 are issued before the MFMAs of chunk kc so LDS latency overlaps
    // matrix-core work.
    auto load_frag = [&](const float *as, const float *bs, int kc, floatx4 (&a)[G::kFM],
                         float (&b)[kFNc][4]) {
#pragma unroll
        for (int r = 0; r < G::kFM; ++r) a[r] = *(const floatx4 *)(as + r * 32 * G::kLdA + kc * 8);
#pragma unroll
        for (int c = 0; c < kFNc; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) b[c][q] = bs[(kc * 8 + q) * G::kLdB + c * 32];
    };
    auto mfma_chunk = [&](const floatx4 (&a)[G::kFM], const float (&b)[kFNc][4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < G::kFM; ++r)
#pragma unroll
                for (int c = 0; c < kFNc; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[r][q], b[c][q], acc[r][c], 0, 0, 0);
    };
    // Scheduling hint: spread the next chunk's LDS reads (kFM 128-bit and
    // 2 kFN paired 32-bit reads) between this chunk's 4 kFM kFN MFMAs.
    auto interleave = [&]() {
        constexpr int kReads = G::kFM + 2 * kFNc;
        constexpr int kMfma = 4 * G::kFM * kFNc;
        constexpr int kPer = kMfma / kReads > 0 ? kMfma / kReads : 1;
#pragma unroll
        for (int i = 0; i < kReads; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);     // one LDS read
            __builtin_amdgcn_sched_group_barrier(0x008, kPer, 0);  // then MFMAs
        }
    };
    // One stage (BK/8 chunks of 8 k) from LDS buffer `buf`, the fragments of
    // chunk kc + 1 requested before the MFMAs of chunk kc.  `hook` runs after
    // the MFMAs of chunk kHook - 1, once the fragments of chunk kHook are in
    // flight, so the matrix core is never left waiting on a fresh LDS read.
    constexpr int kChunks = BK / 8, kHook = kChunks / 2;
    auto compute = [&](int buf, auto &&hook) {
        if constexpr (CK == kWaveNone) {
            hook();
            return;
        }
        const float *as = As + buf * G::kStageA + (om + l32) * G::kLdA + 4 * h;
        const float *bs = Bs + buf * G::kStageB + 4 * h * G::kLdB + on + l32;
        floatx4 a0[G::kFM], a1[G::kFM];
        float b0[kFNc][4], b1[kFNc][4];
        load_frag(as, bs, 0, a0, b0);
#pragma unroll
        for (int kc = 0; kc < kChunks; kc += 2) {
            if (kc + 1 < kChunks) load_frag(as, bs, kc + 1, a1, b1);
            mfma_chunk(a0, b0);
            interleave();
            if (kc + 1 == kHook) hook();
            if (kc + 2 < kChunks) load_frag(as, bs, kc + 2, a0, b0);
            if (kc + 1 < kChunks) mfma_chunk(a1, b1);
            interleave();
            if (kc + 2 == kHook) hook();
        }
    };

    // Pipeline: stage kt + 1 is stored to the other LDS buffer in the middle
    // of step kt and the global loads of stage kt + 2 are issued right after,
    // so every load has a whole step to arrive and the store overlaps MFMAs.
    // Stage kt is in LDS buffer (kt + b0) & 1.
    const int KT = (K + BK - 1) / BK;
    int b0 = 0;
    bool pre = false;
    if constexpr (Chain::kOn) {
        b0 = ch.buf0;
        pre = ch.pre;
    }
    if (!pre) {
        gload(0);
        sstore(b0);
        if (KT > 1) gload(BK);
        __syncthreads();
    }
    int kt = 0;
    for (; kt + 2 < KT; ++kt) {
        const int cur = (kt + b0) & 1;
        compute(cur, [&]() {
            sstore(cur ^ 1);
            gload((kt + 2) * BK);
        });
        if constexpr (Chain::kOn) {
            if (kt == 0) ch.publish();
        }
        __syncthreads();
    }
    bool chain = false;
    if constexpr (Chain::kOn) {
        if (KT >= 3) {
            ch.has_next = ch.peek(ch.next);
            chain = ch.has_next;
        }
    }
    if (kt + 1 < KT) {
        const int cur = (kt + b0) & 1;
        compute(cur, [&]() {
            sstore(cur ^ 1);
            if constexpr (Chain::kOn) {
                if (chain) {
                    retarget(ch.next);
                    gload(0);
                }
            }
        });
        __syncthreads();
        ++kt;
    }
    const int cur = (kt + b0) & 1;
    compute(cur, [&]() {
        if constexpr (Chain::kOn) {
            if (chain) {
                sstore(cur ^ 1);
                if (ch.next.K > BK) gload(BK);
            }
        }
    });
    if constexpr (Chain::kOn) {
        if (chain) {
            __syncthreads();
        } else if (KT < 3) {
            ch.publish();
            __syncthreads();
            ch.has_next = ch.peek(ch.next);
        }
        ch.pre = chain;
        ch.buf0 = cur ^ 1;
    }
}

// The sub-tile of wave `wave` in the tile at (m0, n0): for the 256 x 256 tile
// of 4 x 2 waves outside kTileInterior, the edge-tile arrangement
// (big_wave_tile, so that the valid units are spread evenly over the SIMDs);
// else wave_pos.
template <int BM, int BN, int BK, int WM, int WN, int MODE>
__device__ __forceinline__ WaveTile f32_wave_tile(int M, int N, int m0, int n0, int wave) {
    using G = F32Cfg<BM, BN, BK, WM, WN>;
    if constexpr (MODE != kTileInterior && BM == 256 && BN == 256 && WM == 4 && WN == 2) {
        return big_wave_tile(M, N, m0, n0, wave);
    } else {
        int wm, wn;
        G::wave_pos(wave, wm, wn);
        return WaveTile{wm * G::kWTM, wn * G::kWTN, kWaveFull};
    }
}

// One tile's main loop; each wave runs the loop copy of its sub-tile kind
// (whole functions, so that the main loop stays one scheduling region).
template <int BM, int BN, int BK, int WM, int WN, int MODE, class Chain>
__device__ __forceinline__ void f32_block_mma(
    float *smem, const TileArgs &t,
    floatx16 (&acc)[F32Cfg<BM, BN, BK, WM, WN>::kFM][F32Cfg<BM, BN, BK, WM, WN>::kFN],
    typename F32Cfg<BM, BN, BK, WM, WN>::Regs &regs, Chain &ch) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const WaveTile wt = f32_wave_tile<BM, BN, BK, WM, WN, MODE>(t.M, t.N, t.m0, t.n0, wave);
    if (wt.kind == kWaveFull)
        f32_block_mma_impl<BM, BN, BK, WM, WN, MODE, kWaveFull>(
            smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda, t.ldb, t.m0, t.n0, wt.om, wt.on, acc, regs, ch);
    else if constexpr (MODE != kTileInterior && BM == 256 && BN == 256) {
        if (wt.kind == kWaveHalf)
            f32_block_mma_impl<BM, BN, BK, WM, WN, MODE, kWaveHalf>(
                smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda, t.ldb, t.m0, t.n0, wt.om, wt.on, acc, regs,
                ch);
        else
            f32_block_mma_impl<BM, BN, BK, WM, WN, MODE, kWaveNone>(
                smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda, t.ldb, t.m0, t.n0, wt.om, wt.on, acc, regs,
                ch);
    }
}
template <int BM, int BN, int BK, int WM, int WN, int MODE>
__device__ __forceinline__ void f32_block_mma(
    float *smem, const float *__restrict__ Ag, const float *__restrict__ Bg, int M, int N, int K,
    int lda, int ldb, int m0, int n0,
    floatx16 (&acc)[F32Cfg<BM, BN, BK, WM, WN>::kFM][F32Cfg<BM, BN, BK, WM, WN>::kFN]) {
    const TileArgs t{Ag, Bg, M, N, K, lda, ldb, m0, n0};
    typename F32Cfg<BM, BN, BK, WM, WN>::Regs regs;
    NoChain ch;
    f32_block_mma<BM, BN, BK, WM, WN, MODE>(smem, t, acc, regs, ch);
}

// Epilogue: stores acc to C (Cg = the matrix) at the tile (m0, n0), each wave
// its sub-tile (f32_wave_tile).  32x32 accumulator layout: register q of lane
// l holds row 8*(q/4) + 4*(l/32) + q%4, col l%32 of its block.  Outside
// kTileInterior the stores of a tile that crosses M or N are predicated.
template <int BM, int BN, int BK, int WM, int WN, int MODE>
__device__ __forceinline__ void f32_block_store(
    float *__restrict__ Cg, int M, int N, int ldc, int m0, int n0,
    const floatx16 (&acc)[F32Cfg<BM, BN, BK, WM, WN>::kFM][F32Cfg<BM, BN, BK, WM, WN>::kFN]) {
    using G = F32Cfg<BM, BN, BK, WM, WN>;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const WaveTile wt = f32_wave_tile<BM, BN, BK, WM, WN, MODE>(M, N, m0, n0, wave);
    if (wt.kind == kWaveNone) return;
    const int ncols = wt.kind == kWaveHalf ? G::kFN / 2 : G::kFN;
    if (MODE == kTileInterior || (m0 + BM <= M && n0 + BN <= N))
        store_acc<G::kFM, G::kFN, false>(Cg, ldc, m0 + wt.om, n0 + wt.on, M, N, ncols, acc);
    else
        store_acc<G::kFM, G::kFN, true>(Cg, ldc, m0 + wt.om, n0 + wt.on, M, N, ncols, acc);
}

// One whole output tile: main loop and store.
template <int BM, int BN, int BK, int WM, int WN, int MODE>
__device__ __forceinline__ void f32_block_tile(float *smem, const float *__restrict__ Ag,
                                               const float *__restrict__ Bg,
                                               float *__restrict__ Cg, int M, int N, int K,
                                               int lda, int ldb, int ldc, int m0, int n0) {
    using G = F32Cfg<BM, BN, BK, WM, WN>;
    floatx16 acc[G::kFM][G::kFN];
    f32_block_mma<BM, BN, BK, WM, WN, MODE>(smem, Ag, Bg, M, N, K, lda, ldb, m0, n0, acc);
    f32_block_store<BM, BN, BK, WM, WN, MODE>(Cg, M, N, ldc, m0, n0, acc);
}

}  // namespace hbs
