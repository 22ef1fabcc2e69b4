// tile_bf16x3.h -- one BF16x3 output tile, shared by the square / rectangular
// BF16x3 kernel (sgemm_bf16x3.hip) and the grouped kernel's large-tile launch.
//
// Each float is split x = hi + lo, hi = bf16_rne(x), lo = bf16_rne(x - hi)
// (x - hi is exact), and C = Alo.Bhi + Ahi.Blo + Ahi.Bhi is accumulated in
// FP32 by v_mfma_f32_32x32x16_bf16 or, in configurations with Q16 set, by
// v_mfma_f32_16x16x32_bf16; the lo.lo term (relative size 2^-16) is dropped.
// The split happens in registers on the way to LDS.  Accumulators are handed
// out as one floatx16 per 32 x 32 region, in the layout store_acc names by
// the same Q16 flag.
#pragma once

#include <type_traits>

#include "hbs_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace hbs {

// Q16_: the MFMA shape.  On random data a loop of v_mfma_f32_16x16x32_bf16
// sustains 16% more BF16 flops than one of v_mfma_f32_32x32x16_bf16 at the
// board's power limit (1960 against 1685 TFLOP/s; both about 2490 on zeros),
// so the 16 x 16 form pays in these power-limited kernels: the grouped
// large-tile launch +3.7%, and with the pinned schedule of the main loop
// (q16_schedule) the square kernel +4.9% and stream-K +9% over 32 x 32.
template <int BM_, int BN_, int BK_, int WM_, int WN_, bool Q16_ = false>
struct B3Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
    static constexpr bool kQ16 = Q16_;
    static constexpr int kThreads = WM * WN * 64;
    static constexpr int kWTM = BM / WM, kWTN = BN / WN;
    static constexpr int kFM = kWTM / 32, kFN = kWTN / 32;
    // LDS row of one tile row (A) or column (B): per 8-k group, 8 hi then 8 lo
    // bf16, plus 16 bytes of padding (conflict-free 128-bit fragment reads)
    static constexpr int kLdsRow = BK * 2 + 8;
    static constexpr int kStage = (BM + BN) * kLdsRow;    // bf16 elements per stage
    static constexpr int kGroups = BK / 8;                 // 8-k groups per tile row
    static constexpr int kItemsA = BM * kGroups / kThreads;
    static constexpr int kItemsB = BN * kGroups / kThreads;
    static_assert(kWTM % 32 == 0 && kWTN % 32 == 0, "wave tile");
    static_assert(2 * kStage * 2 <= 160 * 1024, "LDS");
    static_assert(BM * kGroups % kThreads == 0 && BN * kGroups % kThreads == 0, "items");
    static_assert(BN % 64 == 0 && kThreads % BN == 0, "B items: one k group per wave");
    // one stage in registers between its global loads and its split to LDS
    struct Regs {
        float a[kItemsA][8], b[kItemsB][8];
    };
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
// 256 x 256 tile, BK = 32, 8 waves of 64 x 128 (the fastest of the shapes
// measured at N = 8192, NITER = 10), on 32 x 32 MFMAs ...
using B3Main = B3Cfg<256, 256, 32, 4, 2>;
// ... and the production form on 16 x 16 MFMAs: the square kernel, the
// grouped large-tile launch and stream-K.
using B3Main16 = B3Cfg<256, 256, 32, 4, 2, true>;

// hi and lo planes of two floats, each as one dword holding two bf16 (x0 low).
__device__ __forceinline__ void split_pk(float x0, float x1, uint32_t &hi, uint32_t &lo) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const f2 v = {x0, x1};
    const b2 h = __builtin_convertvector(v, b2);
    const b2 l = __builtin_convertvector(v - __builtin_convertvector(h, f2), b2);
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

// Each workgroup loads its A (rows x 32 k) and B (32 k x cols) float32 tiles,
// splits every element in registers and stores the planes to LDS rows that
// are K-contiguous for both operands, as the MFMA fragments read them.
//   A item: one row, 8 consecutive k (two 16-byte loads).
//   B item: one column, 8 consecutive k (8 dword loads; the 64 lanes of a wave
//           read 64 adjacent columns, so each load is one 256-byte row segment).
// MODE (TileMode): kTileInterior, kTileEdge (any shape, A of any alignment;
// out-of-range elements are zero), kTileClamp (ragged M or N; K a multiple of
// 32 and A 16-byte loadable) or kTileClampK (the same with K a multiple of 4).
// Main loop: acc (this lane's kFM x kFN 32x32 blocks of the wave's sub-tile)
// = A[m0.., 0..K) . B[0..K, n0..), with Ag = A + m0 * lda (first row of the
// tile) and Bg = B + n0 (first column); a K range [k0, k1) of a larger
// product is Ag + k0, Bg + k0 * ldb and K = k1 - k0.  M, N, K bound the
// predicates.
// smem holds 2 * G::kStage bf16; regs is the stage in flight.  Every thread of
// the workgroup must call it.  ch: NoChain, or a TileChain to pipeline the
// next tile's first stages into this tile's last steps.  The wave computes
// the sub-tile at (om, on) of the tile: kFM x kFN blocks (CK = kWaveFull),
// kFM x kFN / 2 blocks (kWaveHalf) or none (kWaveNone: it still loads and
// splits its share of each stage and meets every barrier).
template <class G, int MODE, int CK, class Chain>
__device__ __forceinline__ void bf16x3_block_mma_impl(uint16_t *smem, const float *__restrict__ Ag,
                                                      const float *__restrict__ Bg, int M, int N,
                                                      int K, int lda, int ldb, int m0, int n0,
                                                      int om, int on,
                                                      floatx16 (&acc)[G::kFM][G::kFN],
                                                      typename G::Regs &regs, Chain &ch) {
    constexpr int kBM = G::BM, kBN = G::BN, kBK = G::BK, kThreads = G::kThreads;
    constexpr int kFM = G::kFM, kFN = G::kFN;
    constexpr int kFNc = CK == kWaveHalf ? kFN / 2 : kFN;  // column blocks computed
    constexpr int kLdsRow = G::kLdsRow, kStage = G::kStage;
    constexpr int kGroups = G::kGroups, kItemsA = G::kItemsA, kItemsB = G::kItemsB;

    constexpr bool kQ16 = G::kQ16;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int l32 = lane & 31, h = lane >> 5;  // 32 x 32 fragments: row / column, k half
    const int l16 = lane & 15, g = lane >> 4;  // 16 x 16 fragments: row / column, 8-k group

    // Q16: 16 x 16 blocks of v_mfma_f32_16x16x32_bf16, block (rr, cc) at rows
    // 16 rr, columns 16 cc of the wave's sub-tile, packed into acc at the end
    floatx4 a4[kQ16 ? 2 * kFM : 1][kQ16 ? 2 * kFN : 1];
    if constexpr (kQ16) {
#pragma unroll
        for (int r = 0; r < 2 * kFM; ++r)
#pragma unroll
            for (int c = 0; c < 2 * kFN; ++c) a4[r][c] = floatx4{0, 0, 0, 0};
    } else {
#pragma unroll
        for (int r = 0; r < kFM; ++r)
#pragma unroll
            for (int c = 0; c < kFN; ++c) acc[r][c] = floatx16{0};
    }

    // the stage in flight: the caller's registers when chained (they carry
    // the next tile's stage 1 across the call), else two local arrays
    float own_a[kItemsA][8], own_b[kItemsB][8];
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
        for (int i = 0; i < kItemsA; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kGroups, gk = k0 + (f % kGroups) * 8;
            const int srow = LM >= kTileClamp ? min(row, lM - lm0 - 1) : row;
            const float *x = lA + (int64_t)srow * llda + gk;
            if (LM != kTileEdge) {
                const floatx4 zero = {0, 0, 0, 0};
                const bool k0ok = LM != kTileClampK || gk < lK;
                const bool k1ok = LM != kTileClampK || gk + 4 < lK;
                const floatx4 v0 = k0ok ? ld<kG>((const floatx4 *)x) : zero;
                const floatx4 v1 = k1ok ? ld<kG>((const floatx4 *)(x + 4)) : zero;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ra[i][j] = v0[j];
                    ra[i][j + 4] = v1[j];
                }
            } else {
                const bool rok = lm0 + row < lM;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    ra[i][j] = (rok && gk + j < lK) ? ld<kG>(x + j) : 0.0f;
            }
        }
#pragma unroll
        for (int i = 0; i < kItemsB; ++i) {
            const int f = tid + i * kThreads;
            const unsigned col = LM >= kTileClamp ? (unsigned)min(f % kBN, lN - ln0 - 1)
                                                  : (unsigned)(f % kBN);
            // the k group is the same for the whole wave (kBN is a multiple of 64)
            // and each k row pointer is uniform: scalar base + 32-bit lane offset
            const int gk = k0 + __builtin_amdgcn_readfirstlane(f / kBN) * 8;
            const bool cok = LM != kTileEdge || ln0 + (int)col < lN;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float *rowp = lB + (int64_t)(gk + j) * lldb;
                const bool ok = LM == kTileInterior || LM == kTileClamp ||
                                (LM == kTileClampK ? gk + j < lK : (cok && gk + j < lK));
                rb[i][j] = ok ? ld<kG>(rowp + col) : 0.0f;
            }
        }
    };
    // With kTileClampK only the stage that runs past K needs K predicates;
    // every full stage takes the clamped loads (the choice is uniform).
    auto gload = [&](int k0) {
        if (MODE == kTileClampK && k0 + kBK > lK)
            gload_m(std::integral_constant<int, kTileClampK>{}, k0);
        else
            gload_m(std::integral_constant<int, MODE == kTileClampK ? kTileClamp : MODE>{}, k0);
    };
    auto sstore = [&](int buf) {
        uint16_t *as = smem + buf * kStage;
        uint16_t *bsm = as + kBM * kLdsRow;
#pragma unroll
        for (int i = 0; i < kItemsA; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kGroups, g = f % kGroups;
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) split_pk(ra[i][2 * j], ra[i][2 * j + 1], hi[j], lo[j]);
            uint16_t *d = as + row * kLdsRow + g * 16;
            *(uint4 *)d = uint4{hi[0], hi[1], hi[2], hi[3]};
            *(uint4 *)(d + 8) = uint4{lo[0], lo[1], lo[2], lo[3]};
        }
#pragma unroll
        for (int i = 0; i < kItemsB; ++i) {
            const int f = tid + i * kThreads;
            const int col = f % kBN, g = f / kBN;
            uint32_t hi[4], lo[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) split_pk(rb[i][2 * j], rb[i][2 * j + 1], hi[j], lo[j]);
            uint16_t *d = bsm + col * kLdsRow + g * 16;
            *(uint4 *)d = uint4{hi[0], hi[1], hi[2], hi[3]};
            *(uint4 *)(d + 8) = uint4{lo[0], lo[1], lo[2], lo[3]};
        }
    };
    // Half st of the stage in buffer `buf`.
    //   32 x 32 MFMAs: the 16-deep substep st; lane half h reads 8-k group
    //   2 st + h of its fragment row (or column) l & 31.
    //   Q16: the wave's 16-column blocks st kFNc .. st kFNc + kFNc - 1 of the
    //   stage's one 32-deep step against all its 16-row blocks; lane l reads
    //   8-k group l >> 4 of row (or column) l & 15 of a block.  The A
    //   fragments are read once per stage (load_a) and serve both halves:
    //   re-reading them in each half (a third more LDS reads) gave the
    //   grouped launch no gain over 32 x 32 MFMAs.
    // Each fragment is one 16-byte hi and one lo read.
    bf16x8 ahi16[kQ16 ? 2 * kFM : 1], alo16[kQ16 ? 2 * kFM : 1];
    auto load_a = [&](int buf) {
        if constexpr (kQ16) {
            if (CK == kWaveNone) return;
            const uint16_t *as = smem + buf * kStage + (om + l16) * kLdsRow + g * 16;
#pragma unroll
            for (int r = 0; r < 2 * kFM; ++r) {
                ahi16[r] = *(const bf16x8 *)(as + r * 16 * kLdsRow);
                alo16[r] = *(const bf16x8 *)(as + r * 16 * kLdsRow + 8);
            }
        }
    };
    // Consecutive MFMAs share their second operand (the B fragment): per
    // column block, each product term runs over the row blocks.  On the same
    // inputs this order measured 8% faster than issuing the three terms of
    // one block back to back (17% with all-zero data), while sharing the
    // first operand instead measured no gain.  Each accumulator still
    // receives lo.hi, hi.lo, hi.hi in that order.
    auto compute = [&](int buf, int st) {
        if (CK == kWaveNone) return;
        if constexpr (kQ16) {
            const uint16_t *bsm =
                smem + buf * kStage + kBM * kLdsRow + (on + l16) * kLdsRow + g * 16;
            bf16x8 bhi[kFNc], blo[kFNc];
#pragma unroll
            for (int j = 0; j < kFNc; ++j) {
                const int c = st * kFNc + j;
                bhi[j] = *(const bf16x8 *)(bsm + c * 16 * kLdsRow);
                blo[j] = *(const bf16x8 *)(bsm + c * 16 * kLdsRow + 8);
            }
#pragma unroll
            for (int j = 0; j < kFNc; ++j) {
                const int c = st * kFNc + j;
#pragma unroll
                for (int r = 0; r < 2 * kFM; ++r)
                    a4[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo16[r], bhi[j], a4[r][c],
                                                                       0, 0, 0);
#pragma unroll
                for (int r = 0; r < 2 * kFM; ++r)
                    a4[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi16[r], blo[j], a4[r][c],
                                                                       0, 0, 0);
#pragma unroll
                for (int r = 0; r < 2 * kFM; ++r)
                    a4[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi16[r], bhi[j], a4[r][c],
                                                                       0, 0, 0);
            }
        } else {
            const uint16_t *as = smem + buf * kStage + (om + l32) * kLdsRow;
            const uint16_t *bsm = smem + buf * kStage + kBM * kLdsRow + (on + l32) * kLdsRow;
            const int off = (2 * st + h) * 16;  // this lane half's 8-k group
            bf16x8 ahi[kFM], alo[kFM], bhi[kFNc], blo[kFNc];
#pragma unroll
            for (int r = 0; r < kFM; ++r) {
                ahi[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off);
                alo[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int c = 0; c < kFNc; ++c) {
                bhi[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off);
                blo[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int c = 0; c < kFNc; ++c) {
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[r], bhi[c], acc[r][c], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], blo[c], acc[r][c], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], bhi[c], acc[r][c], 0, 0, 0);
            }
        }
    };
    static_assert(kBK == 32, "each stage computed in two halves");

    // Pipeline: the global loads of stage kt + 2 are issued in the middle of
    // step kt, right after stage kt + 1 has been split into the other LDS
    // buffer, so each load has a whole step to arrive; the split of one
    // stage runs between the two substeps of the previous one.  Stage kt is
    // in LDS buffer (kt + b0) & 1.
    const int KT = (K + kBK - 1) / kBK;
    int b0 = 0;
    bool pre = false;
    if constexpr (Chain::kOn) {
        b0 = ch.buf0;
        pre = ch.pre;
    }
    if (!pre) {
        gload(0);
        sstore(b0);
        if (KT > 1) gload(kBK);
        __syncthreads();
    }
    // Q16 steady-state schedule of a full-width 64 x 128 wave (48 MFMAs per
    // half stage).  Left alone, the compiler runs a dozen MFMAs, then the
    // whole split (VALU) and LDS store of the next stage, then the other
    // MFMAs in one run, and the matrix core idles through the split (MFMA
    // busy 55.7% against 67.6% on 32 x 32).  Pinned here: half 0's MFMAs each
    // followed by three VALU instructions of the split, then by one LDS
    // write; then half 1's B reads (they may not pass the LDS writes); then
    // half 1's MFMAs with the next stage's global loads between them.  Not
    // for chained loops (the grouped large-tile launch): applied there, in
    // the same region as the main body, the group ran 7.5% slower (339 ->
    // 314 TFLOP/s, with 16 bytes of spills).
    auto q16_schedule = [&]() {
        if constexpr (kQ16 && !Chain::kOn && CK == kWaveFull && kFM == 2 && kFNc == 4) {
#pragma unroll
            for (int i = 0; i < 40; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // VALU
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // LDS write
            }
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);  // LDS read: half 1's B
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
#pragma unroll
            for (int i = 0; i < 20; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // global load
            }
        }
    };
    int kt = 0;
    for (; kt + 2 < KT; ++kt) {
        const int cur = (kt + b0) & 1;
        load_a(cur);
        compute(cur, 0);
        sstore(cur ^ 1);
        gload((kt + 2) * kBK);
        compute(cur, 1);
        if constexpr (Chain::kOn) {
            if (kt == 0) ch.publish();
        }
        q16_schedule();
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
        load_a(cur);
        compute(cur, 0);
        sstore(cur ^ 1);
        if constexpr (Chain::kOn) {
            if (chain) {
                retarget(ch.next);
                gload(0);
            }
        }
        compute(cur, 1);
        __syncthreads();
        ++kt;
    }
    const int cur = (kt + b0) & 1;
    load_a(cur);
    compute(cur, 0);
    if constexpr (Chain::kOn) {
        if (chain) {
            sstore(cur ^ 1);
            if (ch.next.K > kBK) gload(kBK);
        }
    }
    compute(cur, 1);
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
    if constexpr (kQ16) {
#pragma unroll
        for (int r = 0; r < kFM; ++r)
#pragma unroll
            for (int c = 0; c < kFN; ++c)
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    acc[r][c][q] = a4[2 * r + ((q >> 2) & 1)][2 * c + (q >> 3)][q & 3];
    }
}

// The sub-tile of wave `wave` in the tile t: for the 256 x 256 tile of 4 x 2
// waves outside kTileInterior, the edge-tile arrangement (big_wave_tile, so
// that the valid units are spread evenly over the SIMDs); else wave_pos.
template <class G, int MODE>
__device__ __forceinline__ WaveTile bf16x3_wave_tile(int M, int N, int m0, int n0, int wave) {
    if constexpr (MODE != kTileInterior && G::BM == 256 && G::BN == 256 && G::WM == 4 &&
                  G::WN == 2) {
        return big_wave_tile(M, N, m0, n0, wave);
    } else {
        int wm, wn;
        G::wave_pos(wave, wm, wn);
        return WaveTile{wm * G::kWTM, wn * G::kWTN, kWaveFull};
    }
}

// One tile's main loop; each wave runs the loop copy of its sub-tile kind
// (whole functions, so that the main loop stays one scheduling region).
template <class G, int MODE, class Chain>
__device__ __forceinline__ void bf16x3_block_mma(uint16_t *smem, const TileArgs &t,
                                                 floatx16 (&acc)[G::kFM][G::kFN],
                                                 typename G::Regs &regs, Chain &ch) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const WaveTile wt = bf16x3_wave_tile<G, MODE>(t.M, t.N, t.m0, t.n0, wave);
    if (wt.kind == kWaveFull)
        bf16x3_block_mma_impl<G, MODE, kWaveFull>(smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda, t.ldb,
                                                  t.m0, t.n0, wt.om, wt.on, acc, regs, ch);
    else if constexpr (MODE != kTileInterior && G::BM == 256 && G::BN == 256) {
        if (wt.kind == kWaveHalf)
            bf16x3_block_mma_impl<G, MODE, kWaveHalf>(smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda,
                                                      t.ldb, t.m0, t.n0, wt.om, wt.on, acc, regs,
                                                      ch);
        else
            bf16x3_block_mma_impl<G, MODE, kWaveNone>(smem, t.Ag, t.Bg, t.M, t.N, t.K, t.lda,
                                                      t.ldb, t.m0, t.n0, wt.om, wt.on, acc, regs,
                                                      ch);
    }
}
template <class G, int MODE>
__device__ __forceinline__ void bf16x3_block_mma(uint16_t *smem, const float *__restrict__ Ag,
                                                 const float *__restrict__ Bg, int M, int N, int K,
                                                 int lda, int ldb, int m0, int n0,
                                                 floatx16 (&acc)[G::kFM][G::kFN]) {
    const TileArgs t{Ag, Bg, M, N, K, lda, ldb, m0, n0};
    typename G::Regs regs;
    NoChain ch;
    bf16x3_block_mma<G, MODE>(smem, t, acc, regs, ch);
}

// Epilogue: stores acc to C (Cg = the matrix) at the tile (m0, n0), each wave
// its sub-tile (bf16x3_wave_tile); outside kTileInterior the stores of a tile
// that crosses M or N are predicated.
template <class G, int MODE>
__device__ __forceinline__ void bf16x3_block_store(float *__restrict__ Cg, int M, int N, int ldc,
                                                   int m0, int n0,
                                                   const floatx16 (&acc)[G::kFM][G::kFN]) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const WaveTile wt = bf16x3_wave_tile<G, MODE>(M, N, m0, n0, wave);
    if (wt.kind == kWaveNone) return;
    const int ncols = wt.kind == kWaveHalf ? G::kFN / 2 : G::kFN;
    if (MODE == kTileInterior || (m0 + G::BM <= M && n0 + G::BN <= N))
        store_acc<G::kFM, G::kFN, false, G::kQ16>(Cg, ldc, m0 + wt.om, n0 + wt.on, M, N, ncols,
                                                   acc);
    else
        store_acc<G::kFM, G::kFN, true, G::kQ16>(Cg, ldc, m0 + wt.om, n0 + wt.on, M, N, ncols,
                                                  acc);
}

// One whole output tile: main loop and store.
template <class G, int MODE>
__device__ __forceinline__ void bf16x3_block_tile(uint16_t *smem, const float *__restrict__ Ag,
                                                  const float *__restrict__ Bg,
                                                  float *__restrict__ Cg, int M, int N, int K,
                                                  int lda, int ldb, int ldc, int m0, int n0) {
    floatx16 acc[G::kFM][G::kFN];
    bf16x3_block_mma<G, MODE>(smem, Ag, Bg, M, N, K, lda, ldb, m0, n0, acc);
    bf16x3_block_store<G, MODE>(Cg, M, N, ldc, m0, n0, acc);
}

}  // namespace hbs
