// sgemm_bf16x3.hip -- split-precision SGEMM on gfx950 BF16 matrix cores.
//
// No reference counterpart: a precision variant of apps/sgemm's C = A.B
// (kernel.cpp:279-333 call shape) with FP32-level error.  Each float is split
// x = hi + lo, hi = bf16_rne(x), lo = bf16_rne(x - hi) (x - hi is exact), and
//   C = Alo.Bhi + Ahi.Blo + Ahi.Bhi
// is accumulated in FP32 by v_mfma_f32_32x32x16_bf16; the lo.lo term
// (relative size 2^-16) is dropped.
//
// Two launches per call:
//  1. pack: A -> Ap[Mp][Kp/8][hi8|lo8] and B -> Bp[Np][Kp/8][hi8|lo8] (B is
//     transposed so both operands are K-contiguous); rows and K are zero-padded
//     to the tile multiples, so the GEMM needs no load bounds checks;
//  2. GEMM (B3Main): 256x256 output tile per workgroup, BK = 32, global ->
//     registers -> LDS double buffer of the packed rows, 8 waves each 64x128.
#include "hbs_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short ushort8 __attribute__((ext_vector_type(8)));

namespace {

// Packed operands are padded to these multiples so that every GEMM tile
// configuration below can read them without bounds checks.
constexpr int kRowPad = 256, kKPad = 64;

template <int BM_, int BN_, int BK_, int WM_, int WN_>
struct B3Cfg {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
    static constexpr int kThreads = WM * WN * 64;
    static constexpr int kWTM = BM / WM, kWTN = BN / WN;
    static constexpr int kFM = kWTM / 32, kFN = kWTN / 32;
    static constexpr int kRowElems = BK * 2;              // packed bf16 per tile row (hi + lo)
    static constexpr int kLdsRow = kRowElems + 8;         // + 16 B pad
    static constexpr int kStage = (BM + BN) * kLdsRow;    // bf16 elements per stage
    static constexpr int kChunksPerRow = kRowElems / 8;   // 16-byte chunks per tile row
    static constexpr int kLoadsA = BM * kChunksPerRow / kThreads;
    static constexpr int kLoadsB = BN * kChunksPerRow / kThreads;
    static_assert(kWTM % 32 == 0 && kWTN % 32 == 0, "wave tile");
    static_assert(BM * kChunksPerRow % kThreads == 0 && BN * kChunksPerRow % kThreads == 0, "loads");
    static_assert(kRowPad % BM == 0 && kRowPad % BN == 0 && kKPad % BK == 0, "padding");
    static_assert(2 * kStage * 2 <= 160 * 1024, "LDS");
};
// Production configuration: 256 x 256 tile, BK = 32, 8 waves of 64 x 128 (the
// fastest of the shapes measured at N = 8192, NITER = 10).
using B3Main = B3Cfg<256, 256, 32, 4, 2>;

// x = hi + lo with hi = bf16_rne(x), lo = bf16_rne(x - hi); x - hi is exact.
// v_cvt_pk_bf16_f32 rounds to nearest even, two floats per instruction.
__device__ __forceinline__ void split2(float x0, float x1, uint16_t &h0, uint16_t &h1,
                                       uint16_t &l0, uint16_t &l1) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const f2 v = {x0, x1};
    const b2 h = __builtin_convertvector(v, b2);
    const b2 l = __builtin_convertvector(v - __builtin_convertvector(h, f2), b2);
    const uint32_t hu = __builtin_bit_cast(uint32_t, h), lu = __builtin_bit_cast(uint32_t, l);
    h0 = (uint16_t)hu;
    h1 = (uint16_t)(hu >> 16);
    l0 = (uint16_t)lu;
    l1 = (uint16_t)(lu >> 16);
}

struct PackArgs {
    const float *X;
    uint16_t *P;
    int rows, cols;      // logical matrix (rows x cols for A; K x N source for B)
    int ld;              // leading dimension of X
    int rows_p, cols_p;  // padded packed extents (rows_p x cols_p, cols_p % 8 == 0)
    int64_t sX, sP;      // slice strides (elements)
    int vec;             // X, ld and sX allow 16-byte loads
};

// A [M][K] -> Ap[Mp][Kp/8][hi8 lo8]; one thread per 8-element group.
__global__ __launch_bounds__(256) void pack_rows_kernel(PackArgs a, int64_t groups_per_slice,
                                                        int64_t total) {
    const int gpr = a.cols_p / 8;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = g / groups_per_slice;
        const int64_t r = g - s * groups_per_slice;
        const int row = (int)(r / gpr), kg = (int)(r % gpr);
        const float *x = a.X + s * a.sX + (int64_t)row * a.ld + kg * 8;
        ushort8 hi, lo;
        float v[8];
        if (a.vec && row < a.rows && kg * 8 + 8 <= a.cols) {
            const floatx4 p0 = *(const floatx4 *)x, p1 = *(const floatx4 *)(x + 4);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                v[j] = p0[j];
                v[j + 4] = p1[j];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (row < a.rows && kg * 8 + j < a.cols) ? x[j] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            uint16_t h0, h1, l0, l1;
            split2(v[j], v[j + 1], h0, h1, l0, l1);
            hi[j] = h0; hi[j + 1] = h1;
            lo[j] = l0; lo[j + 1] = l1;
        }
        ushort8 *dst = (ushort8 *)(a.P + s * a.sP + ((int64_t)row * gpr + kg) * 16);
        dst[0] = hi;
        dst[1] = lo;
    }
}

// B [K][N] -> Bp[Np][Kp/8][hi8 lo8] (transposed) through a 64x64 LDS tile.
__global__ __launch_bounds__(256) void pack_cols_kernel(PackArgs a, int tiles_k, int tiles_n) {
    __shared__ float t[64][65];
    const int per_slice = tiles_k * tiles_n;
    const int s = blockIdx.x / per_slice;
    const int r = blockIdx.x % per_slice;
    const int tk = r / tiles_n, tn = r % tiles_n;
    const int k0 = tk * 64, n0 = tn * 64;
    const float *X = a.X + (int64_t)s * a.sX;
    const int tid = threadIdx.x;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int kk = e / 64, nn = e % 64;
        const int gk = k0 + kk, gn = n0 + nn;
        t[kk][nn] = (gk < a.rows && gn < a.cols) ? X[(int64_t)gk * a.ld + gn] : 0.0f;
    }
    __syncthreads();
    const int gpr = a.rows_p / 8;  // k groups per packed row (rows_p = Kp here)
    uint16_t *P = a.P + (int64_t)s * a.sP;
    for (int e = tid; e < 64 * 8; e += 256) {
        const int nn = e / 8, kg = e % 8;
        const int gn = n0 + nn, gkg = k0 / 8 + kg;
        if (gn >= a.cols_p || gkg >= gpr) continue;
        ushort8 hi, lo;
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            uint16_t h0, h1, l0, l1;
            split2(t[kg * 8 + j][nn], t[kg * 8 + j + 1][nn], h0, h1, l0, l1);
            hi[j] = h0; hi[j + 1] = h1;
            lo[j] = l0; lo[j + 1] = l1;
        }
        ushort8 *dst = (ushort8 *)(P + ((int64_t)gn * gpr + gkg) * 16);
        dst[0] = hi;
        dst[1] = lo;
    }
}

struct GemmArgs {
    const uint16_t *Ap;
    const uint16_t *Bp;
    float *C;
    int M, N, Kp, ldc;
    int64_t sAp, sBp, sC;
    int tiles_m, tiles_n, group_m;
};

template <class G, bool EDGE>
__global__ __launch_bounds__(G::kThreads) void gemm_bf16x3_kernel(GemmArgs p) {
    constexpr int kBM = G::BM, kBN = G::BN, kBK = G::BK, kWM = G::WM, kThreads = G::kThreads;
    constexpr int kWTM = G::kWTM, kWTN = G::kWTN, kFM = G::kFM, kFN = G::kFN;
    constexpr int kLdsRow = G::kLdsRow, kStage = G::kStage, kChunksPerRow = G::kChunksPerRow;
    __shared__ __attribute__((aligned(16))) uint16_t smem[2 * kStage];

    const int tiles = p.tiles_m * p.tiles_n;
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    const int bs = bid / tiles;
    int tm, tn;
    hbs::tile_coords(bid - bs * tiles, p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * kBM, n0 = tn * kBN;
    const int64_t row_elems = (int64_t)p.Kp * 2;
    const uint16_t *__restrict__ Ag = p.Ap + bs * p.sAp + (int64_t)m0 * row_elems;
    const uint16_t *__restrict__ Bg = p.Bp + bs * p.sBp + (int64_t)n0 * row_elems;
    float *__restrict__ Cg = p.C + bs * p.sC;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kWM, wn = wave / kWM;
    const int l32 = lane & 31, h = lane >> 5;

    floatx16 acc[kFM][kFN];
#pragma unroll
    for (int r = 0; r < kFM; ++r)
#pragma unroll
        for (int c = 0; c < kFN; ++c) acc[r][c] = floatx16{0};

    floatx4 ra[G::kLoadsA], rb[G::kLoadsB];

    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < G::kLoadsA; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kChunksPerRow, c = f % kChunksPerRow;
            ra[i] = *(const floatx4 *)(Ag + (int64_t)row * row_elems + (int64_t)k0 * 2 + c * 8);
        }
#pragma unroll
        for (int i = 0; i < G::kLoadsB; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kChunksPerRow, c = f % kChunksPerRow;
            rb[i] = *(const floatx4 *)(Bg + (int64_t)row * row_elems + (int64_t)k0 * 2 + c * 8);
        }
    };
    auto sstore = [&](int buf) {
        uint16_t *as = smem + buf * kStage;
        uint16_t *bsm = as + kBM * kLdsRow;
#pragma unroll
        for (int i = 0; i < G::kLoadsA; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kChunksPerRow, c = f % kChunksPerRow;
            *(floatx4 *)(as + row * kLdsRow + c * 8) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < G::kLoadsB; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kChunksPerRow, c = f % kChunksPerRow;
            *(floatx4 *)(bsm + row * kLdsRow + c * 8) = rb[i];
        }
    };
    auto compute = [&](int buf) {
        const uint16_t *as = smem + buf * kStage + (wm * kWTM + l32) * kLdsRow;
        const uint16_t *bsm = smem + buf * kStage + kBM * kLdsRow + (wn * kWTN + l32) * kLdsRow;
#pragma unroll
        for (int st = 0; st < kBK / 16; ++st) {
            const int off = (2 * st + h) * 16;  // this lane half's 8-k group
            bf16x8 ahi[kFM], alo[kFM], bhi[kFN], blo[kFN];
#pragma unroll
            for (int r = 0; r < kFM; ++r) {
                ahi[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off);
                alo[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int c = 0; c < kFN; ++c) {
                bhi[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off);
                blo[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int r = 0; r < kFM; ++r)
#pragma unroll
                for (int c = 0; c < kFN; ++c) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[r], bhi[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], blo[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], bhi[c], acc[r][c], 0, 0, 0);
                }
        }
    };

    const int KT = p.Kp / kBK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < KT) gload((kt + 1) * kBK);
        compute(cur);
        if (kt + 1 < KT) sstore(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int r = 0; r < kFM; ++r) {
#pragma unroll
        for (int c = 0; c < kFN; ++c) {
            const int col = n0 + wn * kWTN + c * 32 + l32;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = m0 + wm * kWTM + r * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
                if (!EDGE || (row < p.M && col < p.N)) Cg[(int64_t)row * p.ldc + col] = acc[r][c][q];
            }
        }
    }
}

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

struct FusedArgs {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    int64_t sA, sB, sC;
    int tiles_m, tiles_n, group_m;
};

// The same GEMM reading float32 A and B directly: each workgroup loads its
// A (rows x 32 k) and B (32 k x cols) tiles, splits every element in
// registers and stores the split planes to LDS in the packed row layout the
// packed-operand kernel uses, so the MFMA sequence and the result are
// identical to pack + gemm_bf16x3_kernel, with no pre-pass and no workspace.
//   A item: one row, 8 consecutive k (two 16-byte loads).
//   B item: one column, 8 consecutive k (8 dword loads; the 64 lanes of a wave
//           read 64 adjacent columns, so each load is one 256-byte row segment).
// EDGE: ragged M, N or K, or A not 16-byte loadable; out-of-range elements are
// zero.
template <class G, bool EDGE>
__global__ __launch_bounds__(G::kThreads) void gemm_bf16x3_fused_kernel(FusedArgs p) {
    constexpr int kBM = G::BM, kBN = G::BN, kBK = G::BK, kWM = G::WM, kThreads = G::kThreads;
    constexpr int kWTM = G::kWTM, kWTN = G::kWTN, kFM = G::kFM, kFN = G::kFN;
    constexpr int kLdsRow = G::kLdsRow, kStage = G::kStage;
    constexpr int kGroups = kBK / 8;                       // 8-k groups per tile row
    constexpr int kItemsA = kBM * kGroups / kThreads;
    constexpr int kItemsB = kBN * kGroups / kThreads;
    static_assert(kBM * kGroups % kThreads == 0 && kBN * kGroups % kThreads == 0, "items");
    static_assert(kBN % 64 == 0 && kThreads % kBN == 0, "B items: one k group per wave");
    __shared__ __attribute__((aligned(16))) uint16_t smem[2 * kStage];

    const int tiles = p.tiles_m * p.tiles_n;
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    const int bs = bid / tiles;
    int tm, tn;
    hbs::tile_coords(bid - bs * tiles, p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * kBM, n0 = tn * kBN;
    const float *__restrict__ Ag = p.A + bs * p.sA + (int64_t)m0 * p.lda;
    const float *__restrict__ Bg = p.B + bs * p.sB + n0;
    float *__restrict__ Cg = p.C + bs * p.sC;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kWM, wn = wave / kWM;
    const int l32 = lane & 31, h = lane >> 5;

    floatx16 acc[kFM][kFN];
#pragma unroll
    for (int r = 0; r < kFM; ++r)
#pragma unroll
        for (int c = 0; c < kFN; ++c) acc[r][c] = floatx16{0};

    float ra[kItemsA][8], rb[kItemsB][8];

    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kItemsA; ++i) {
            const int f = tid + i * kThreads;
            const int row = f / kGroups, gk = k0 + (f % kGroups) * 8;
            const float *x = Ag + (int64_t)row * p.lda + gk;
            if (!EDGE) {
                const floatx4 v0 = *(const floatx4 *)x, v1 = *(const floatx4 *)(x + 4);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ra[i][j] = v0[j];
                    ra[i][j + 4] = v1[j];
                }
            } else {
                const bool rok = m0 + row < p.M;
#pragma unroll
                for (int j = 0; j < 8; ++j) ra[i][j] = (rok && gk + j < p.K) ? x[j] : 0.0f;
            }
        }
#pragma unroll
        for (int i = 0; i < kItemsB; ++i) {
            const int f = tid + i * kThreads;
            const unsigned col = f % kBN;
            // the k group is the same for the whole wave (kBN is a multiple of 64)
            // and each k row pointer is uniform: scalar base + 32-bit lane offset
            const int gk = k0 + __builtin_amdgcn_readfirstlane(f / kBN) * 8;
            const bool cok = !EDGE || n0 + col < p.N;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float *rowp = Bg + (int64_t)(gk + j) * p.ldb;
                rb[i][j] = (!EDGE || (cok && gk + j < p.K)) ? rowp[col] : 0.0f;
            }
        }
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
    auto compute = [&](int buf) {
        const uint16_t *as = smem + buf * kStage + (wm * kWTM + l32) * kLdsRow;
        const uint16_t *bsm = smem + buf * kStage + kBM * kLdsRow + (wn * kWTN + l32) * kLdsRow;
#pragma unroll
        for (int st = 0; st < kBK / 16; ++st) {
            const int off = (2 * st + h) * 16;  // this lane half's 8-k group
            bf16x8 ahi[kFM], alo[kFM], bhi[kFN], blo[kFN];
#pragma unroll
            for (int r = 0; r < kFM; ++r) {
                ahi[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off);
                alo[r] = *(const bf16x8 *)(as + r * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int c = 0; c < kFN; ++c) {
                bhi[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off);
                blo[c] = *(const bf16x8 *)(bsm + c * 32 * kLdsRow + off + 8);
            }
#pragma unroll
            for (int r = 0; r < kFM; ++r)
#pragma unroll
                for (int c = 0; c < kFN; ++c) {
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[r], bhi[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], blo[c], acc[r][c], 0, 0, 0);
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[r], bhi[c], acc[r][c], 0, 0, 0);
                }
        }
    };

    const int KT = (p.K + kBK - 1) / kBK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < KT) gload((kt + 1) * kBK);
        compute(cur);
        if (kt + 1 < KT) sstore(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int r = 0; r < kFM; ++r) {
#pragma unroll
        for (int c = 0; c < kFN; ++c) {
            const int col = n0 + wn * kWTN + c * 32 + l32;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = m0 + wm * kWTM + r * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
                if (!EDGE || (row < p.M && col < p.N)) Cg[(int64_t)row * p.ldc + col] = acc[r][c][q];
            }
        }
    }
}

template <class G>
int launch_fused(const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb,
                 int ldc, int64_t sA, int64_t sB, int64_t sC, int batch, int group_m,
                 hipStream_t stream) {
    FusedArgs g{A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC,
                (M + G::BM - 1) / G::BM, (N + G::BN - 1) / G::BN, group_m};
    const int64_t gblocks = (int64_t)g.tiles_m * g.tiles_n * batch;
    if (gblocks > 0x7fffffffLL) return HBS_ERR_ARG;
    const bool interior = M % G::BM == 0 && N % G::BN == 0 && K % G::BK == 0;
    const bool vec_a = hbs::aligned16(A) && lda % 4 == 0 && sA % 4 == 0;
    if (interior && vec_a)
        gemm_bf16x3_fused_kernel<G, false><<<(unsigned)gblocks, G::kThreads, 0, stream>>>(g);
    else
        gemm_bf16x3_fused_kernel<G, true><<<(unsigned)gblocks, G::kThreads, 0, stream>>>(g);
    return hbs::launch_status();
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct Layout {
    int Mp, Np, Kp;
    int64_t sAp, sBp;  // elements per slice
    size_t bytes_a, bytes_total;
};

Layout layout(int M, int N, int K, int batch) {
    Layout L;
    L.Mp = (int)round_up(M, kRowPad);
    L.Np = (int)round_up(N, kRowPad);
    L.Kp = (int)round_up(K, kKPad);
    L.sAp = (int64_t)L.Mp * L.Kp * 2;
    L.sBp = (int64_t)L.Np * L.Kp * 2;
    L.bytes_a = (size_t)round_up(L.sAp * batch * 2, 256);
    L.bytes_total = L.bytes_a + (size_t)L.sBp * batch * 2;
    return L;
}

template <class G>
int launch_gemm(const uint16_t *Ap, const uint16_t *Bp, float *C, int M, int N, const Layout &L,
                int ldc, int64_t sC, int batch, int group_m, hipStream_t stream) {
    GemmArgs g{Ap, Bp, C, M, N, L.Kp, ldc, L.sAp, L.sBp, sC, L.Mp / G::BM, L.Np / G::BN, group_m};
    const int64_t gblocks = (int64_t)g.tiles_m * g.tiles_n * batch;
    if (gblocks > 0x7fffffffLL) return HBS_ERR_ARG;
    if (M % G::BM == 0 && N % G::BN == 0)
        gemm_bf16x3_kernel<G, false><<<(unsigned)gblocks, G::kThreads, 0, stream>>>(g);
    else
        gemm_bf16x3_kernel<G, true><<<(unsigned)gblocks, G::kThreads, 0, stream>>>(g);
    return hbs::launch_status();
}

int run_bf16x3(const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb,
               int ldc, int64_t sA, int64_t sB, int64_t sC, int batch, void *ws, size_t ws_bytes,
               hipStream_t stream) {
    const Layout L = layout(M, N, K, batch);
    if (!ws || ws_bytes < L.bytes_total) return HBS_ERR_WORKSPACE;
    if (!hbs::aligned16(ws)) return HBS_ERR_ALIGN;
    uint16_t *Ap = (uint16_t *)ws;
    uint16_t *Bp = (uint16_t *)((char *)ws + L.bytes_a);

    const int vec_a = hbs::aligned16(A) && lda % 4 == 0 && sA % 4 == 0;
    PackArgs pa{A, Ap, M, K, lda, L.Mp, L.Kp, sA, L.sAp, vec_a};
    const int64_t gps = (int64_t)L.Mp * (L.Kp / 8);
    const int64_t total = gps * batch;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 262144) blocks = 262144;
    pack_rows_kernel<<<(unsigned)blocks, 256, 0, stream>>>(pa, gps, total);
    int st = hbs::launch_status();
    if (st) return st;

    PackArgs pb{B, Bp, K, N, ldb, L.Kp, L.Np, sB, L.sBp, 0};
    const int tk = L.Kp / 64 + (L.Kp % 64 ? 1 : 0), tnb = L.Np / 64;
    const int64_t pblocks = (int64_t)tk * tnb * batch;
    if (pblocks > 0x7fffffffLL) return HBS_ERR_ARG;
    pack_cols_kernel<<<(unsigned)pblocks, 256, 0, stream>>>(pb, tk, tnb);
    st = hbs::launch_status();
    if (st) return st;

    return launch_gemm<B3Main>(Ap, Bp, C, M, N, L, ldc, sC, batch, 8, stream);
}

}  // namespace

extern "C" size_t hbs_gemm_bf16x3_workspace_size(int M, int N, int K) {
    if (M < 1 || N < 1 || K < 1) return 0;
    return layout(M, N, K, 1).bytes_total;
}

extern "C" int hbs_gemm_bf16x3(const float *A, const float *B, float *C, int M, int N, int K,
                               int lda, int ldb, int ldc, void *workspace, size_t workspace_bytes,
                               void *stream) {
    if (M < 1 || N < 1 || K < 1 || lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    if (!workspace || workspace_bytes < layout(M, N, K, 1).bytes_total) return HBS_ERR_WORKSPACE;
    if (!hbs::aligned16(workspace)) return HBS_ERR_ALIGN;
    const int st = hbs::check_device();
    if (st) return st;
    return run_bf16x3(A, B, C, M, N, K, lda, ldb, ldc, 0, 0, 0, 1, workspace, workspace_bytes,
                      (hipStream_t)stream);
}

extern "C" size_t hbs_sgemm_bf16x3_workspace_size(int N, int niter) {
    if (N < 1 || niter < 1) return 0;
    return layout(N, N, N, niter).bytes_total;
}

extern "C" int hbs_sgemm_bf16x3(const float *A, const float *B, float *C, int N, int niter,
                                void *workspace, size_t workspace_bytes, void *stream) {
    if (N < 1 || niter < 1) return HBS_ERR_ARG;
    const int64_t nn = (int64_t)N * N;
    if (nn > ((int64_t)1 << 62) / niter) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    if (!workspace || workspace_bytes < layout(N, N, N, niter).bytes_total) return HBS_ERR_WORKSPACE;
    if (!hbs::aligned16(workspace)) return HBS_ERR_ALIGN;
    const int st = hbs::check_device();
    if (st) return st;
    return run_bf16x3(A, B, C, N, N, N, N, N, N, nn, nn, nn, niter, workspace, workspace_bytes,
                      (hipStream_t)stream);
}
