// sgemm_f32.hip -- plain FP32 SGEMM on gfx950 FP32 matrix cores.
//
// Replaces the HammerBlade SPMD kernel of apps/sgemm/kernel.cpp:279-333 (16x16
// output blocks strided over a 16x8 tile group, DRAM->scratchpad block loads
// at :36-109, a 4x4 register-blocked fmaf micro-kernel at :186-275).  The
// MI355X design instead gives each workgroup one BM x BN output tile:
//   * global -> registers -> LDS double buffer, 128-bit coalesced loads; each
//     K stage is loaded a full step ahead and stored to LDS in the middle of
//     the previous step;
//   * each wave owns a (BM/WM) x (BN/WN) sub-tile held as 32x32 accumulator
//     blocks of v_mfma_f32_32x32x2_f32 (FP32 in, FP32 accumulate);
//   * an A fragment for four consecutive MFMAs comes from one 128-bit LDS read
//     (lane half h takes k = 4h..4h+3 of each 8-wide k chunk);
//   * XCD-aware workgroup order plus a grouped raster for L2 reuse;
//   * NITER slices (kernel.cpp:293-297) are extra tiles of the same launch.
#include "hbs_internal.h"

namespace {

struct GemmArgs {
    const float *A;
    const float *B;
    float *C;
    int M, N, K;
    int lda, ldb, ldc;
    int64_t sA, sB, sC;
    int tiles_m, tiles_n, batch, group_m;
};

template <int BM, int BN, int BK, int WM, int WN>
struct Cfg {
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
    static_assert(kWTM % 32 == 0 && kWTN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert((BM * BK / 4) % kThreads == 0, "A tile must split evenly over threads");
    static_assert((BK * BN / 4) % kThreads == 0, "B tile must split evenly over threads");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16 (two or more 8-k chunks)");
};

template <int BM, int BN, int BK, int WM, int WN, bool EDGE>
__global__ __launch_bounds__(WM *WN * 64) void sgemm_f32_kernel(GemmArgs p) {
    using G = Cfg<BM, BN, BK, WM, WN>;
    __shared__ __attribute__((aligned(16))) float smem[G::kLdsFloats];
    float *const As = smem;
    float *const Bs = smem + 2 * G::kStageA;

    const int tiles = p.tiles_m * p.tiles_n;
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    const int bslice = bid / tiles;
    int tm, tn;
    hbs::tile_coords(bid - bslice * tiles, p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    const float *__restrict__ Ag = p.A + bslice * p.sA + (int64_t)m0 * p.lda;
    const float *__restrict__ Bg = p.B + bslice * p.sB + n0;
    float *__restrict__ Cg = p.C + bslice * p.sC;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int l32 = lane & 31, h = lane >> 5;

    floatx16 acc[G::kFM][G::kFN];
#pragma unroll
    for (int r = 0; r < G::kFM; ++r)
#pragma unroll
        for (int c = 0; c < G::kFN; ++c) acc[r][c] = floatx16{0};

    floatx4 ra[G::kAF4], rb[G::kBF4];

    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < G::kAF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int row = f / (BK / 4), c4 = f % (BK / 4);
            if constexpr (!EDGE) {
                ra[i] = *(const floatx4 *)(Ag + (int64_t)row * p.lda + k0 + c4 * 4);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gk = k0 + c4 * 4 + j;
                    ra[i][j] = (m0 + row < p.M && gk < p.K) ? Ag[(int64_t)row * p.lda + gk] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < G::kBF4; ++i) {
            const int f = tid + i * G::kThreads;
            const int kr = f / (BN / 4), c4 = f % (BN / 4);
            if constexpr (!EDGE) {
                rb[i] = *(const floatx4 *)(Bg + (int64_t)(k0 + kr) * p.ldb + c4 * 4);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int gn = n0 + c4 * 4 + j;
                    rb[i][j] = (k0 + kr < p.K && gn < p.N) ? Bg[(int64_t)(k0 + kr) * p.ldb + c4 * 4 + j]
                                                          : 0.0f;
                }
            }
        }
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
                         float (&b)[G::kFN][4]) {
#pragma unroll
        for (int r = 0; r < G::kFM; ++r) a[r] = *(const floatx4 *)(as + r * 32 * G::kLdA + kc * 8);
#pragma unroll
        for (int c = 0; c < G::kFN; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) b[c][q] = bs[(kc * 8 + q) * G::kLdB + c * 32];
    };
    auto mfma_chunk = [&](const floatx4 (&a)[G::kFM], const float (&b)[G::kFN][4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < G::kFM; ++r)
#pragma unroll
                for (int c = 0; c < G::kFN; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[r][q], b[c][q], acc[r][c], 0, 0, 0);
    };
    // Scheduling hint: spread the next chunk's LDS reads (kFM 128-bit and
    // 2 kFN paired 32-bit reads) between this chunk's 4 kFM kFN MFMAs.
    auto interleave = [&]() {
        constexpr int kReads = G::kFM + 2 * G::kFN;
        constexpr int kMfma = 4 * G::kFM * G::kFN;
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
        const float *as = As + buf * G::kStageA + (wm * G::kWTM + l32) * G::kLdA + 4 * h;
        const float *bs = Bs + buf * G::kStageB + 4 * h * G::kLdB + wn * G::kWTN + l32;
        floatx4 a0[G::kFM], a1[G::kFM];
        float b0[G::kFN][4], b1[G::kFN][4];
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
    const int KT = (p.K + BK - 1) / BK;
    gload(0);
    sstore(0);
    if (KT > 1) gload(BK);
    __syncthreads();
    int kt = 0;
    for (; kt + 2 < KT; ++kt) {
        const int cur = kt & 1;
        compute(cur, [&]() {
            sstore(cur ^ 1);
            gload((kt + 2) * BK);
        });
        __syncthreads();
    }
    if (kt + 1 < KT) {
        const int cur = kt & 1;
        compute(cur, [&]() { sstore(cur ^ 1); });
        __syncthreads();
        ++kt;
    }
    compute(kt & 1, [&]() {});

    // Epilogue: 32x32 accumulator layout -- register q of lane l holds
    // row 8*(q/4) + 4*(l/32) + q%4, col l%32 of its block.
#pragma unroll
    for (int r = 0; r < G::kFM; ++r) {
#pragma unroll
        for (int c = 0; c < G::kFN; ++c) {
            const int col = n0 + wn * G::kWTN + c * 32 + l32;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = m0 + wm * G::kWTM + r * 32 + 8 * (q >> 2) + 4 * h + (q & 3);
                if (!EDGE || (row < p.M && col < p.N)) Cg[(int64_t)row * p.ldc + col] = acc[r][c][q];
            }
        }
    }
}

template <int BM, int BN, int BK, int WM, int WN>
int launch_cfg(GemmArgs p, bool edge, hipStream_t stream) {
    using G = Cfg<BM, BN, BK, WM, WN>;
    p.tiles_m = (p.M + BM - 1) / BM;
    p.tiles_n = (p.N + BN - 1) / BN;
    const int64_t blocks = (int64_t)p.tiles_m * p.tiles_n * p.batch;
    if (blocks > 0x7fffffffLL) return HBS_ERR_ARG;
    if (edge)
        sgemm_f32_kernel<BM, BN, BK, WM, WN, true><<<(unsigned)blocks, G::kThreads, 0, stream>>>(p);
    else
        sgemm_f32_kernel<BM, BN, BK, WM, WN, false><<<(unsigned)blocks, G::kThreads, 0, stream>>>(p);
    return hbs::launch_status();
}

int run_f32(GemmArgs p, hipStream_t stream) {
    const bool vec_ok = hbs::aligned16(p.A) && hbs::aligned16(p.B) && (p.lda % 4 == 0) &&
                        (p.ldb % 4 == 0) && (p.sA % 4 == 0) && (p.sB % 4 == 0);
    // Largest tile whose grid still covers the 256 CUs.
    const int64_t big_tiles = (int64_t)((p.M + 255) / 256) * ((p.N + 255) / 256) * p.batch;
    if (big_tiles >= 256) {
        const bool edge = !(vec_ok && p.M % 256 == 0 && p.N % 256 == 0 && p.K % 32 == 0);
        return launch_cfg<256, 256, 32, 4, 2>(p, edge, stream);
    }
    const int64_t mid_tiles = (int64_t)((p.M + 127) / 128) * ((p.N + 127) / 128) * p.batch;
    if (mid_tiles >= 128) {
        const bool edge = !(vec_ok && p.M % 128 == 0 && p.N % 128 == 0 && p.K % 16 == 0);
        return launch_cfg<128, 128, 16, 2, 2>(p, edge, stream);
    }
    const bool edge = !(vec_ok && p.M % 64 == 0 && p.N % 64 == 0 && p.K % 16 == 0);
    return launch_cfg<64, 64, 16, 2, 2>(p, edge, stream);
}

}  // namespace

extern "C" int hbs_sgemm_strided(const float *A, const float *B, float *C, int M, int N, int K,
                                 int lda, int ldb, int ldc, int64_t strideA, int64_t strideB,
                                 int64_t strideC, int batch, void *stream) {
    if (M < 1 || N < 1 || K < 1 || batch < 1) return HBS_ERR_ARG;
    if (lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (strideA < 0 || strideB < 0 || strideC < 0) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    const int st = hbs::check_device();
    if (st) return st;
    GemmArgs p{A, B, C, M, N, K, lda, ldb, ldc, strideA, strideB, strideC, 0, 0, batch, 8};
    return run_f32(p, (hipStream_t)stream);
}

extern "C" int hbs_sgemm(const float *A, const float *B, float *C, int N, int niter,
                         void *stream) {
    if (N < 1 || niter < 1) return HBS_ERR_ARG;
    const int64_t nn = (int64_t)N * N;
    if (nn > ((int64_t)1 << 62) / niter) return HBS_ERR_ARG;
    return hbs_sgemm_strided(A, B, C, N, N, N, N, N, N, nn, nn, nn, niter, stream);
}
