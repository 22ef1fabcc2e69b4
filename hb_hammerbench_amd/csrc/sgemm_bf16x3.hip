// sgemm_bf16x3.hip -- split-precision SGEMM on gfx950 BF16 matrix cores.
//
// No reference counterpart: a precision variant of apps/sgemm's C = A.B
// (kernel.cpp:279-333 call shape) with FP32-level error.  Each float is split
// x = hi + lo, hi = bf16_rne(x), lo = bf16_rne(x - hi) (x - hi is exact), and
//   C = Alo.Bhi + Ahi.Blo + Ahi.Bhi
// is accumulated in FP32 by v_mfma_f32_32x32x16_bf16; the lo.lo term
// (relative size 2^-16) is dropped.
//
// One launch, no workspace: each workgroup computes one output tile
// (tile_bf16x3.h) from float32 tiles of A and B split in registers; 256 x 256
// tile, BK = 32, LDS double buffer, 8 waves each 64 x 128 (grids under 256
// such tiles take 128 x 128 or 64 x 64 tiles of 2 x 2 waves).
#include "tile_bf16x3.h"

namespace {

struct GemmArgs {
    const float *A, *B;
    float *C;
    int M, N, K, lda, ldb, ldc;
    int64_t sA, sB, sC;
    int tiles_m, tiles_n, group_m;
};

template <class G, int MODE>
__global__ __launch_bounds__(G::kThreads) void gemm_bf16x3_kernel(GemmArgs p) {
    __shared__ __attribute__((aligned(16))) uint16_t smem[2 * G::kStage];
    const int tiles = p.tiles_m * p.tiles_n;
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    const int bs = bid / tiles;
    int tm, tn;
    hbs::tile_coords(bid - bs * tiles, p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * G::BM, n0 = tn * G::BN;
    hbs::bf16x3_block_tile<G, MODE>(smem, p.A + bs * p.sA + (int64_t)m0 * p.lda,
                                    p.B + bs * p.sB + n0, p.C + bs * p.sC, p.M, p.N, p.K,
                                    p.lda, p.ldb, p.ldc, m0, n0);
}

template <class G>
int launch_gemm(const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb,
                 int ldc, int64_t sA, int64_t sB, int64_t sC, int batch, int group_m,
                 hipStream_t stream) {
    GemmArgs g{A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC,
               (M + G::BM - 1) / G::BM, (N + G::BN - 1) / G::BN, group_m};
    const int64_t gblocks = (int64_t)g.tiles_m * g.tiles_n * batch;
    if (gblocks > 0x7fffffffLL) return HBS_ERR_ARG;
    // interior tiles when the shape divides evenly, clamped loads for ragged
    // M or N (with K predicates when K is not a multiple of the stage; K a
    // multiple of 4, A 16-byte loadable), else predicated loads
    const bool vec_a = hbs::aligned16(A) && lda % 4 == 0 && sA % 4 == 0;
    const unsigned nb = (unsigned)gblocks;
    if (vec_a && K % G::BK == 0 && M % G::BM == 0 && N % G::BN == 0)
        gemm_bf16x3_kernel<G, hbs::kTileInterior><<<nb, G::kThreads, 0, stream>>>(g);
    else if (vec_a && K % G::BK == 0)
        gemm_bf16x3_kernel<G, hbs::kTileClamp><<<nb, G::kThreads, 0, stream>>>(g);
    else if (vec_a && K % 4 == 0)
        gemm_bf16x3_kernel<G, hbs::kTileClampK><<<nb, G::kThreads, 0, stream>>>(g);
    else
        gemm_bf16x3_kernel<G, hbs::kTileEdge><<<nb, G::kThreads, 0, stream>>>(g);
    return hbs::launch_status();
}

// The 256 x 256 production tile unless 128 x 128 tiles of 2 x 2 waves (the
// same main loop) finish in fewer weighted rounds (a round of them measured at
// 0.66 of a 256 x 256 round, N = 4096); grids of under 128 such tiles take
// 64 x 64 tiles.
int run_bf16x3(const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb,
               int ldc, int64_t sA, int64_t sB, int64_t sC, int batch, hipStream_t stream) {
    const int64_t big_tiles = (int64_t)((M + 255) / 256) * ((N + 255) / 256) * batch;
    const int64_t mid_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128) * batch;
    if (hbs::prefer_big_tiles(big_tiles, mid_tiles, hbs::cu_count(), 0.66))
        return launch_gemm<hbs::B3Main16>(A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC, batch, 8,
                                        stream);
    if (mid_tiles >= 128)
        return launch_gemm<hbs::B3Cfg<128, 128, 32, 2, 2>>(A, B, C, M, N, K, lda, ldb, ldc, sA,
                                                          sB, sC, batch, 8, stream);
    return launch_gemm<hbs::B3Cfg<64, 64, 32, 2, 2>>(A, B, C, M, N, K, lda, ldb, ldc, sA, sB, sC,
                                                    batch, 8, stream);
}

}  // namespace

extern "C" int hbs_gemm_bf16x3(const float *A, const float *B, float *C, int M, int N, int K,
                               int lda, int ldb, int ldc, void *stream) {
    if (M < 1 || N < 1 || K < 1 || lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(M, N, K, lda, ldb, ldc)) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    const int st = hbs::check_device();
    if (st) return st;
    return run_bf16x3(A, B, C, M, N, K, lda, ldb, ldc, 0, 0, 0, 1, (hipStream_t)stream);
}

extern "C" int hbs_sgemm_bf16x3(const float *A, const float *B, float *C, int N, int niter,
                                void *stream) {
    if (N < 1 || niter < 1 || N > hbs::kMaxDim) return HBS_ERR_ARG;
    const int64_t nn = (int64_t)N * N;
    if (nn > ((int64_t)1 << 62) / niter) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    const int st = hbs::check_device();
    if (st) return st;
    return run_bf16x3(A, B, C, N, N, N, N, N, N, nn, nn, nn, niter, (hipStream_t)stream);
}
