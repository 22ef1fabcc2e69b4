// sgemm_f32.hip -- plain FP32 SGEMM on gfx950 FP32 matrix cores.
//
// Replaces the HammerBlade SPMD kernel of apps/sgemm/kernel.cpp:279-333 (16x16
// output blocks strided over a 16x8 tile group, DRAM->scratchpad block loads
// at :36-109, a 4x4 register-blocked fmaf micro-kernel at :186-275).  The
// MI355X design instead gives each workgroup one BM x BN output tile
// (tile_f32.h), with
//   * XCD-aware workgroup order plus a grouped raster for L2 reuse;
//   * NITER slices (kernel.cpp:293-297) as extra tiles of the same launch.
#include "tile_f32.h"

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

template <int BM, int BN, int BK, int WM, int WN, int MODE>
__global__ __launch_bounds__(WM *WN * 64) void sgemm_f32_kernel(GemmArgs p) {
    using G = hbs::F32Cfg<BM, BN, BK, WM, WN>;
    __shared__ __attribute__((aligned(16))) float smem[G::kLdsFloats];
    const int tiles = p.tiles_m * p.tiles_n;
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    const int bslice = bid / tiles;
    int tm, tn;
    hbs::tile_coords(bid - bslice * tiles, p.tiles_m, p.tiles_n, p.group_m, tm, tn);
    const int m0 = tm * BM, n0 = tn * BN;
    hbs::f32_block_tile<BM, BN, BK, WM, WN, MODE>(
        smem, p.A + bslice * p.sA + (int64_t)m0 * p.lda, p.B + bslice * p.sB + n0,
        p.C + bslice * p.sC, p.M, p.N, p.K, p.lda, p.ldb, p.ldc, m0, n0);
}

// Interior tiles when the shape divides evenly; ragged M or N with aligned
// operands take the clamped loads (with K predicates when K is not a multiple
// of BK, K and N multiples of 4); anything else the predicated ones.
template <int BM, int BN, int BK, int WM, int WN>
int launch_cfg(GemmArgs p, bool vec_ok, hipStream_t stream) {
    using G = hbs::F32Cfg<BM, BN, BK, WM, WN>;
    p.tiles_m = (p.M + BM - 1) / BM;
    p.tiles_n = (p.N + BN - 1) / BN;
    const int64_t blocks = (int64_t)p.tiles_m * p.tiles_n * p.batch;
    if (blocks > 0x7fffffffLL) return HBS_ERR_ARG;
    const unsigned nb = (unsigned)blocks;
    if (vec_ok && p.K % BK == 0 && p.M % BM == 0 && p.N % BN == 0)
        sgemm_f32_kernel<BM, BN, BK, WM, WN, hbs::kTileInterior><<<nb, G::kThreads, 0, stream>>>(p);
    else if (vec_ok && p.K % BK == 0 && p.N % 4 == 0)
        sgemm_f32_kernel<BM, BN, BK, WM, WN, hbs::kTileClamp><<<nb, G::kThreads, 0, stream>>>(p);
    else if (vec_ok && p.K % 4 == 0 && p.N % 4 == 0)
        sgemm_f32_kernel<BM, BN, BK, WM, WN, hbs::kTileClampK><<<nb, G::kThreads, 0, stream>>>(p);
    else
        sgemm_f32_kernel<BM, BN, BK, WM, WN, hbs::kTileEdge><<<nb, G::kThreads, 0, stream>>>(p);
    return hbs::launch_status();
}

int run_f32(GemmArgs p, hipStream_t stream) {
    const bool vec_ok = hbs::aligned16(p.A) && hbs::aligned16(p.B) && (p.lda % 4 == 0) &&
                        (p.ldb % 4 == 0) && (p.sA % 4 == 0) && (p.sB % 4 == 0);
    // 256 x 256 tiles unless 128 x 128 tiles finish in fewer weighted rounds
    // (a round of them measured at 0.52 of a 256 x 256 round, N = 4096); grids
    // of under 128 such tiles take 64 x 64 tiles.
    const int cus = hbs::cu_count();
    const int64_t big_tiles = (int64_t)((p.M + 255) / 256) * ((p.N + 255) / 256) * p.batch;
    const int64_t mid_tiles = (int64_t)((p.M + 127) / 128) * ((p.N + 127) / 128) * p.batch;
    if (hbs::prefer_big_tiles(big_tiles, mid_tiles, cus, 0.52))
        return launch_cfg<256, 256, 32, 4, 2>(p, vec_ok, stream);
    if (mid_tiles >= 128) return launch_cfg<128, 128, 16, 2, 2>(p, vec_ok, stream);
    return launch_cfg<64, 64, 16, 2, 2>(p, vec_ok, stream);
}

}  // namespace

extern "C" int hbs_sgemm_strided(const float *A, const float *B, float *C, int M, int N, int K,
                                 int lda, int ldb, int ldc, int64_t strideA, int64_t strideB,
                                 int64_t strideC, int batch, void *stream) {
    if (M < 1 || N < 1 || K < 1 || batch < 1) return HBS_ERR_ARG;
    if (lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(M, N, K, lda, ldb, ldc)) return HBS_ERR_ARG;
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
