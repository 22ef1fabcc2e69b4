// verify_f64.hip -- FP64 verification of a float32 GEMM result on the GPU.
//
// The validation step of apps/sgemm/main.cpp:102-135 compares the device
// result with a float host golden by SSE.  At N = 8192 that golden costs a
// host core tens of minutes, so the library verifies on the device instead:
// R = A.B and S = |A|.|B| are accumulated in FP64 (every float x float product
// is exact in double) and compared with C.  This is verification only; no GEMM
// entry point uses FP64.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "hbs_internal.h"

namespace {

constexpr int kTile = 64;  // output tile edge
constexpr int kKb = 16;    // k step held in LDS
constexpr int kStats = 6;  // max_cw, max_err, max_ref, sum_sq_err, sum_sq_ref, nonfinite

__global__ __launch_bounds__(256) void verify_f64_kernel(const float *__restrict__ A,
                                                         const float *__restrict__ B,
                                                         const float *__restrict__ C, int M,
                                                         int N, int K, int lda, int ldb, int ldc,
                                                         int row_begin, int row_end,
                                                         double *__restrict__ partial) {
    __shared__ float As[kKb][kTile + 1];
    __shared__ float Bs[kKb][kTile + 1];
    __shared__ double red[kStats][256];

    const int tiles_n = (N + kTile - 1) / kTile;
    const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
    const int r0 = row_begin + tm * kTile, c0 = tn * kTile;
    const int tid = threadIdx.x;
    const int tr = tid / 16, tc = tid % 16;  // 4x4 outputs per thread, strided by 16

    double R[4][4], S[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) R[i][j] = S[i][j] = 0.0;

    for (int k0 = 0; k0 < K; k0 += kKb) {
        for (int e = tid; e < kKb * kTile; e += 256) {
            const int ar = e / kKb, ak = e % kKb;  // A tile row-major over k
            const int gr = r0 + ar, gk = k0 + ak;
            As[ak][ar] = (gr < row_end && gk < K) ? A[(int64_t)gr * lda + gk] : 0.0f;
            const int bk = e / kTile, bc = e % kTile;
            const int gk2 = k0 + bk, gc = c0 + bc;
            Bs[bk][bc] = (gk2 < K && gc < N) ? B[(int64_t)gk2 * ldb + gc] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kKb; ++k) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = (double)As[k][tr + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = (double)Bs[k][tc + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    R[i][j] = fma(a[i], b[j], R[i][j]);
                    S[i][j] = fma(fabs(a[i]), fabs(b[j]), S[i][j]);
                }
        }
        __syncthreads();
    }

    double st[kStats] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gr = r0 + tr + 16 * i, gc = c0 + tc + 16 * j;
            if (gr >= row_end || gc >= N) continue;
            const float cf = C[(int64_t)gr * ldc + gc];
            if (!isfinite(cf)) {
                st[5] += 1.0;
                continue;
            }
            const double e = fabs((double)cf - R[i][j]);
            const double s = S[i][j];
            const double cw = s > 0.0 ? e / s : (e > 0.0 ? INFINITY : 0.0);
            st[0] = fmax(st[0], cw);
            st[1] = fmax(st[1], e);
            st[2] = fmax(st[2], fabs(R[i][j]));
            st[3] += e * e;
            st[4] += R[i][j] * R[i][j];
        }
    }
#pragma unroll
    for (int s = 0; s < kStats; ++s) red[s][tid] = st[s];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] = fmax(red[0][tid], red[0][tid + w]);
            red[1][tid] = fmax(red[1][tid], red[1][tid + w]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + w]);
            red[3][tid] += red[3][tid + w];
            red[4][tid] += red[4][tid + w];
            red[5][tid] += red[5][tid + w];
        }
        __syncthreads();
    }
    if (tid < kStats) partial[(int64_t)blockIdx.x * kStats + tid] = red[tid][0];
}

}  // namespace

extern "C" int hbs_verify_f64(const float *A, const float *B, const float *C, int M, int N, int K,
                              int lda, int ldb, int ldc, int row_begin, int row_end,
                              hbs_errstats *out, void *stream) {
    if (M < 1 || N < 1 || K < 1 || lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(M, N, K, lda, ldb, ldc)) return HBS_ERR_ARG;
    if (row_begin < 0 || row_end > M || row_begin >= row_end) return HBS_ERR_ARG;
    if (!A || !B || !C || !out) return HBS_ERR_NULL;
    const int st = hbs::check_device();
    if (st) return st;
    const int rows = row_end - row_begin;
    const int64_t blocks = (int64_t)((rows + kTile - 1) / kTile) * ((N + kTile - 1) / kTile);
    if (blocks > 0x7fffffffLL) return HBS_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    double *partial = nullptr;
    const size_t bytes = (size_t)blocks * kStats * sizeof(double);
    HBS_HIP_TRY(hipMalloc(&partial, bytes));
    verify_f64_kernel<<<(unsigned)blocks, 256, 0, s>>>(A, B, C, M, N, K, lda, ldb, ldc, row_begin,
                                                       row_end, partial);
    int rc = hbs::launch_status();
    std::vector<double> host((size_t)blocks * kStats);
    if (rc == HBS_SUCCESS) {
        hipError_t e = hipMemcpyAsync(host.data(), partial, bytes, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = (int)e;
    }
    hipError_t fe = hipFree(partial);
    if (rc != HBS_SUCCESS) return rc;
    if (fe != hipSuccess) return (int)fe;

    hbs_errstats r;
    memset(&r, 0, sizeof r);
    double nonfinite = 0.0;
    for (int64_t b = 0; b < blocks; ++b) {
        const double *p = &host[(size_t)b * kStats];
        r.e_cw = p[0] > r.e_cw ? p[0] : r.e_cw;
        r.max_abs_err = p[1] > r.max_abs_err ? p[1] : r.max_abs_err;
        r.max_abs_ref = p[2] > r.max_abs_ref ? p[2] : r.max_abs_ref;
        r.sum_sq_err += p[3];
        r.sum_sq_ref += p[4];
        nonfinite += p[5];
    }
    r.nonfinite = (int64_t)nonfinite;
    r.e_nw = r.max_abs_ref > 0.0 ? r.max_abs_err / r.max_abs_ref : r.max_abs_err;
    r.e_F = r.sum_sq_ref > 0.0 ? sqrt(r.sum_sq_err / r.sum_sq_ref) : sqrt(r.sum_sq_err);
    *out = r;
    return HBS_SUCCESS;
}
