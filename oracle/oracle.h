/*
 * oracle.h -- CPU restatement of HammerBench's apps/sgemm path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and never as the thing that is measured or shipped.  The product library
 * (libhbsgemm.so) does not link it and has no CPU fallback.
 *
 * Every function names the reference file:line it restates.  Functions with
 * no reference counterpart (generators, FP64 oracle, BF16/FP8 emulation) are
 * restatements of this project's own published algorithms and are pinned by
 * closed forms and library routines in tests/test_oracle.py.
 */
#ifndef HBS_ORACLE_H
#define HBS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* apps/sgemm/main.cpp:15-26 -- golden host_mm for one N x N slice. */
void orc_host_mm(float *result, const float *mat1, const float *mat2, int N);
/* Same loop nest restricted to rows [row_begin, row_end) (timing samples). */
void orc_host_mm_rows(float *result, const float *mat1, const float *mat2, int N,
                      int row_begin, int row_end);
/* Row-parallel OpenMP variant of host_mm (bit-identical per element). */
void orc_host_mm_omp(float *result, const float *mat1, const float *mat2, int N,
                     int row_begin, int row_end, int threads);

/* apps/sgemm/kernel.cpp:186-275,293-321 -- device arithmetic: one sequential
 * k = 0..N-1 chain of fmaf per output element, starting from 0.0f. */
void orc_fmaf_chain_mm(float *result, const float *mat1, const float *mat2, int N);

/* apps/sgemm/main.cpp:46-52 -- mat1[i] = i % 7, mat2[i] = i % 3 over the
 * whole NITER*N*N array (global index i, offset by `first`). */
void orc_pattern_init(float *mat1, float *mat2, int64_t first, int64_t n);
/* Closed form of C_it[y][x] for the main.cpp:46-52 pattern (exact integer). */
int64_t orc_pattern_value(int N, int it, int y, int x);
/* Fill one N x N slice of the closed form as float. */
void orc_pattern_slice(float *C, int N, int it);

/* apps/sgemm/main.cpp:120-134 -- SSE accumulated in float over n elements;
 * *n_diff receives the number of elements with diff != 0.0f. */
float orc_sse(const float *expected, const float *actual, int64_t n, int64_t *n_diff);

/* FP64 oracle: R = A.B and S = |A|.|B| in double, row-major with leading
 * dimensions.  Products of two floats are exact in double. */
void orc_gemm_f64(int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                  double *R, double *S, int threads);

/* Error metrics of C against (R, S):
 * out[0] e_cw = max |C-R| / S, out[1] e_nw = max |C-R| / max |R|,
 * out[2] e_F = ||C-R||_F / ||R||_F, out[3] max |C-R|, out[4] max |R|,
 * out[5] number of non-finite C entries. */
void orc_errors(int M, int N, const float *C, int ldc, const double *R, const double *S,
                double out[6]);

/* Counter-based generator: element i of seed's stream is the (i+1)-th output
 * of SplitMix64 seeded with `seed`. */
uint64_t orc_rng64(uint64_t seed, uint64_t i);
/* uniform[-1,1): u = (r >> 40) * 2^-24, x = 2u - 1 (exact in float). */
void orc_gen_uniform(uint64_t seed, uint64_t first, int64_t n, float *out);
/* |x| = 10^(-3 + 6u), u = (r >> 11) * 2^-53, sign = bit 0 of r. */
void orc_gen_loguniform(uint64_t seed, uint64_t first, int64_t n, float *out);
/* Box-Muller normal(0,1) from draws 3i, 3i+1; draw 3i+2 < 0.01 scales by 1e3. */
void orc_gen_normal_outlier(uint64_t seed, uint64_t first, int64_t n, float *out);

/* BF16 round-to-nearest-even of a float (bits). */
uint16_t orc_bf16_rne(float x);
float orc_bf16_to_float(uint16_t h);
/* x = hi + lo with hi = bf16(x), lo = bf16(x - hi). */
void orc_split_bf16(const float *x, int64_t n, uint16_t *hi, uint16_t *lo);

/* OCP FP8 e4m3 (bias 7, max 448) round-to-nearest-even, |x| <= 448 assumed. */
float orc_fp8_e4m3_round(float x);
uint8_t orc_fp8_e4m3_encode(float x);
float orc_fp8_e4m3_decode(uint8_t b);
/* Power-of-two tile scale exponent e (scale 2^e) with amax / 2^e <= 448 and
 * amax / 2^e > 224; 0 when amax < 2^-100 (all-zero or negligible tile). */
int orc_fp8_tile_exponent(float amax);
/* Quantise row-major rows x cols matrix X (leading dim ld) with per-tile
 * (tile x tile) power-of-two scales: Xq = 2^e * e4m3(X / 2^e), Xr = bf16(X - Xq).
 * Writes Xq, Xr as floats and the exponents per tile (row-major tile grid). */
void orc_fp8_quantize(int rows, int cols, const float *X, int ld, int tile,
                      float *Xq, float *Xr, int *exps);

#ifdef __cplusplus
}
#endif
#endif
