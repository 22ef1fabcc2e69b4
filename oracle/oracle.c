/*
 * oracle.c -- CPU restatement of HammerBench's apps/sgemm path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Never linked into
 * libhbsgemm.so.  Built with -O2 -ffp-contract=off so that every `a*b` and
 * `+` in host_mm rounds separately, as the reference's x86 host build does
 * (apps/sgemm/template.mk:33-35 passes no -mfma and no -ffast-math).
 *
 * Parity pins for every function are in tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* apps/sgemm/main.cpp:15-26                                                 */
/* ------------------------------------------------------------------------ */

void orc_host_mm_rows(float *result, const float *mat1, const float *mat2, int N,
                      int row_begin, int row_end)
{
    /* Same y -> x -> z loop nest, float accumulator, sequential z, separate
     * multiply and add (main.cpp:17-23). */
    for (int y = row_begin; y < row_end; y++) {
        for (int x = 0; x < N; x++) {
            float sum = 0.0f;
            for (int z = 0; z < N; z++) {
                sum += mat1[(int64_t)N * y + z] * mat2[(int64_t)N * z + x];
            }
            result[(int64_t)N * y + x] = sum;
        }
    }
}

void orc_host_mm(float *result, const float *mat1, const float *mat2, int N)
{
    orc_host_mm_rows(result, mat1, mat2, N, 0, N);
}

void orc_host_mm_omp(float *result, const float *mat1, const float *mat2, int N,
                     int row_begin, int row_end, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    /* Rows are independent; each row runs the exact main.cpp:17-23 nest. */
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = row_begin; y < row_end; y++) {
        orc_host_mm_rows(result, mat1, mat2, N, y, y + 1);
    }
}

/* ------------------------------------------------------------------------ */
/* apps/sgemm/kernel.cpp:186-275 (compute_block) and :293-321 (block loop)   */
/* ------------------------------------------------------------------------ */

void orc_fmaf_chain_mm(float *result, const float *mat1, const float *mat2, int N)
{
    /* reset_out() zeroes the block (kernel.cpp:24-32); each z block reloads
     * psum and continues the fmaf chain over sz = 0..15 (kernel.cpp:201-262),
     * so every element is one sequential k = 0..N-1 chain of fmaf. */
    for (int y = 0; y < N; y++) {
        for (int x = 0; x < N; x++) {
            float psum = 0.0f;
            for (int z = 0; z < N; z++) {
                psum = fmaf(mat1[(int64_t)N * y + z], mat2[(int64_t)N * z + x], psum);
            }
            result[(int64_t)N * y + x] = psum;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* apps/sgemm/main.cpp:46-52 -- input pattern and its closed-form product    */
/* ------------------------------------------------------------------------ */

void orc_pattern_init(float *mat1, float *mat2, int64_t first, int64_t n)
{
    for (int64_t k = 0; k < n; k++) {
        int64_t i = first + k;
        mat1[k] = (float)(i % 7);
        mat2[k] = (float)(i % 3);
    }
}

int64_t orc_pattern_value(int N, int it, int y, int x)
{
    /* C[y][x] = sum_z ((o + N*y + z) mod 7) * ((o + N*z + x) mod 3), o = it*N*N.
     * The first factor has period 7 in z, the second period 3, so the sum
     * runs over the 21 residues r = z mod 21 with their multiplicities. */
    int64_t o = (int64_t)it * N * N;
    int64_t c1 = (o + (int64_t)N * y) % 7;
    int64_t c2 = (o + x) % 3;
    int64_t n3 = N % 3;
    int64_t q = N / 21, rem = N % 21;
    int64_t sum = 0;
    for (int64_t r = 0; r < 21; r++) {
        int64_t cnt = q + (r < rem ? 1 : 0);
        sum += cnt * ((c1 + r) % 7) * ((c2 + n3 * r) % 3);
    }
    return sum;
}

void orc_pattern_slice(float *C, int N, int it)
{
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++)
            C[(int64_t)N * y + x] = (float)orc_pattern_value(N, it, y, x);
}

/* ------------------------------------------------------------------------ */
/* apps/sgemm/main.cpp:120-134 -- SSE validation                             */
/* ------------------------------------------------------------------------ */

float orc_sse(const float *expected, const float *actual, int64_t n, int64_t *n_diff)
{
    float sse = 0.0f;
    int64_t nd = 0;
    for (int64_t i = 0; i < n; i++) {
        float diff = expected[i] - actual[i];
        sse += diff * diff;
        if (diff != 0.0f) nd++;
    }
    if (n_diff) *n_diff = nd;
    return sse;
}

/* ------------------------------------------------------------------------ */
/* FP64 oracle and error metrics (no reference counterpart)                 */
/* ------------------------------------------------------------------------ */

void orc_gemm_f64(int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                  double *R, double *S, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; i++) {
        double *r = R + (int64_t)i * N;
        double *s = S ? S + (int64_t)i * N : NULL;
        for (int j = 0; j < N; j++) {
            r[j] = 0.0;
            if (s) s[j] = 0.0;
        }
        for (int k = 0; k < K; k++) {
            double a = (double)A[(int64_t)i * lda + k];
            double aa = fabs(a);
            const float *b = B + (int64_t)k * ldb;
            for (int j = 0; j < N; j++) r[j] += a * (double)b[j];
            if (s)
                for (int j = 0; j < N; j++) s[j] += aa * fabs((double)b[j]);
        }
    }
}

void orc_errors(int M, int N, const float *C, int ldc, const double *R, const double *S,
                double out[6])
{
    double e_cw = 0.0, max_err = 0.0, max_ref = 0.0, se = 0.0, sr = 0.0, nonfinite = 0.0;
    for (int i = 0; i < M; i++) {
        for (int j = 0; j < N; j++) {
            double c = (double)C[(int64_t)i * ldc + j];
            double r = R[(int64_t)i * N + j];
            if (!isfinite(c)) {
                nonfinite += 1.0;
                continue;
            }
            double e = fabs(c - r);
            double s = S ? S[(int64_t)i * N + j] : fabs(r);
            double cw = (s > 0.0) ? e / s : (e > 0.0 ? INFINITY : 0.0);
            if (cw > e_cw) e_cw = cw;
            if (e > max_err) max_err = e;
            if (fabs(r) > max_ref) max_ref = fabs(r);
            se += e * e;
            sr += r * r;
        }
    }
    out[0] = e_cw;
    out[1] = max_ref > 0.0 ? max_err / max_ref : max_err;
    out[2] = sr > 0.0 ? sqrt(se / sr) : sqrt(se);
    out[3] = max_err;
    out[4] = max_ref;
    out[5] = nonfinite;
}

/* ------------------------------------------------------------------------ */
/* Seeded generators (this project's own published algorithm)               */
/* ------------------------------------------------------------------------ */

#define ORC_GAMMA 0x9e3779b97f4a7c15ULL

uint64_t orc_rng64(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * ORC_GAMMA;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void orc_gen_uniform(uint64_t seed, uint64_t first, int64_t n, float *out)
{
    for (int64_t k = 0; k < n; k++) {
        uint64_t r = orc_rng64(seed, first + (uint64_t)k);
        float u = (float)(r >> 40) * 0x1p-24f;
        out[k] = 2.0f * u - 1.0f;
    }
}

void orc_gen_loguniform(uint64_t seed, uint64_t first, int64_t n, float *out)
{
    for (int64_t k = 0; k < n; k++) {
        uint64_t r = orc_rng64(seed, first + (uint64_t)k);
        double u = (double)(r >> 11) * 0x1p-53;
        double mag = exp(2.302585092994045684 * (-3.0 + 6.0 * u));
        float m = (float)mag;
        out[k] = (r & 1ULL) ? -m : m;
    }
}

void orc_gen_normal_outlier(uint64_t seed, uint64_t first, int64_t n, float *out)
{
    for (int64_t k = 0; k < n; k++) {
        uint64_t i = first + (uint64_t)k;
        uint64_t r1 = orc_rng64(seed, 3 * i), r2 = orc_rng64(seed, 3 * i + 1),
                 r3 = orc_rng64(seed, 3 * i + 2);
        double u1 = (double)((r1 >> 11) + 1) * 0x1p-53;
        double u2 = (double)(r2 >> 11) * 0x1p-53;
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586477 * u2);
        double u3 = (double)(r3 >> 11) * 0x1p-53;
        if (u3 < 0.01) z *= 1000.0;
        out[k] = (float)z;
    }
}

/* ------------------------------------------------------------------------ */
/* BF16 split (no reference counterpart)                                     */
/* ------------------------------------------------------------------------ */

static uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

uint16_t orc_bf16_rne(float x)
{
    uint32_t u = f2u(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float orc_bf16_to_float(uint16_t h) { return u2f((uint32_t)h << 16); }

void orc_split_bf16(const float *x, int64_t n, uint16_t *hi, uint16_t *lo)
{
    for (int64_t i = 0; i < n; i++) {
        uint16_t h = orc_bf16_rne(x[i]);
        float r = x[i] - orc_bf16_to_float(h); /* exact: x and hi share a binade or adjoin */
        hi[i] = h;
        lo[i] = orc_bf16_rne(r);
    }
}

/* ------------------------------------------------------------------------ */
/* FP8 e4m3 with power-of-two tile scales (no reference counterpart)         */
/* ------------------------------------------------------------------------ */

float orc_fp8_e4m3_round(float x)
{
    float a = fabsf(x);
    if (a == 0.0f) return x;
    if (a < 0x1p-6f) {
        /* subnormal range: quantum 2^-9 */
        return rintf(x * 512.0f) * 0x1p-9f;
    }
    int E;
    frexpf(a, &E); /* a = f * 2^E, f in [0.5, 1): leading bit weight 2^(E-1) */
    /* three fraction bits: quantum 2^(E-1-3) */
    return ldexpf(rintf(ldexpf(x, 4 - E)), E - 4);
}

uint8_t orc_fp8_e4m3_encode(float x)
{
    float q = orc_fp8_e4m3_round(x);
    uint8_t s = signbit(q) ? 0x80 : 0x00;
    float a = fabsf(q);
    if (a == 0.0f) return s;
    if (a < 0x1p-6f) return (uint8_t)(s | (uint8_t)(a * 512.0f));
    int E;
    float f = frexpf(a, &E);
    int ef = E - 1 + 7;
    int m = (int)((f * 2.0f - 1.0f) * 8.0f);
    return (uint8_t)(s | (ef << 3) | m);
}

float orc_fp8_e4m3_decode(uint8_t b)
{
    int s = b >> 7, ef = (b >> 3) & 15, m = b & 7;
    float v;
    if (ef == 0) v = (float)m * 0x1p-9f;
    else v = ldexpf(1.0f + (float)m / 8.0f, ef - 7);
    return s ? -v : v;
}

int orc_fp8_tile_exponent(float amax)
{
    if (!(amax >= 0x1p-100f)) return 0; /* all-zero or negligible tile: scale 1 */
    int E;
    float f = frexpf(amax, &E); /* amax = f * 2^E, f in [0.5,1) */
    /* amax / 2^(E-9) = f * 512 in [256, 512) */
    return (f * 512.0f > 448.0f) ? E - 8 : E - 9;
}

void orc_fp8_quantize(int rows, int cols, const float *X, int ld, int tile,
                      float *Xq, float *Xr, int *exps)
{
    int tr = (rows + tile - 1) / tile, tc = (cols + tile - 1) / tile;
    for (int bi = 0; bi < tr; bi++) {
        for (int bj = 0; bj < tc; bj++) {
            int r0 = bi * tile, c0 = bj * tile;
            int r1 = r0 + tile < rows ? r0 + tile : rows;
            int c1 = c0 + tile < cols ? c0 + tile : cols;
            float amax = 0.0f;
            for (int i = r0; i < r1; i++)
                for (int j = c0; j < c1; j++) {
                    float a = fabsf(X[(int64_t)i * ld + j]);
                    if (isfinite(a) && a > amax) amax = a; /* finite entries only */
                }
            int e = orc_fp8_tile_exponent(amax);
            exps[bi * tc + bj] = e;
            for (int i = r0; i < r1; i++)
                for (int j = c0; j < c1; j++) {
                    float x = X[(int64_t)i * ld + j];
                    float q = ldexpf(orc_fp8_e4m3_round(ldexpf(x, -e)), e);
                    Xq[(int64_t)i * cols + j] = q;
                    Xr[(int64_t)i * cols + j] = orc_bf16_to_float(orc_bf16_rne(x - q));
                }
        }
    }
}
