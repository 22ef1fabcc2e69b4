/*
 * hbsgemm.h -- C-ABI of libhbsgemm.so, the MI355X (gfx950) implementation of
 * HammerBench's apps/sgemm path.
 *
 * Boundary it replaces (reference, apps/sgemm):
 *   device entry  extern "C" int kernel(float *mat1, float *mat2, float *result,
 *                                       int pod_id)            kernel.cpp:279-280
 *   host runtime  hb_mc_device_malloc / transfer_data_to_device /
 *                 kernel_enqueue / pods_kernels_execute        main.cpp:61-143
 * On one GPU the runtime calls become HIP allocation, copies and stream order,
 * owned by the caller; this library only launches work on a caller's stream.
 *
 * Conventions (all entry points):
 *  - matrices are row-major float32 in device memory owned by the caller;
 *  - every call is asynchronous on `stream` (a hipStream_t, NULL = default
 *    stream) unless stated otherwise; the caller synchronises, as
 *    hb_mc_device_pods_kernels_execute blocks in main.cpp:98;
 *  - the return value is HBS_SUCCESS (0), a negative HBS_ERR_* code for an
 *    invalid argument (nothing is launched), or a positive HIP runtime error
 *    code -- the counterpart of HB_MC_SUCCESS / BSG_CUDA_CALL (main.cpp:61-148);
 *  - no entry point allocates device memory on the GEMM path; scratch space
 *    is passed in as a workspace whose size the *_workspace_size call returns
 *    (grouped plans allocate once, when they are created, and
 *    hbs_gemm_fp8res_noalloc uses a fixed scratch array in the library);
 *  - calls on distinct streams are independent, and the entry points may be
 *    called from several host threads at once (the library's per-device
 *    state is locked; grouped graphs are captured in thread-local mode);
 *    calls sharing a workspace, a grouped plan or the allocation-free
 *    scratch must be ordered by the caller.
 */
#ifndef HBSGEMM_H
#define HBSGEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HBS_SUCCESS = 0,
    HBS_ERR_ARG = -1,         /* size or argument out of range */
    HBS_ERR_NULL = -2,        /* null pointer */
    HBS_ERR_ALIGN = -3,       /* pointer or leading dimension misaligned */
    HBS_ERR_WORKSPACE = -4,   /* workspace missing or too small */
    HBS_ERR_DEVICE = -5,      /* current device is not a gfx950 GPU */
    HBS_ERR_UNSUPPORTED = -6, /* valid request this build does not handle */
    HBS_ERR_SCRATCH = -7,     /* problem needs more than an allocation-free entry
                                 point's fixed scratch */
    HBS_ERR_IN_USE = -8       /* grouped plan still replayed by a graph */
};

/* Every entry point rejects M, N, K or a leading dimension above
 * HBS_MAX_DIM with HBS_ERR_ARG (tile arithmetic is done in 32-bit ints). */
#define HBS_MAX_DIM (0x7fffffff - 4096)

/* Human-readable text for any status code (never NULL). */
const char *hbs_strerror(int code);
/* Library version, major * 10000 + minor * 100 + patch. */
int hbs_version(void);

/* ------------------------------------------------------------------------ */
/* Plain FP32 SGEMM on FP32 matrix cores.                                     */
/* ------------------------------------------------------------------------ */

/* Replaces kernel(mat1, mat2, result, pod_id), apps/sgemm/kernel.cpp:279-333:
 * C_i = A_i . B_i for i < niter, each an N x N row-major slice stored back to
 * back with stride N*N (kernel.cpp:293-297), all in one launch.  N and NITER
 * are compile-time macros in the reference (template.mk:30-31,54-55); here
 * they are arguments.  Any N >= 1 is accepted (the reference needs N % 16 == 0,
 * kernel.cpp:9-10).  Requires niter >= 1 and niter * N * N < 2^62. */
int hbs_sgemm(const float *A, const float *B, float *C, int N, int niter, void *stream);

/* Rectangular, strided generalisation (apps/legacy/gemm/sgemm_data_parallel
 * main.cpp:42-74 uses M x N x P tensors with row strides):
 * C_b[M x N] = A_b[M x K] . B_b[K x N] for b < batch, row-major with leading
 * dimensions lda >= K, ldb >= N, ldc >= N and element strides between batch
 * entries.  M, N, K >= 1. */
int hbs_sgemm_strided(const float *A, const float *B, float *C, int M, int N, int K,
                      int lda, int ldb, int ldc, int64_t strideA, int64_t strideB,
                      int64_t strideC, int batch, void *stream);

/* ------------------------------------------------------------------------ */
/* BF16x3 split-precision SGEMM (no reference counterpart): x = hi + lo with  */
/* hi = bf16(x), lo = bf16(x - hi); C = Ahi.Bhi + Ahi.Blo + Alo.Bhi on BF16   */
/* matrix cores with FP32 accumulation.  Same call shape as hbs_sgemm.        */
/* ------------------------------------------------------------------------ */

/* C_i = A_i . B_i for i < niter N x N slices, as hbs_sgemm, on BF16 matrix
 * cores.  One launch; the split happens in registers, so no workspace and no
 * allocation.  Asynchronous on `stream`. */
int hbs_sgemm_bf16x3(const float *A, const float *B, float *C, int N, int niter, void *stream);
/* Rectangular form: C[M x N] = A[M x K] . B[K x N], row-major with leading
 * dimensions lda >= K, ldb >= N, ldc >= N. */
int hbs_gemm_bf16x3(const float *A, const float *B, float *C, int M, int N, int K, int lda,
                    int ldb, int ldc, void *stream);


/* ------------------------------------------------------------------------ */
/* Grouped persistent GEMM (no reference counterpart): many independent        */
/* problems of uneven M/N/K, FP32-MFMA or BF16x3 arithmetic, one launch.       */
/* Generalises the reference's NITER equal slices in one launch               */
/* (apps/sgemm/kernel.cpp:293-297) to uneven problems.                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    const float *A; /* M x K, row-major, leading dimension lda             */
    const float *B; /* K x N, row-major, leading dimension ldb             */
    float *C;       /* M x N, row-major, leading dimension ldc (written)   */
    int M, N, K;
    int lda, ldb, ldc;
} hbs_gemm_problem;

enum { HBS_PREC_FP32 = 0, HBS_PREC_BF16X3 = 1 };
/* Tile configurations of the grouped kernel (BM x BN, waves, K stage).
 *  - 256-thread persistent launch, 4 waves, two workgroups per CU, optional
 *    split-K: 0..3 (2 x 2 waves; K stage 16 for FP32, 32 for BF16x3),
 *    HBS_TILE_128x128_4x1 (FP32 only: waves of 32 x 128), HBS_TILE_128x128_BK32
 *    (FP32 only: 2 x 2 waves, K stage 32);
 *  - 512-thread large-tile launches, 8 waves, one workgroup per CU, whole
 *    problems (no split-K), tiles chained through one pipeline, K stage 32:
 *    HBS_TILE_256x256 (the default; edge tiles spread their valid part over
 *    the SIMDs), HBS_TILE_128x128_8W, HBS_TILE_256x128, HBS_TILE_128x256.
 * Each large-tile shape has one launch for problems whose K is a multiple of
 * 32 (clamped loads) and one for other K (zero fill past K); problems whose
 * operands allow no 16-byte loads run on a 256 x 256 launch with per-element
 * predicated loads. */
enum {
    HBS_TILE_128x128 = 0,
    HBS_TILE_64x64 = 1,
    HBS_TILE_128x64 = 2,
    HBS_TILE_64x128 = 3,
    HBS_TILE_256x256 = 4,
    HBS_TILE_128x128_4x1 = 5,
    HBS_TILE_128x128_BK32 = 6,
    HBS_TILE_128x128_8W = 7,
    HBS_TILE_256x128 = 8,
    HBS_TILE_128x256 = 9
};

typedef struct hbs_grouped_plan hbs_grouped_plan;

typedef struct {
    int problems;
    int workgroups;       /* persistent workgroups of one run (both launches) */
    int64_t tiles;        /* output tiles over all problems                 */
    int64_t work_items;   /* tiles x K splits                               */
    int64_t split_tiles;  /* tiles computed with split-K                    */
    double est_imbalance; /* max / mean modelled workgroup load             */
} hbs_grouped_info;

/* Builds a plan for `count` problems (host array; the matrices are device
 * pointers that must stay valid while the plan is used).  Requirements per
 * problem: 1 <= M, N, K <= HBS_MAX_DIM, lda >= K, ldb >= N, ldc >= N; any
 * alignment.  Tile shape and split-K come from `tuning_file` (written by the
 * autotuner) when it is non-NULL -- HBS_ERR_ARG if it cannot be read -- for
 * the size classes it names; every other problem takes HBS_TILE_256x256.
 * Problems whose A or B is not 16-byte aligned, or whose N, K, lda or ldb is
 * not a multiple of 4, always take HBS_TILE_256x256 (on the launch with
 * per-element predicated loads).  Allocates and uploads the plan's device
 * buffers (synchronous). */
int hbs_grouped_plan_create(const hbs_gemm_problem *problems, int count, int precision,
                            const char *tuning_file, hbs_grouped_plan **plan);
/* Same with one tile configuration and split-K factor (1..64) for every
 * problem (the autotuner's probe; split-K applies to the 256-thread tiles
 * only).  HBS_ERR_ARG for a configuration the precision does not have. */
int hbs_grouped_plan_create_forced(const hbs_gemm_problem *problems, int count, int precision,
                                   int tile, int split_k, hbs_grouped_plan **plan);
int hbs_grouped_plan_info(const hbs_grouped_plan *plan, hbs_grouped_info *info);
/* Dry run of plan creation, no reference counterpart: fills `info` with the
 * counts a plan for these problems would have on a GPU of `cus` compute units,
 * without a device, device memory or any HIP call (capacity planning, and
 * tests of the planner on machines without a GPU).  tile = -1 takes the
 * tuning file when `tuning_file` is non-NULL, else the default tile, as
 * hbs_grouped_plan_create does; tile >= 0 forces tile and split_k as
 * hbs_grouped_plan_create_forced does (then `tuning_file` must be NULL).  The
 * problems are validated as for a real plan; their pointers are not
 * dereferenced. */
int hbs_grouped_plan_describe(const hbs_gemm_problem *problems, int count, int precision,
                              const char *tuning_file, int tile, int split_k, int cus,
                              hbs_grouped_info *info);
/* Computes every problem of the plan on `stream`: the large-tile launch (when
 * the plan has 256 x 256 tiles), then the 256-thread persistent launch (when
 * it has smaller tiles).  Workgroups take their items from queues in plan
 * memory, which reset themselves at the end of each launch.  Both precisions
 * read the operands in place; BF16x3 tiles of the large-tile launch run on
 * 16 x 16 x 32 MFMAs, so they round differently from hbs_gemm_bf16x3 (same
 * error bounds).  Deterministic; runs of one plan must not overlap (they
 * share the queues, split-K counters and partial buffers). */
int hbs_grouped_run(const hbs_grouped_plan *plan, void *stream);
/* Frees the plan's device buffers (synchronous).  Returns HBS_ERR_IN_USE and
 * frees nothing while a graph made from the plan by hbs_grouped_graph_create
 * has not been destroyed with hbs_graph_destroy. */
int hbs_grouped_plan_destroy(hbs_grouped_plan *plan);

/* Captures one hbs_grouped_run of `plan` into an executable hipGraph
 * (returned as hipGraphExec_t in *graph_exec); replay with hbs_graph_launch.
 * The graph holds the plan's device buffers (problem table, queues, split-K
 * counters and partials), so the plan must outlive it: destroy every graph of
 * a plan before the plan.  Replays of one plan's graphs, and hbs_grouped_run
 * of that plan, must not overlap. */
int hbs_grouped_graph_create(const hbs_grouped_plan *plan, void **graph_exec);
int hbs_graph_launch(void *graph_exec, void *stream);
int hbs_graph_destroy(void *graph_exec);

/* ------------------------------------------------------------------------ */
/* FP8-e4m3 GEMM with per-tile scaling and a BF16 residual (no reference      */
/* counterpart).  Per 128 x 128 tile of A and of B, with amax its max |x|     */
/* over the tile's finite entries:                                            */
/*   e = power-of-two exponent with 224 < amax / 2^e <= 448,                  */
/*   q = 2^e * e4m3_rne(x / 2^e),  r = bf16_rne(x - q);                       */
/*   C = Aq.Bq (FP8 matrix cores, E8M0 tile scales) + Aq.Br + Ar.Bq (BF16),  */
/* FP32 accumulation; the Ar.Br term is dropped.                              */
/* ------------------------------------------------------------------------ */

/* Workspace bytes for an M x N x K problem.  Layout (byte offsets aligned to
 * 256; Mp = M rounded up to 256, Np = N rounded up to 256, Kp = K rounded up
 * to 128): A8 [Mp][Kp] e4m3 | Ar [Mp][Kp] bf16 | eA [Mp/128][Kp/128] int32 |
 * B8 [Np][Kp] | Br [Np][Kp] | eB [Np/128][Kp/128] (B stored transposed: row n
 * holds column n of B).  q = decode(X8) * 2^e is rebuilt on the fly. */
size_t hbs_gemm_fp8res_workspace_size(int M, int N, int K);
/* C[M x N] = A[M x K] . B[K x N] (row-major, leading dimensions lda >= K,
 * ldb >= N, ldc >= N) by the scheme above.  A NaN or infinity in row i of A
 * (column j of B) makes every output of row i (column j) of C non-finite and
 * changes no other output: tile scales ignore non-finite entries. */
int hbs_gemm_fp8res(const float *A, const float *B, float *C, int M, int N, int K, int lda,
                    int ldb, int ldc, void *workspace, size_t workspace_bytes, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hbs_gemm_fp8res with its workspace in a fixed scratch
 * array inside the library instead of a caller buffer.  Takes problems whose
 * hbs_gemm_fp8res_workspace_size(M, N, K) is at most
 * HBS_FP8RES_NOALLOC_SCRATCH_BYTES (64 MiB, e.g. M = N = 4096 with K <= 2688)
 * and returns HBS_ERR_SCRATCH for larger ones.  Calls no allocating routine;
 * each GPU the library runs on holds its own copy of the scratch until the
 * process ends.  Every call on one GPU uses that GPU's one scratch, so calls
 * that overlap on one GPU (different streams, or host threads) race and must be
 * ordered by the caller.  Results are identical to hbs_gemm_fp8res's. */
#define HBS_FP8RES_NOALLOC_SCRATCH_BYTES (64u << 20)
int hbs_gemm_fp8res_noalloc(const float *A, const float *B, float *C, int M, int N, int K,
                            int lda, int ldb, int ldc, void *stream);

/* ------------------------------------------------------------------------ */
/* Stream-K GEMM (no reference counterpart): the rectangular FP32 or BF16x3   */
/* product for grids of output tiles that leave CUs idle.                     */
/* ------------------------------------------------------------------------ */

/* Workspace bytes of hbs_gemm_streamk for an M x N x K product in `precision`
 * (HBS_PREC_FP32 or HBS_PREC_BF16X3) on the current device: a counter region
 * of 256 KiB, then two 256 KiB partial tiles per workgroup (at most one
 * workgroup per CU: at most 128.25 MiB on a 256-CU MI355X).  0 for invalid
 * arguments or for products of more than 65536 output tiles of 256 x 256
 * (those fill the GPU many times over; hbs_sgemm_strided and hbs_gemm_bf16x3
 * serve them). */
size_t hbs_gemm_streamk_workspace_size(int precision, int M, int N, int K);
/* C[M x N] = A[M x K] . B[K x N] (row-major, leading dimensions lda >= K,
 * ldb >= N, ldc >= N) with FP32-MFMA (HBS_PREC_FP32) or BF16x3
 * (HBS_PREC_BF16X3) arithmetic, the tiles and main loops of hbs_sgemm_strided
 * and hbs_gemm_bf16x3 (BF16x3 here on 16 x 16 x 32 MFMAs instead of
 * 32 x 32 x 16), scheduled stream-K: G workgroups (G <= CUs) each take
 * an equal run of (256 x 256 tile, 32-deep K stage) units, and a tile shared
 * by several workgroups is summed from their partials in K order by the last
 * to finish.  G minimises the longest run plus the measured cost of summing
 * the most-split tile.  Pays where those entry points launch fewer 256 x 256
 * tiles than CUs, most over a deep K (M = N = 2048, K = 16384: 1.1x FP32,
 * 1.5x BF16x3; M = N = 2560 or 3072, K >= M); loses up to 12% on small
 * cubes (M = N = K <= 2048).  Results differ from theirs in rounding order
 * only (the split tiles' sums, and in BF16x3 the MFMA shape), are
 * deterministic, and meet the same error bounds plus one addition per extra
 * split.  The workspace (hbs_gemm_streamk_workspace_size bytes, 16-byte
 * aligned) must be all zero before its first use; every call leaves its
 * counter region zero again, so it can be reused by later calls of any shape
 * and precision whose size fits, but calls sharing it must be ordered. */
int hbs_gemm_streamk(int precision, const float *A, const float *B, float *C, int M, int N, int K,
                     int lda, int ldb, int ldc, void *workspace, size_t workspace_bytes,
                     void *stream);

/* ------------------------------------------------------------------------ */
/* Input generation on the device.                                            */
/* ------------------------------------------------------------------------ */

/* apps/sgemm/main.cpp:46-52: mat1[i] = i % 7, mat2[i] = i % 3 for global index
 * i = first .. first + n - 1 (written to mat1[0..n), mat2[0..n)). */
int hbs_fill_pattern(float *mat1, float *mat2, int64_t n, int64_t first, void *stream);

/* Seeded counter-based inputs: element j is a function of (seed, first + j).
 * kind HBS_GEN_UNIFORM: uniform[-1,1); HBS_GEN_LOGUNIFORM: |x| = 10^(-3+6u)
 * with random sign; HBS_GEN_NORMAL_OUTLIER: normal(0,1) with 1% of entries
 * scaled by 1e3. */
enum { HBS_GEN_UNIFORM = 0, HBS_GEN_LOGUNIFORM = 1, HBS_GEN_NORMAL_OUTLIER = 2 };
int hbs_fill_random(float *dst, int64_t n, int kind, uint64_t seed, uint64_t first,
                    void *stream);

/* ------------------------------------------------------------------------ */
/* FP64 verification on the GPU (the validation of main.cpp:102-135 with an   */
/* FP64 reference instead of the float host golden).  Synchronous.           */
/* ------------------------------------------------------------------------ */
typedef struct {
    double e_cw;        /* max_ij |C - R| / (|A||B|)_ij                 */
    double e_nw;        /* max_ij |C - R| / max_ij |R|                  */
    double e_F;         /* ||C - R||_F / ||R||_F                        */
    double max_abs_err; /* max_ij |C - R|                               */
    double max_abs_ref; /* max_ij |R|                                   */
    double sum_sq_err;  /* sum_ij (C - R)^2                             */
    double sum_sq_ref;  /* sum_ij R^2                                   */
    int64_t nonfinite;  /* entries of C that are Inf or NaN             */
} hbs_errstats;

/* R = A.B and (|A||B|) computed in FP64 on the device for rows
 * [row_begin, row_end) of C (row_end <= M; pass 0, M for all), compared with C.
 * Blocks the calling thread until the statistics are in *out. */
int hbs_verify_f64(const float *A, const float *B, const float *C, int M, int N, int K,
                   int lda, int ldb, int ldc, int row_begin, int row_end,
                   hbs_errstats *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
