// gemm_fp8res.hip -- FP8-e4m3 GEMM with per-tile scaling and a BF16 residual.
//
// No reference counterpart: a precision variant of apps/sgemm's C = A.B
// (kernel.cpp:279-333 call shape, generalised to M x N x K).
//
// Quantisation (per 128 x 128 tile T of A, and of B, with amax = max |x| over
// the finite entries of T):
//   e    = the power-of-two exponent with 224 < amax / 2^e <= 448
//   q(x) = 2^e * e4m3_rne(x / 2^e)                 (exact in bf16)
//   r(x) = bf16_rne(x - q(x))                      (x - q(x) is exact)
// Product: C = Aq.Bq + Aq.Br + Ar.Bq, FP32 accumulate; Ar.Br is dropped.
//   Aq.Bq runs on the FP8 matrix cores with the tile exponents applied as
//   E8M0 block scales by v_mfma_scale_f32_32x32x64_f8f6f4; the two residual
//   products run on v_mfma_f32_32x32x16_bf16.  Power-of-two scales keep small
//   integers exact, so the reference pattern (0..6 x 0..2) stays bit-exact.
//
// Launches: quantise A, quantise B (transposed, K-contiguous), GEMM.  The
// workspace holds, per operand X (rows R = M or N, padded to Rp, a multiple of
// 256; Kp = K padded to 128):  X8 [Rp][Kp] e4m3 bytes, Xr [Rp][Kp] bf16 residuals, and the tile
// exponents eX [Rp/128][Kp/128] int32 (for B indexed [n tile][k tile]) -- three
// bytes per element.  The bf16 copy of q that the residual products need is
// rebuilt from the e4m3 bytes in registers (v_cvt_scalef32_pk_bf16_fp8, exact:
// q has four significant bits and the scale is a power of two).
#include <mutex>

#include "hbs_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kQT = 128;  // quantisation tile edge

__device__ __forceinline__ int tile_exponent(float amax) {
    if (!(amax >= 0x1p-100f)) return 0;  // all-zero or negligible tile: scale 1
    const uint32_t u = __float_as_uint(amax);
    const int E = (int)((u >> 23) & 0xff) - 126;  // amax = f * 2^E, f in [0.5, 1)
    const float f = __uint_as_float((u & 0x807fffffu) | (126u << 23));
    return (f * 512.0f > 448.0f) ? E - 8 : E - 9;
}

struct QuantArgs {
    const float *X;
    int rows, cols, ld;  // logical matrix as stored (A: M x K; B: K x N)
    int vec;             // X and ld allow 16-byte loads
    int Rp, Kp;          // padded output rows / K
    uint8_t *X8;
    uint16_t *Xr;
    int *ex;
};

// One workgroup of 256 threads per 128 x 128 quantisation tile, held in
// registers: load i of wave w covers tile rows 32w + 2i + lane/32 and columns
// 4 (lane % 32) .. +3 of the stored matrix (512 contiguous bytes per row).
// TRANS = false: stored rows are output rows (A).  TRANS = true: X is K x N
// and output rows are its columns (B); the tile is transposed through LDS so
// that each thread then holds 4 consecutive k of one output row, as for A.
// Conversions: v_cvt_pk_fp8_f32 (round to nearest even to OCP e4m3, |x| <= 448
// after scaling), v_cvt_scalef32_pk_f32_fp8 (exact decode times 2^e) and
// v_cvt_pk_bf16_f32 (round to nearest even) for the residual.
template <bool TRANS>
__global__ __launch_bounds__(256) void quantize_kernel(QuantArgs a) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    __shared__ float t[TRANS ? kQT : 1][kQT + 1];
    __shared__ float red[4];
    const int tiles_k = a.Kp / kQT;
    const int tr = blockIdx.x / tiles_k, tk = blockIdx.x % tiles_k;
    const int r0 = tr * kQT, k0 = tk * kQT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c4 = (lane & 31) * 4;
    const int rsub = 32 * wave + (lane >> 5);  // tile row of load 0; load i adds 2i
    const int sr0 = TRANS ? k0 : r0, sc0 = TRANS ? r0 : k0;

    float v[16][4];
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int sr = sr0 + rsub + 2 * i, sc = sc0 + c4;
        const float *x = a.X + (int64_t)sr * a.ld + sc;
        if (a.vec && sr < a.rows && sc + 4 <= a.cols) {
            const floatx4 q = *(const floatx4 *)x;
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = q[j];
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = (sr < a.rows && sc + j < a.cols) ? x[j] : 0.0f;
        }
        // amax over the finite entries: a NaN or infinity leaves the tile's
        // scale alone and makes only its own row (A) or column (B) of C
        // non-finite
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float av = fabsf(v[i][j]);
            m = fmaxf(m, av < INFINITY ? av : 0.0f);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (lane == 0) red[wave] = m;
    if (TRANS) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) t[c4 + j][rsub + 2 * i] = v[i][j];
    }
    __syncthreads();
    if (TRANS) {
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[i][j] = t[rsub + 2 * i][c4 + j];
    }
    const float amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    const int e = tile_exponent(amax);
    if (tid == 0) a.ex[tr * tiles_k + tk] = e;
    const float inv = ldexpf(1.0f, -e), scl = ldexpf(1.0f, e);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[i][0] * inv, v[i][1] * inv, 0, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[i][2] * inv, v[i][3] * inv, w, true);
        const f2 q01 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp8(w, scl, false);
        const f2 q23 = __builtin_amdgcn_cvt_scalef32_pk_f32_fp8(w, scl, true);
        const b2 r01 = __builtin_convertvector(f2{v[i][0], v[i][1]} - q01, b2);
        const b2 r23 = __builtin_convertvector(f2{v[i][2], v[i][3]} - q23, b2);
        const int64_t o = (int64_t)(r0 + rsub + 2 * i) * a.Kp + k0 + c4;
        *(int *)(a.X8 + o) = w;
        *(uint2 *)(a.Xr + o) =
            make_uint2(__builtin_bit_cast(uint32_t, r01), __builtin_bit_cast(uint32_t, r23));
    }
}

// GEMM: 256 x 128 output tile, 8 waves as 4 x 2 of 64 x 64, BK = 64 (one FP8
// MFMA step and four BF16 steps per stage), LDS double buffer.  LDS rows are
// unpadded; 16-byte chunks are XOR-swizzled by row so that the 16 lanes of a
// 128-bit read touch 16 distinct bank groups.  Each stage also carries the
// three tile exponents it needs (two A row tiles, one B column tile), loaded
// with the stage's data, so that no step waits on a separate memory load.
constexpr int kBM = 256, kBN = 128, kBK = 64;
constexpr int kWM = 4, kWN = 2, kThreads = kWM * kWN * 64;
constexpr int kFM = 2, kFN = 2;
constexpr int kRow8 = kBK;          // bytes per LDS row, e4m3 plane (4 chunks)
constexpr int kRowR = kBK * 2;      // bytes per LDS row, residual plane (8 chunks)
constexpr int kA8 = 0;
constexpr int kB8 = kA8 + kBM * kRow8;
constexpr int kAR = kB8 + kBN * kRow8;
constexpr int kBR = kAR + kBM * kRowR;
constexpr int kExp = kBR + kBN * kRowR;    // 3 int32 tile exponents (+ pad)
constexpr int kStage = kExp + 16;          // bytes per stage
// 16-byte chunks per thread per stage
constexpr int kCA8 = kBM * 4 / kThreads;   // 2
constexpr int kCB8 = kBN * 4 / kThreads;   // 1
constexpr int kCAR = kBM * 8 / kThreads;   // 4
constexpr int kCBR = kBN * 8 / kThreads;   // 2
static_assert(2 * kStage <= 160 * 1024, "LDS");

// swizzled byte offset of 16-byte chunk c of row r
__device__ __forceinline__ int sw8(int r, int c) { return r * kRow8 + ((c ^ ((r >> 2) & 3)) << 4); }
__device__ __forceinline__ int swr(int r, int c) { return r * kRowR + ((c ^ ((r >> 1) & 7)) << 4); }

struct GemmArgs {
    const uint8_t *A8, *B8;
    const uint16_t *Ar, *Br;
    const int *eA, *eB;
    float *C;
    int M, N, Kp, ldc;
    int tiles_m, tiles_n;
};

// 8 e4m3 bytes (two dwords) -> 8 bf16 values times 2^e
__device__ __forceinline__ bf16x8 fp8x8_to_bf16(uint32_t w0, uint32_t w1, float scale) {
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const b2 p0 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w0, scale, false);
    const b2 p1 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w0, scale, true);
    const b2 p2 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w1, scale, false);
    const b2 p3 = __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(w1, scale, true);
    return bf16x8{p0[0], p0[1], p1[0], p1[1], p2[0], p2[1], p3[0], p3[1]};
}

template <bool EDGE>
__global__ __launch_bounds__(kThreads) void gemm_fp8res_kernel(GemmArgs p) {
    __shared__ __attribute__((aligned(16))) uint8_t smem[2 * kStage];
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    int tm, tn;
    hbs::tile_coords(bid, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * kBM, n0 = tn * kBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % kWM, wn = wave / kWM;
    const int l32 = lane & 31, h = lane >> 5;
    const int ktiles = p.Kp / kQT;

    floatx16 acc[kFM][kFN];
#pragma unroll
    for (int r = 0; r < kFM; ++r)
#pragma unroll
        for (int c = 0; c < kFN; ++c) acc[r][c] = floatx16{0};

    // One stage in registers between its global loads and its store to LDS.
    // Two of them are in flight, so a stage's loads are issued two steps
    // before its data is stored to LDS: one step of MFMAs does not cover the
    // loads' latency (72 KB per workgroup per stage), two do.
    struct Stage {
        floatx4 a8[kCA8], b8[kCB8], ar[kCAR], br[kCBR];
        int ex;
    };
    const int KT = p.Kp / kBK;
    // lanes 0, 1: exponents of the two 128-row A tiles; lane 2: the B column tile
    const int *exp_src = tid < 2 ? p.eA + (int64_t)(m0 / kQT + tid) * ktiles
                                 : p.eB + (int64_t)(n0 / kQT) * ktiles;
    // Loads are unconditional (a stage past the end re-reads the last one,
    // and every lane reads an exponent; lanes 0..2 keep theirs), so that the
    // register sets stay in registers.
    auto gload = [&](Stage &r, int kt) {
        const int k0 = min(kt, KT - 1) * kBK;
        r.ex = exp_src[k0 / kQT];
#pragma unroll
        for (int i = 0; i < kCA8; ++i) {
            const int f = tid + i * kThreads, row = f >> 2, c = f & 3;
            r.a8[i] = *(const floatx4 *)(p.A8 + (int64_t)(m0 + row) * p.Kp + k0 + c * 16);
        }
#pragma unroll
        for (int i = 0; i < kCB8; ++i) {
            const int f = tid + i * kThreads, row = f >> 2, c = f & 3;
            r.b8[i] = *(const floatx4 *)(p.B8 + (int64_t)(n0 + row) * p.Kp + k0 + c * 16);
        }
#pragma unroll
        for (int i = 0; i < kCAR; ++i) {
            const int f = tid + i * kThreads, row = f >> 3, c = f & 7;
            r.ar[i] = *(const floatx4 *)(p.Ar + (int64_t)(m0 + row) * p.Kp + k0 + c * 8);
        }
#pragma unroll
        for (int i = 0; i < kCBR; ++i) {
            const int f = tid + i * kThreads, row = f >> 3, c = f & 7;
            r.br[i] = *(const floatx4 *)(p.Br + (int64_t)(n0 + row) * p.Kp + k0 + c * 8);
        }
    };
    auto sstore = [&](const Stage &r, int buf) {
        uint8_t *st = smem + buf * kStage;
        if (tid < 3) ((int *)(st + kExp))[tid] = r.ex;
#pragma unroll
        for (int i = 0; i < kCA8; ++i) {
            const int f = tid + i * kThreads, row = f >> 2, c = f & 3;
            *(floatx4 *)(st + kA8 + sw8(row, c)) = r.a8[i];
        }
#pragma unroll
        for (int i = 0; i < kCB8; ++i) {
            const int f = tid + i * kThreads, row = f >> 2, c = f & 3;
            *(floatx4 *)(st + kB8 + sw8(row, c)) = r.b8[i];
        }
#pragma unroll
        for (int i = 0; i < kCAR; ++i) {
            const int f = tid + i * kThreads, row = f >> 3, c = f & 7;
            *(floatx4 *)(st + kAR + swr(row, c)) = r.ar[i];
        }
#pragma unroll
        for (int i = 0; i < kCBR; ++i) {
            const int f = tid + i * kThreads, row = f >> 3, c = f & 7;
            *(floatx4 *)(st + kBR + swr(row, c)) = r.br[i];
        }
    };

    const int arow = wm * 64 + l32, brow = wn * 64 + l32;
    auto compute = [&](int buf) {
        const uint8_t *st = smem + buf * kStage;
        const int ea = ((const int *)(st + kExp))[wm >> 1];
        const int eb = ((const int *)(st + kExp))[2];
        // e4m3 fragments: lane half h holds k = 32h .. 32h + 31 (chunks 2h,
        // 2h + 1); used by the FP8 product and, converted, by the residual one
        intx8 a8[kFM], b8[kFN];
#pragma unroll
        for (int r = 0; r < kFM; ++r) {
            const int row = arow + r * 32;
            const floatx4 lo = *(const floatx4 *)(st + kA8 + sw8(row, 2 * h));
            const floatx4 hi = *(const floatx4 *)(st + kA8 + sw8(row, 2 * h + 1));
            a8[r] = intx8{__float_as_int(lo[0]), __float_as_int(lo[1]), __float_as_int(lo[2]),
                          __float_as_int(lo[3]), __float_as_int(hi[0]), __float_as_int(hi[1]),
                          __float_as_int(hi[2]), __float_as_int(hi[3])};
        }
#pragma unroll
        for (int c = 0; c < kFN; ++c) {
            const int row = brow + c * 32;
            const floatx4 lo = *(const floatx4 *)(st + kB8 + sw8(row, 2 * h));
            const floatx4 hi = *(const floatx4 *)(st + kB8 + sw8(row, 2 * h + 1));
            b8[c] = intx8{__float_as_int(lo[0]), __float_as_int(lo[1]), __float_as_int(lo[2]),
                          __float_as_int(lo[3]), __float_as_int(hi[0]), __float_as_int(hi[1]),
                          __float_as_int(hi[2]), __float_as_int(hi[3])};
        }
        // BF16 residual: Ar.Bq + Aq.Br, four 16-deep steps.  In step s lane
        // half h supplies k = 32h + 8s .. +7 for both operands -- a permutation
        // of the BF16 instruction's k slots applied to A and B alike, so the
        // products pair up unchanged -- which lets q come from the e4m3 bytes
        // already in a8 / b8 (dwords 2s, 2s + 1) and the residual from chunk
        // 4h + s.
        const float sa = ldexpf(1.0f, ea), sb = ldexpf(1.0f, eb);
#pragma unroll
        for (int s = 0; s < kBK / 16; ++s) {
            bf16x8 aq[kFM], ar[kFM], bq[kFN], br[kFN];
#pragma unroll
            for (int r = 0; r < kFM; ++r) {
                ar[r] = *(const bf16x8 *)(st + kAR + swr(arow + r * 32, 4 * h + s));
                aq[r] = fp8x8_to_bf16(a8[r][2 * s], a8[r][2 * s + 1], sa);
            }
#pragma unroll
            for (int c = 0; c < kFN; ++c) {
                br[c] = *(const bf16x8 *)(st + kBR + swr(brow + c * 32, 4 * h + s));
                bq[c] = fp8x8_to_bf16(b8[c][2 * s], b8[c][2 * s + 1], sb);
            }
#pragma unroll
            for (int c = 0; c < kFN; ++c) {
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[r], bq[c], acc[r][c], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aq[r], br[c], acc[r][c], 0, 0, 0);
            }
        }
        // FP8: Aq.Bq with the tile exponents as E8M0 block scales, after the
        // residual steps (that order measured 2% faster than before them);
        // consecutive MFMAs share their second operand
#pragma unroll
        for (int c = 0; c < kFN; ++c)
#pragma unroll
            for (int r = 0; r < kFM; ++r)
                acc[r][c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    a8[r], b8[c], acc[r][c], 0, 0, 0, ea + 127, 0, eb + 127);
    };

    // Pipeline: stage kt is in LDS buffer kt & 1; at the start of step kt the
    // register sets hold stages kt + 1 and kt + 2, in flight.  Step kt computes
    // from LDS, stores stage kt + 1 to the other buffer (waiting for its
    // loads, issued two steps earlier), meets the barrier and issues the loads
    // of stage kt + 3 into the set just stored.  The loop is unrolled by two so
    // that each register set has a fixed role.
    Stage s0, s1;
    gload(s0, 0);
    gload(s1, 1);
    sstore(s0, 0);
    __syncthreads();
    gload(s0, 2);
    for (int kt = 0; kt < KT; kt += 2) {
        compute(0);  // stage kt; s1 = stage kt + 1, s0 = stage kt + 2
        if (kt + 1 < KT) sstore(s1, 1);
        __syncthreads();
        if (kt + 1 >= KT) break;
        gload(s1, kt + 3);
        compute(1);  // stage kt + 1; s0 = stage kt + 2, s1 = stage kt + 3
        if (kt + 2 < KT) sstore(s0, 0);
        __syncthreads();
        gload(s0, kt + 4);
    }

    if (!EDGE || (m0 + kBM <= p.M && n0 + kBN <= p.N))
        hbs::store_acc<kFM, kFN, false>(p.C, p.ldc, m0 + wm * 64, n0 + wn * 64, p.M, p.N, kFN, acc);
    else
        hbs::store_acc<kFM, kFN, true>(p.C, p.ldc, m0 + wm * 64, n0 + wn * 64, p.M, p.N, kFN, acc);
}

// Wide GEMM: 256 x 256 output tile, 8 waves as 4 x 2 of 64 x 128, BK = 64,
// chosen over the 256 x 128 form by use_wide below.  A third fewer bytes are
// loaded per flop than with the 256 x 128 tile.  Its stages do not fit twice
// in the 160 KB LDS (96 KB each), so the e4m3 planes have one LDS buffer and
// the residual planes two:
//   e4m3 buffer   A8 [256][64 B], B8 [256][64 B]          32 KB
//   residual x 2  Ar [256][128 B], Br [256][128 B]     2 x 64 KB
// The operands go from the workspace straight into LDS
// (global_load_lds_dwordx4: lane l of a load writes bytes 16 l .. 16 l + 15
// of a 1 KB run), so no registers hold a stage in flight.  Each load picks
// its source chunk so that the runs land in the swizzled layout above (sw8,
// swr).  Step kt:
//   1. loads the residual planes of stage kt + 1 into the other residual
//      buffer and the exponents of stage kt + 1 into a register;
//   2. reads its e4m3 fragments from the e4m3 buffer, runs residual step 0;
//   3. barrier (every wave holds its e4m3 fragments), then loads the e4m3
//      planes of stage kt + 1 into the e4m3 buffer;
//   4. runs residual steps 1-3 and the FP8 products;
//   5. waits for its loads, barrier.
// The MFMAs of each 32 x 32 block run in the same order as in the 256 x 128
// kernel, so both give the same bits.
constexpr int kWBN = 256, kWFN = 4;                  // tile width, column blocks per wave
constexpr int kWA8 = 0, kWB8 = kWA8 + kBM * kRow8;   // e4m3 buffer
constexpr int kWRes = kWB8 + kWBN * kRow8;           // first residual buffer
constexpr int kWAR = 0, kWBR = kWAR + kBM * kRowR;   // within a residual buffer
constexpr int kWRStage = kWBR + kWBN * kRowR;
constexpr int kWLds = kWRes + 2 * kWRStage;
static_assert(kWLds <= 160 * 1024, "LDS");

// one 16-byte lane of a global -> LDS run (the LDS address is the run's start)
__device__ __forceinline__ void dma16(const void *g, void *lds) {
    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)g,
                                     (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}

template <bool EDGE>
__global__ __launch_bounds__(kThreads) void gemm_fp8res_wide_kernel(GemmArgs p) {
    __shared__ __attribute__((aligned(16))) uint8_t smem[kWLds];
    const int bid = hbs::xcd_remap(blockIdx.x, gridDim.x);
    int tm, tn;
    hbs::tile_coords(bid, p.tiles_m, p.tiles_n, 8, tm, tn);
    const int m0 = tm * kBM, n0 = tn * kWBN;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % 4, wn = wave / 4;
    const int l32 = lane & 31, h = lane >> 5;
    const int ktiles = p.Kp / kQT;
    const int KT = p.Kp / kBK;

    floatx16 acc[kFM][kWFN];
#pragma unroll
    for (int r = 0; r < kFM; ++r)
#pragma unroll
        for (int c = 0; c < kWFN; ++c) acc[r][c] = floatx16{0};

    // Sources of this lane's loads.  e4m3 planes: load i (0, 1) of wave w
    // fills rows 16 (2w + i) .. +15 (4 chunks a row); row & ~3 is a multiple of
    // 4, so the swizzle of lane l's chunk depends on l alone.  Residual planes:
    // load i (0..3) fills rows 8 (4w + i) .. +7 (8 chunks a row).
    const uint8_t *a8g = p.A8 + (int64_t)m0 * p.Kp;
    const uint8_t *b8g = p.B8 + (int64_t)n0 * p.Kp;
    const uint16_t *arg = p.Ar + (int64_t)m0 * p.Kp;
    const uint16_t *brg = p.Br + (int64_t)n0 * p.Kp;
    const int r8 = 32 * wave + (lane >> 2);                    // row of e4m3 load 0
    const int c8 = ((lane & 3) ^ ((lane >> 4) & 3)) * 16;      // its byte column
    const int rr = 32 * wave + (lane >> 3);                    // row of residual load 0
    auto dma_q = [&](int kt) {  // e4m3 planes of stage kt -> e4m3 buffer
        const int k0 = kt * kBK;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t o = (int64_t)(r8 + 16 * i) * p.Kp + k0 + c8;
            dma16(a8g + o, smem + kWA8 + (2 * wave + i) * 1024);
            dma16(b8g + o, smem + kWB8 + (2 * wave + i) * 1024);
        }
    };
    auto dma_r = [&](int kt, int buf) {  // residual planes of stage kt -> buffer buf
        const int k0 = kt * kBK;
        uint8_t *rs = smem + kWRes + buf * kWRStage;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int cr = ((lane & 7) ^ ((4 * (i & 1) + (lane >> 4)) & 7)) * 8;
            const int64_t o = (int64_t)(rr + 8 * i) * p.Kp + k0 + cr;
            dma16(arg + o, rs + kWAR + (4 * wave + i) * 1024);
            dma16(brg + o, rs + kWBR + (4 * wave + i) * 1024);
        }
    };
    // lanes 0, 1: exponents of the two 128-row A tiles, lanes 2, 3: of the two
    // 128-column B tiles; the other lanes repeat lane 0's load
    const int *exp_src = lane == 1 ? p.eA + (int64_t)(m0 / kQT + 1) * ktiles
                         : lane == 2 ? p.eB + (int64_t)(n0 / kQT) * ktiles
                         : lane == 3 ? p.eB + (int64_t)(n0 / kQT + 1) * ktiles
                                     : p.eA + (int64_t)(m0 / kQT) * ktiles;
    auto eload = [&](int kt) { return hbs::ldg(exp_src + (kt * kBK) / kQT); };

    const int arow = wm * 64 + l32, brow = wn * 128 + l32;
    auto compute = [&](int buf, int ex, auto &&mid) {
        const uint8_t *rs = smem + kWRes + buf * kWRStage;
        const int ea = __builtin_amdgcn_readlane(ex, wm >> 1);
        const int eb = __builtin_amdgcn_readlane(ex, 2 + wn);
        intx8 a8[kFM], b8[kWFN];
#pragma unroll
        for (int r = 0; r < kFM; ++r) {
            const int row = arow + r * 32;
            const floatx4 lo = *(const floatx4 *)(smem + kWA8 + sw8(row, 2 * h));
            const floatx4 hi = *(const floatx4 *)(smem + kWA8 + sw8(row, 2 * h + 1));
            a8[r] = intx8{__float_as_int(lo[0]), __float_as_int(lo[1]), __float_as_int(lo[2]),
                          __float_as_int(lo[3]), __float_as_int(hi[0]), __float_as_int(hi[1]),
                          __float_as_int(hi[2]), __float_as_int(hi[3])};
        }
#pragma unroll
        for (int c = 0; c < kWFN; ++c) {
            const int row = brow + c * 32;
            const floatx4 lo = *(const floatx4 *)(smem + kWB8 + sw8(row, 2 * h));
            const floatx4 hi = *(const floatx4 *)(smem + kWB8 + sw8(row, 2 * h + 1));
            b8[c] = intx8{__float_as_int(lo[0]), __float_as_int(lo[1]), __float_as_int(lo[2]),
                          __float_as_int(lo[3]), __float_as_int(hi[0]), __float_as_int(hi[1]),
                          __float_as_int(hi[2]), __float_as_int(hi[3])};
        }
        const float sa = ldexpf(1.0f, ea), sb = ldexpf(1.0f, eb);
#pragma unroll
        for (int s = 0; s < kBK / 16; ++s) {
            if (s == 1) mid();
            bf16x8 aq[kFM], ar[kFM];
#pragma unroll
            for (int r = 0; r < kFM; ++r) {
                ar[r] = *(const bf16x8 *)(rs + kWAR + swr(arow + r * 32, 4 * h + s));
                aq[r] = fp8x8_to_bf16(a8[r][2 * s], a8[r][2 * s + 1], sa);
            }
#pragma unroll
            for (int c = 0; c < kWFN; ++c) {
                const bf16x8 br = *(const bf16x8 *)(rs + kWBR + swr(brow + c * 32, 4 * h + s));
                const bf16x8 bq = fp8x8_to_bf16(b8[c][2 * s], b8[c][2 * s + 1], sb);
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ar[r], bq, acc[r][c], 0, 0, 0);
#pragma unroll
                for (int r = 0; r < kFM; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(aq[r], br, acc[r][c], 0, 0, 0);
            }
        }
#pragma unroll
        for (int c = 0; c < kWFN; ++c)
#pragma unroll
            for (int r = 0; r < kFM; ++r)
                acc[r][c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                    a8[r], b8[c], acc[r][c], 0, 0, 0, ea + 127, 0, eb + 127);
    };

    dma_q(0);
    dma_r(0, 0);
    int ex = eload(0);
    __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0)
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
        const bool more = kt + 1 < KT;
        int exn = ex;
        if (more) {
            dma_r(kt + 1, (kt + 1) & 1);
            exn = eload(kt + 1);
        }
        compute(kt & 1, ex, [&]() {
            // every wave holds its e4m3 fragments of stage kt.  A bare
            // barrier after this wave's LDS reads: __syncthreads() would also
            // wait for the residual loads in flight (they write LDS).
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_barrier();
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (more) dma_q(kt + 1);
        });
        __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0): stage kt + 1 is in LDS
        __syncthreads();
        ex = exn;
    }

    if (!EDGE || (m0 + kBM <= p.M && n0 + kWBN <= p.N))
        hbs::store_acc<kFM, kWFN, false>(p.C, p.ldc, m0 + wm * 64, n0 + wn * 128, p.M, p.N, kWFN,
                                         acc);
    else
        hbs::store_acc<kFM, kWFN, true>(p.C, p.ldc, m0 + wm * 64, n0 + wn * 128, p.M, p.N, kWFN,
                                        acc);
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct Layout {
    int Mp, Np, Kp;
    size_t a8, ar, ea, b8, br, eb, total;  // byte offsets
};

// M, N, K <= hbs::kMaxDim, so the padded sizes fit in int.  B is padded to
// a multiple of the wide tile's 256 columns whichever form runs: the 256 x 128
// form reads only its first round_up(N, 128) rows of the B planes.
Layout layout(int M, int N, int K) {
    Layout L;
    L.Mp = (int)round_up(M, kBM);
    L.Np = (int)round_up(N, kWBN);
    L.Kp = (int)round_up(K, kQT);
    auto al = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    size_t o = 0;
    L.a8 = o; o = al(o + (size_t)L.Mp * L.Kp);
    L.ar = o; o = al(o + (size_t)L.Mp * L.Kp * 2);
    L.ea = o; o = al(o + (size_t)(L.Mp / kQT) * (L.Kp / kQT) * 4);
    L.b8 = o; o = al(o + (size_t)L.Np * L.Kp);
    L.br = o; o = al(o + (size_t)L.Np * L.Kp * 2);
    L.eb = o; o = al(o + (size_t)(L.Np / kQT) * (L.Kp / kQT) * 4);
    L.total = o;
    return L;
}

int check_args(const float *A, const float *B, const float *C, int M, int N, int K, int lda,
               int ldb, int ldc) {
    if (M < 1 || N < 1 || K < 1 || lda < K || ldb < N || ldc < N) return HBS_ERR_ARG;
    if (!hbs::dims_ok(M, N, K, lda, ldb, ldc)) return HBS_ERR_ARG;
    if (!A || !B || !C) return HBS_ERR_NULL;
    return HBS_SUCCESS;
}

// The three launches of one product: quantise A, quantise B, GEMM.  `ws`
// holds L.total bytes, 16-byte aligned; every byte the GEMM reads is written
// by the quantisers first (padding rows and columns as zeros), so nothing in
// it needs clearing beforehand.
// Whether the 256 x 256 form runs: always when round_up(N, 128) is already a
// multiple of 256; otherwise when its rounds of one workgroup per CU, each
// costing 1.75 rounds of the 256 x 128 form (its tile does twice the work at
// 0.52 against 0.46 of the roofline; 1.72 measured at 16384 x 16512 x 4096),
// take less time than the 256 x 128 form's rounds.  Measured at M = 16384,
// K = 4096 (wide against 256 x 128 time): N = 384 0.79x and 1152 0.92x
// (256 x 128 kept), 2176 1.00x, 640 1.12x, 4224 1.06x, 16512 1.13x (wide).
bool use_wide(const Layout &L, int N) {
    const int64_t narrow_w = round_up(N, kBN);
    if (narrow_w % kWBN == 0) return true;
    const int64_t cus = hbs::cu_count(), tm = L.Mp / kBM;
    const int64_t narrow_rounds = (tm * (narrow_w / kBN) + cus - 1) / cus;
    const int64_t wide_rounds = (tm * (L.Np / kWBN) + cus - 1) / cus;
    return 7 * wide_rounds < 4 * narrow_rounds;
}

int run_fp8res(const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb,
               int ldc, const Layout &L, char *ws, hipStream_t s) {
    const int64_t qa_blocks = (int64_t)(L.Mp / kQT) * (L.Kp / kQT);
    const int64_t qb_blocks = (int64_t)(L.Np / kQT) * (L.Kp / kQT);
    const int narrow_tn = (int)(round_up(N, kBN) / kBN);
    const int64_t blocks = (int64_t)(L.Mp / kBM) * narrow_tn;
    if (qa_blocks > 0x7fffffffLL || qb_blocks > 0x7fffffffLL || blocks > 0x7fffffffLL)
        return HBS_ERR_ARG;
    QuantArgs qa{A, M, K, lda, hbs::aligned16(A) && lda % 4 == 0, L.Mp, L.Kp,
                 (uint8_t *)(ws + L.a8), (uint16_t *)(ws + L.ar), (int *)(ws + L.ea)};
    quantize_kernel<false><<<(unsigned)qa_blocks, 256, 0, s>>>(qa);
    int rc = hbs::launch_status();
    if (rc) return rc;
    QuantArgs qb{B, K, N, ldb, hbs::aligned16(B) && ldb % 4 == 0, L.Np, L.Kp,
                 (uint8_t *)(ws + L.b8), (uint16_t *)(ws + L.br), (int *)(ws + L.eb)};
    quantize_kernel<true><<<(unsigned)qb_blocks, 256, 0, s>>>(qb);
    rc = hbs::launch_status();
    if (rc) return rc;

    GemmArgs g{(const uint8_t *)(ws + L.a8), (const uint8_t *)(ws + L.b8),
               (const uint16_t *)(ws + L.ar), (const uint16_t *)(ws + L.br),
               (const int *)(ws + L.ea), (const int *)(ws + L.eb), C, M, N, L.Kp, ldc,
               L.Mp / kBM, narrow_tn};
    if (use_wide(L, N)) {
        g.tiles_n = L.Np / kWBN;
        const unsigned wblocks = (unsigned)(g.tiles_m * g.tiles_n);
        if (M % kBM == 0 && N % kWBN == 0)
            gemm_fp8res_wide_kernel<false><<<wblocks, kThreads, 0, s>>>(g);
        else
            gemm_fp8res_wide_kernel<true><<<wblocks, kThreads, 0, s>>>(g);
    } else if (M % kBM == 0 && N % kBN == 0) {
        gemm_fp8res_kernel<false><<<(unsigned)blocks, kThreads, 0, s>>>(g);
    } else {
        gemm_fp8res_kernel<true><<<(unsigned)blocks, kThreads, 0, s>>>(g);
    }
    return hbs::launch_status();
}

// Scratch of the allocation-free entry point: the workspace of one product of
// up to HBS_FP8RES_NOALLOC_SCRATCH_BYTES, held by each GPU the library runs
// on until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: kernels get its device address, looked up once
// per device with hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[HBS_FP8RES_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
char *scratch_ptr[kMaxDevices];

// Device address of the current device's scratch (looked up on first use).
int scratch_address(char **out) {
    int dev = 0;
    HBS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return HBS_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        HBS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p || !hbs::aligned16(p)) return HBS_ERR_DEVICE;
        scratch_ptr[dev] = (char *)p;
    }
    *out = scratch_ptr[dev];
    return HBS_SUCCESS;
}

}  // namespace

extern "C" size_t hbs_gemm_fp8res_workspace_size(int M, int N, int K) {
    if (M < 1 || N < 1 || K < 1) return 0;
    if (M > hbs::kMaxDim || N > hbs::kMaxDim || K > hbs::kMaxDim) return 0;
    return layout(M, N, K).total;
}

extern "C" int hbs_gemm_fp8res(const float *A, const float *B, float *C, int M, int N, int K,
                               int lda, int ldb, int ldc, void *workspace,
                               size_t workspace_bytes, void *stream) {
    int st = check_args(A, B, C, M, N, K, lda, ldb, ldc);
    if (st) return st;
    const Layout L = layout(M, N, K);
    if (!workspace || workspace_bytes < L.total) return HBS_ERR_WORKSPACE;
    if (!hbs::aligned16(workspace)) return HBS_ERR_ALIGN;
    st = hbs::check_device();
    if (st) return st;
    return run_fp8res(A, B, C, M, N, K, lda, ldb, ldc, L, (char *)workspace, (hipStream_t)stream);
}

extern "C" int hbs_gemm_fp8res_noalloc(const float *A, const float *B, float *C, int M, int N,
                                       int K, int lda, int ldb, int ldc, void *stream) {
    int st = check_args(A, B, C, M, N, K, lda, ldb, ldc);
    if (st) return st;
    const Layout L = layout(M, N, K);
    if (L.total > (size_t)HBS_FP8RES_NOALLOC_SCRATCH_BYTES) return HBS_ERR_SCRATCH;
    st = hbs::check_device();
    if (st) return st;
    char *ws = nullptr;
    st = scratch_address(&ws);
    if (st) return st;
    return run_fp8res(A, B, C, M, N, K, lda, ldb, ldc, L, ws, (hipStream_t)stream);
}
