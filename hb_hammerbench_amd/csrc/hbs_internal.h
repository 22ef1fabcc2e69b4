// hbs_internal.h -- shared host/device helpers for libhbsgemm (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hbsgemm.h"
#include "hbs_host.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

#define HBS_HIP_TRY(expr)                                   \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return (int)_e;               \
    } while (0)

namespace hbs {

// Load modes of a block tile (tile_f32.h, tile_bf16x3.h):
//   kTileInterior  the tile lies inside M x N, K is a multiple of the stage
//                  depth and the operands allow 16-byte loads: no predicates;
//   kTileEdge      any shape and alignment: per-element predicates, zero fill;
//   kTileClamp     ragged M or N, with K a multiple of the stage depth and
//                  16-byte-loadable operands: rows and columns beyond M or N
//                  are read from the last valid one (clamped addresses at full
//                  load width) and only the stores are predicated;
//   kTileClampK    kTileClamp with K any multiple of 4: whole 4-k groups
//                  beyond K are also zero.
enum TileMode { kTileInterior = 0, kTileEdge = 1, kTileClamp = 2, kTileClampK = 3 };

// Loads and stores of matrix elements in device memory, typed as global
// accesses: an operand pointer read from memory (a grouped problem table)
// would otherwise be accessed with flat instructions, which also count against
// the LDS wait counter and so make LDS fragment reads wait for global loads.
template <class T>
__device__ __forceinline__ T ldg(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <bool GLOBAL, class T>
__device__ __forceinline__ T ld(const T *p) {
    if constexpr (GLOBAL)
        return ldg(p);
    else
        return *p;
}
template <class T>
__device__ __forceinline__ void stg(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}

// Where a lane's accumulator registers sit in C.  A wave's sub-tile is held
// as FM x FN regions of 32 x 32, one floatx16 per region, in one of two
// layouts:
//   Q16 = false (v_mfma_f32_32x32x*): lane l holds column (l & 31) and, in
//     register q, row 8 (q >> 2) + 4 (l >> 5) + (q & 3);
//   Q16 = true (four v_mfma_f32_16x16x* blocks, block s = q >> 2 at rows
//     16 (s & 1), columns 16 (s >> 1) of the region): lane l holds column
//     16 (q >> 3) + (l & 15) and row 16 ((q >> 2) & 1) + 4 (l >> 4) + (q & 3).
// acc_lane_row / acc_lane_col: the lane's own offset; acc_row / acc_col: the
// wave-uniform offset of register q.
template <bool Q16>
__device__ __forceinline__ int acc_lane_row(int lane) {
    return Q16 ? 4 * (lane >> 4) : 4 * (lane >> 5);
}
template <bool Q16>
__device__ __forceinline__ int acc_lane_col(int lane) {
    return Q16 ? (lane & 15) : (lane & 31);
}
template <bool Q16>
__host__ __device__ constexpr int acc_row(int q) {
    return Q16 ? 16 * ((q >> 2) & 1) + (q & 3) : 8 * (q >> 2) + (q & 3);
}
template <bool Q16>
__host__ __device__ constexpr int acc_col(int q) {
    return Q16 ? 16 * (q >> 3) : 0;
}

// Stores a wave's FM x FN accumulator regions (layout Q16, above) to C at the
// wave's sub-tile origin (rb, cb), row stride ldc.  The lane's own address is
// formed once; each store adds a wave-uniform offset to it, so a store costs
// one 64-bit add and no 64-bit multiply.  PRED: only rows < M and columns < N
// are written.  Column regions from ncols on are skipped (a half-width wave).
// (A 16-byte-store form, each group of four registers transposed across its
// lane quad by DPP, was bit-identical but 2.5% slower on the grouped BF16x3
// group: its 16 extra VALU instructions per four values cost more than the
// three quarters of the store instructions they save.)
template <int FM, int FN, bool PRED, bool Q16 = false>
__device__ __forceinline__ void store_acc(float *__restrict__ C, int ldc, int rb, int cb, int M,
                                          int N, int ncols, const floatx16 (&acc)[FM][FN]) {
    const int lane = threadIdx.x & 63;
    const int lr = acc_lane_row<Q16>(lane), lc = acc_lane_col<Q16>(lane);
    float *lanep = C + ((int64_t)rb + lr) * ldc + cb + lc;
    const int vm = M - rb - lr, vn = N - cb - lc;  // valid rows / columns from the lane's own
#pragma unroll
    for (int r = 0; r < FM; ++r) {
#pragma unroll
        for (int c = 0; c < FN; ++c) {
            if (c >= ncols) continue;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int dr = r * 32 + acc_row<Q16>(q), dc = c * 32 + acc_col<Q16>(q);
                if (!PRED || (dr < vm && dc < vn))
                    stg(lanep + ((int64_t)dr * ldc + dc), acc[r][c][q]);
            }
        }
    }
}

// Operands of one block tile's main loop: Ag = A + m0 * lda (the tile's first
// row), Bg = B + n0 (its first column); M, N, K bound the predicates.
struct TileArgs {
    const float *Ag, *Bg;
    int M, N, K, lda, ldb, m0, n0;
};

// Cross-tile pipelining of a workgroup that runs block tiles back to back
// (the grouped kernel's large-tile launch).  A main loop given a chain with
// kOn set, on a tile of at least three K stages:
//   * calls publish() before the barrier that ends step 0 (the caller makes
//     the next tile's identity visible in LDS there) and peek(next) after it,
//     which returns whether there is a next tile and fills its operands;
//   * issues the next tile's stage-0 loads in the second-to-last step and, in
//     the last step, stores that stage to the free LDS buffer and issues the
//     stage-1 loads, then meets one more barrier;
//   * returns pre = true and buf0 = that buffer: the next main loop skips its
//     prologue (stage 0 already in LDS, stage 1 in the load registers).
// On a shorter tile the loop runs unchained and ends with publish(), a barrier
// and peek(next).  Every thread of the workgroup takes the same branches.
// kGlobal: the operand pointers were read from memory (a problem table), so
// the loads are typed as global (ldg); kernel-argument pointers need no cast.
struct NoChain {
    static constexpr bool kOn = false, kGlobal = false;
};
struct NoChainGlobal {
    static constexpr bool kOn = false, kGlobal = true;
};
template <class Publish, class Peek>
struct TileChain {
    static constexpr bool kOn = true, kGlobal = true;
    Publish publish;
    Peek peek;
    bool pre;       // in: this tile's stages 0 and 1 are already staged
    int buf0;       // in: LDS buffer of stage 0; out: the next tile's
    bool has_next;  // out: peek found a next tile (its operands in next)
    TileArgs next;
};
template <class Publish, class Peek>
__device__ __forceinline__ TileChain<Publish, Peek> make_chain(Publish p, Peek k) {
    return TileChain<Publish, Peek>{p, k, false, 0, false, TileArgs{}};
}

// Wave arrangement of a 256 x 256 tile of 8 waves (4 x 2 sub-tiles of 64 x 128;
// waves w and w + 4 share a SIMD).  Interior tiles, and edge tiles whose valid
// rows and columns still reach all four 64-row and all four 64-column units,
// use the fixed arrangement of wave_pos (the tile configs in tile_f32.h and
// tile_bf16x3.h).  Any other edge tile is covered by the mu x nu units of
// 64 x 64 that hold valid rows and columns: each wave takes a piece of one
// unit row -- two units (64 x 128, kWaveFull), one unit (64 x 64, kWaveHalf)
// or nothing (kWaveNone) -- so that no SIMD holds more than ceil(mu nu / 4)
// units, the least possible, with the fewest pieces.  The arrangements below
// came from an exhaustive search over every split of the unit rows and every
// assignment of the pieces to the 8 waves; edge_arrangements_valid() checks
// them at compile time.
enum WaveKind { kWaveNone = 0, kWaveFull = 1, kWaveHalf = 2 };
struct WaveTile {
    int om, on;  // origin of the wave's sub-tile in the tile (elements)
    int kind;    // WaveKind
};

// mu x nu valid units (1..4 each, not both 4): byte w = the piece of wave w,
// (unit row << 4) | (first unit column << 2) | WaveKind.
__host__ __device__ constexpr uint64_t edge_arrangement(int mu, int nu) {
    switch ((mu - 1) * 4 + nu - 1) {
        case 0: return 0x0000000000000002ull;   // 1 x 1
        case 1: return 0x0000000000000602ull;   // 1 x 2
        case 2: return 0x00000000000a0602ull;   // 1 x 3
        case 3: return 0x000000000e0a0602ull;   // 1 x 4
        case 4: return 0x0000000000001202ull;   // 2 x 1
        case 5: return 0x0000000016120602ull;   // 2 x 2
        case 6: return 0x0000001200150502ull;   // 2 x 3
        case 7: return 0x0000000019110901ull;   // 2 x 4
        case 8: return 0x0000000000221202ull;   // 3 x 1
        case 9: return 0x0000000000211101ull;   // 3 x 2
        case 10: return 0x0025150500221202ull;  // 3 x 3
        case 11: return 0x292119110e0a0602ull;  // 3 x 4
        case 12: return 0x0000000032221202ull;  // 4 x 1
        case 13: return 0x0000000031211101ull;  // 4 x 2
        case 14: return 0x3525150532221202ull;  // 4 x 3
        default: return 0;
    }
}

// Every valid unit covered exactly once, every piece inside the valid units,
// and no SIMD with more than ceil(mu nu / 4) units.
constexpr bool edge_arrangements_valid() {
    for (int mu = 1; mu <= 4; ++mu) {
        for (int nu = 1; nu <= 4; ++nu) {
            if (mu == 4 && nu == 4) continue;
            int cover[4][4] = {};
            int load[4] = {};
            const uint64_t a = edge_arrangement(mu, nu);
            for (int w = 0; w < 8; ++w) {
                const int b = (int)((a >> (8 * w)) & 0xff);
                const int kind = b & 3, um = b >> 4, un = (b >> 2) & 3;
                if (kind == kWaveNone) {
                    if (b != 0) return false;
                    continue;
                }
                if (kind != kWaveFull && kind != kWaveHalf) return false;
                const int n = kind == kWaveFull ? 2 : 1;
                if (um >= mu || un + n > nu) return false;
                for (int j = 0; j < n; ++j) ++cover[um][un + j];
                load[w % 4] += n;
            }
            for (int r = 0; r < mu; ++r)
                for (int c = 0; c < nu; ++c)
                    if (cover[r][c] != 1) return false;
            for (int q = 0; q < 4; ++q)
                if (load[q] > (mu * nu + 3) / 4) return false;
        }
    }
    return true;
}
static_assert(edge_arrangements_valid(), "edge-tile wave arrangements");

// The sub-tile of wave `wave` (wave-uniform) in the 256 x 256 tile at
// (m0, n0) of an M x N product, for a tile that may cross M or N.
__device__ __forceinline__ WaveTile big_wave_tile(int M, int N, int m0, int n0, int wave) {
    const int vm = min(M - m0, 256), vn = min(N - n0, 256);
    const int mu = (vm + 63) >> 6, nu = (vn + 63) >> 6;
    if (mu == 4 && nu == 4) {
        const int wn = wave >> 2, wm = (wave + 2 * wn) & 3;
        return WaveTile{wm * 64, wn * 128, kWaveFull};
    }
    const int b = (int)(edge_arrangement(mu, nu) >> (8 * wave)) & 0xff;
    return WaveTile{(b >> 4) * 64, ((b >> 2) & 3) * 64, b & 3};
}

// Checks that the current device is a gfx950 GPU (cached per device).
int check_device();

// Compute units of the current device (cached per device; 256 on MI355X).
int cu_count();

// Whether a square or rectangular GEMM should take 256 x 256 tiles (one
// workgroup per CU) rather than 128 x 128 tiles (two per CU): the one that
// finishes in less time counted in rounds of resident workgroups, a round of
// 128 x 128 tiles costing `mid_round` of a round of 256 x 256 tiles.
inline bool prefer_big_tiles(int64_t big_tiles, int64_t mid_tiles, int cus, double mid_round) {
    const int64_t big_rounds = (big_tiles + cus - 1) / cus;
    const int64_t mid_rounds = (mid_tiles + 2 * cus - 1) / (2 * cus);
    return (double)big_rounds <= mid_round * (double)mid_rounds;
}

// Returns the launch error of the last kernel launch as a status code.
inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? HBS_SUCCESS : (int)e;
}

// Workgroup id remap so that each XCD (workgroups are dealt to the 8 XCDs
// round-robin by id) receives one contiguous range of logical tile ids, which
// keeps the operand panels a group of neighbouring tiles shares in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    constexpr int kXcd = 8;
    const int per = nblocks / kXcd;
    const int full = per * kXcd;
    if (bid >= full) return bid;  // tail that does not divide evenly
    return (bid % kXcd) * per + bid / kXcd;
}

// Grouped raster: logical tile id -> (tile row, tile col) walking `group`
// tile-rows column by column, so concurrently resident tiles share A panels
// (same rows) and B panels (same cols).
__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int group, int &tm,
                                            int &tn) {
    const int per_group = group * tiles_n;
    const int g = t / per_group;
    const int first_m = g * group;
    const int gm = min(tiles_m - first_m, group);
    const int r = t - g * per_group;
    tm = first_m + r % gm;
    tn = r / gm;
}

}  // namespace hbs
