// hbs_internal.h -- shared host/device helpers for libhbsgemm (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hbsgemm.h"

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

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

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
