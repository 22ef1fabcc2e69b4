// hbs_host.h -- host-only helpers shared by the device sources and the grouped
// planner (no HIP headers: the planner also builds with a plain C++ compiler,
// as in the sanitizer build).
#pragma once

#include <stdint.h>

namespace hbs {

// Largest M, N, K and leading dimension any entry point accepts.  Tile and
// stage arithmetic runs on int (a tile origin plus up to 256 rows, a K range
// rounded up to a 32-deep stage or a 128-deep quantisation tile, split-K chunk
// ends), which must not overflow; larger values get HBS_ERR_ARG.
constexpr int kMaxDim = 0x7fffffff - 4096;

inline int64_t ceil_div(int64_t x, int64_t b) { return (x + b - 1) / b; }

// Whether every size and leading dimension is at most kMaxDim.
inline bool dims_ok(int M, int N, int K, int lda, int ldb, int ldc) {
    return M <= kMaxDim && N <= kMaxDim && K <= kMaxDim && lda <= kMaxDim && ldb <= kMaxDim &&
           ldc <= kMaxDim;
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace hbs
