// common.cuh -- constants and plain types shared by the host code and the
// kernels of the skycube engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sky {

constexpr int kBlock = 256;         // threads per workgroup (4 wave64)
constexpr uint32_t kPruners = 128;  // default: lowest-score points used as a prefilter
constexpr int kXCap = 2048;         // tie-touched subspaces resolved per round
constexpr int kCodePivots = 15;     // 16 quantile buckets per column, 4-bit codes

// Columns of a subspace, ascending (the reference's dimension_set).
struct DimList {
  int32_t idx[16];
};

constexpr uint32_t kCover = 2048;  // default: lowest-score candidates tested first
// The default cover of an extended skyline of m points: it grows with m, about
// m / 32 as a power of two in [kCover, 16384] (measured best on the four
// large workloads).
__host__ __device__ __forceinline__ uint32_t auto_cover(uint32_t m) {
  uint32_t c = kCover;
  while (c < 16384 && (uint64_t)c * 2 * 32 <= m) c *= 2;
  return c;
}
// Default: k_member visits the candidates through the median-mask buckets
// when they hold at least this many candidates on average (each bucket
// costs one bound test and a share of a 64-lane round), else scans them.
constexpr uint32_t kMinBucket = 16;

}  // namespace sky
