// prims.cuh -- device-wide primitives of the batched path that issue no
// memset and no host synchronisation, so that a pass built from them can be
// captured in a HIP graph: an exclusive scan, a flagged select and an LSD
// radix sort of (key, value) pairs.  Counters and histograms are cleared by
// the kernels that fill them.  Tiles are 256 threads x 16 items.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sky {
namespace {

constexpr int kPrimThreads = 256;
constexpr int kPrimItems = 16;
constexpr uint32_t kPrimTile = kPrimThreads * kPrimItems;  // 4096
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;

__host__ __device__ constexpr uint32_t prim_tiles(uint32_t n) { return (n + kPrimTile - 1) / kPrimTile; }

// Exclusive scan of the 256 values v of a workgroup (one per thread, in
// thread order); *total receives their sum.  Uses 4 words of `red`.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* red, uint32_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) red[wv] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kPrimThreads / 64; ++w) {
    const uint32_t s = red[w];
    before += w < wv ? s : 0u;
    all += s;
  }
  *total = all;
  __syncthreads();
  return before + inc - v;
}

// Scan phase 1: sum of each tile of in[0..n).
__global__ __launch_bounds__(kPrimThreads) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint32_t n,
                                                             uint32_t* __restrict__ tsum) {
  __shared__ uint32_t red[kPrimThreads / 64];
  const uint32_t base = blockIdx.x * kPrimTile;
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) {
    const uint32_t i = base + r * kPrimThreads + threadIdx.x;
    s += i < n ? in[i] : 0u;
  }
  uint32_t total;
  (void)block_excl_scan(s, red, &total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// Scan phase 2 (one workgroup): exclusive scan of the nt tile sums in place;
// *total (if given) receives the sum of all.
__global__ __launch_bounds__(kPrimThreads) void k_scan_sums(uint32_t* __restrict__ tsum, uint32_t nt,
                                                            uint32_t* __restrict__ total) {
  __shared__ uint32_t red[kPrimThreads / 64];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nt; b0 += kPrimThreads) {  // uniform
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t v = i < nt ? tsum[i] : 0u;
    uint32_t all;
    const uint32_t ex = block_excl_scan(v, red, &all);
    if (i < nt) tsum[i] = carry + ex;
    carry += all;
  }
  if (total && threadIdx.x == 0) *total = carry;
}

// Scan phase 3: out[i] = tile offset + exclusive prefix inside the tile.
// Each thread owns kPrimItems consecutive entries (read before any write, so
// out may alias in).
__global__ __launch_bounds__(kPrimThreads) void k_scan_apply(const uint32_t* in, uint32_t n,
                                                             const uint32_t* __restrict__ tsum,
                                                             uint32_t* out) {
  __shared__ uint32_t red[kPrimThreads / 64];
  const uint32_t base = blockIdx.x * kPrimTile + threadIdx.x * kPrimItems;
  uint32_t v[kPrimItems];
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) {
    v[r] = base + r < n ? in[base + r] : 0u;
    s += v[r];
  }
  uint32_t total;
  uint32_t run = tsum[blockIdx.x] + block_excl_scan(s, red, &total);
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) {
    if (base + r < n) out[base + r] = run;
    run += v[r];
  }
}

// Select phase 1: count of set flags per tile (thread-contiguous items).
__global__ __launch_bounds__(kPrimThreads) void k_sel_tiles(const uint8_t* __restrict__ flags,
                                                            uint32_t n,
                                                            uint32_t* __restrict__ tsum) {
  __shared__ uint32_t red[kPrimThreads / 64];
  const uint32_t base = blockIdx.x * kPrimTile + threadIdx.x * kPrimItems;
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) s += (base + r < n && flags[base + r]) ? 1u : 0u;
  uint32_t total;
  (void)block_excl_scan(s, red, &total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// Select phase 3: out[k] = the k-th index with a set flag, ascending.
__global__ __launch_bounds__(kPrimThreads) void k_sel_scatter(const uint8_t* __restrict__ flags,
                                                              uint32_t n,
                                                              const uint32_t* __restrict__ tsum,
                                                              uint32_t* __restrict__ out) {
  __shared__ uint32_t red[kPrimThreads / 64];
  const uint32_t base = blockIdx.x * kPrimTile + threadIdx.x * kPrimItems;
  uint32_t f = 0, s = 0;
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) {
    const bool on = base + r < n && flags[base + r];
    f |= (uint32_t)on << r;
    s += on;
  }
  uint32_t total;
  uint32_t o = tsum[blockIdx.x] + block_excl_scan(s, red, &total);
  for (uint32_t m = f; m; m &= m - 1u) out[o++] = base + (uint32_t)(__ffs(m) - 1);
}

// Segmented count of set flags: cnt[t] += flags set in [off[t], off[t+1])
// for t < T; cnt must start at zero.  Workgroup (t, y) of a grid of T x Y
// takes every Y-th stretch of kPrimThreads flags of segment t, so that a
// long segment is spread over Y workgroups.
__global__ __launch_bounds__(kPrimThreads) void k_seg_count(const uint8_t* __restrict__ flags,
                                                            const uint32_t* __restrict__ off,
                                                            uint32_t* __restrict__ cnt) {
  __shared__ uint32_t red[kPrimThreads / 64];
  const uint32_t t = blockIdx.x;
  const uint32_t step = kPrimThreads * gridDim.y;
  uint32_t s = 0;
  for (uint32_t i = off[t] + blockIdx.y * kPrimThreads + threadIdx.x; i < off[t + 1]; i += step)
    s += flags[i] ? 1u : 0u;
  uint32_t total;
  (void)block_excl_scan(s, red, &total);
  if (threadIdx.x == 0 && total) atomicAdd(cnt + t, total);
}

// Workgroups per segment for the segmented kernels: one per 16,384 items
// of the longest segment, at most 64.
inline unsigned seg_slices(uint32_t max_len) {
  const uint32_t y = (max_len + 16383u) / 16384u;
  return y < 1u ? 1u : (y > 64u ? 64u : y);
}

// Radix sort.  Tile b holds items b*kPrimTile + r*256 + t (r = round, t =
// thread), so that tile order is round-major, thread-minor.

template <typename K>
__device__ __forceinline__ uint32_t radix_digit(K key, int shift) {
  return (uint32_t)(key >> shift) & (kRadixBins - 1);
}

// hist[digit * ntiles + b] = items of tile b with that digit.
template <typename K>
__global__ __launch_bounds__(kPrimThreads) void k_rs_hist(const K* __restrict__ keys, uint32_t n,
                                                          int shift, uint32_t ntiles,
                                                          uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kRadixBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kPrimTile;
#pragma unroll
  for (int r = 0; r < kPrimItems; ++r) {
    const uint32_t i = base + r * kPrimThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[radix_digit(keys[i], shift)], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of tile b: an item goes to the scanned offset of (digit, b)
// plus the number of items of its tile with the same digit before it.  Per
// round, each wave finds the lanes sharing a lane's digit with eight
// ballots; the waves' counts per digit sit in LDS.
template <typename K>
__global__ __launch_bounds__(kPrimThreads) void k_rs_scatter(const K* __restrict__ kin,
                                                             const uint32_t* __restrict__ vin,
                                                             uint32_t n, int shift, uint32_t ntiles,
                                                             const uint32_t* __restrict__ offs,
                                                             K* __restrict__ kout,
                                                             uint32_t* __restrict__ vout) {
  constexpr int NW = kPrimThreads / 64;
  __shared__ uint32_t run[kRadixBins];
  __shared__ uint32_t wcnt[NW][kRadixBins];
  __shared__ uint32_t boff[kRadixBins];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  run[threadIdx.x] = 0;
  boff[threadIdx.x] = offs[threadIdx.x * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < NW; ++w) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kPrimTile;
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll 1
  for (int r = 0; r < kPrimItems; ++r) {
    const uint32_t i = base + r * kPrimThreads + threadIdx.x;
    const bool valid = i < n;
    K key = 0;
    uint32_t val = 0;
    if (valid) {
      key = kin[i];
      val = vin[i];
    }
    const uint32_t dg = radix_digit(key, shift);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const uint64_t on = __ballot((dg >> b) & 1u);
      peers &= ((dg >> b) & 1u) ? on : ~on;
    }
    const uint32_t lrank = (uint32_t)__popcll(peers & below);
    if (valid && lrank == 0) wcnt[wv][dg] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t p = boff[dg] + run[dg] + lrank;
#pragma unroll
      for (int w = 0; w < NW; ++w) p += w < wv ? wcnt[w][dg] : 0u;
      kout[p] = key;
      vout[p] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      add += wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = 0;
    }
    run[threadIdx.x] += add;
    __syncthreads();
  }
}

}  // namespace
}  // namespace sky

namespace skyi {
namespace {

// Temporary bytes of the primitives for n items: scan tile sums (with the
// sums of the radix histogram, 256 per tile) and the histogram itself.
inline size_t prim_tmp_bytes(uint64_t n) {
  const uint64_t nt = sky::prim_tiles((uint32_t)std::min<uint64_t>(n, 0xFFFFFFFFull)) + 1;
  const uint64_t hist = (uint64_t)sky::kRadixBins * nt;
  const uint64_t hist_tiles = sky::prim_tiles((uint32_t)hist) + 1;
  return sizeof(uint32_t) * (hist + hist_tiles + nt + 64);
}

// Exclusive scan of in[0..n) into out (may alias in); total (device, may be
// null) receives the sum.  tmp: at least prim_tmp_bytes(n) bytes.
inline hipError_t prim_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                            void* tmp, hipStream_t s) {
  uint32_t* tsum = static_cast<uint32_t*>(tmp);
  const uint32_t nt = sky::prim_tiles(n);
  if (nt) {
    hipLaunchKernelGGL(sky::k_scan_tiles, dim3(nt), dim3(sky::kPrimThreads), 0, s, in, n, tsum);
  }
  hipLaunchKernelGGL(sky::k_scan_sums, dim3(1), dim3(sky::kPrimThreads), 0, s, tsum, nt, total);
  if (nt) {
    hipLaunchKernelGGL(sky::k_scan_apply, dim3(nt), dim3(sky::kPrimThreads), 0, s, in, n, tsum, out);
  }
  return hipGetLastError();
}

// out[0..*count) = indices i < n with flags[i] != 0, ascending; *count on
// the device.
inline hipError_t prim_select(const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count,
                              void* tmp, hipStream_t s) {
  uint32_t* tsum = static_cast<uint32_t*>(tmp);
  const uint32_t nt = sky::prim_tiles(n);
  if (nt) {
    hipLaunchKernelGGL(sky::k_sel_tiles, dim3(nt), dim3(sky::kPrimThreads), 0, s, flags, n, tsum);
  }
  hipLaunchKernelGGL(sky::k_scan_sums, dim3(1), dim3(sky::kPrimThreads), 0, s, tsum, nt, count);
  if (nt) {
    hipLaunchKernelGGL(sky::k_sel_scatter, dim3(nt), dim3(sky::kPrimThreads), 0, s, flags, n, tsum,
                       out);
  }
  return hipGetLastError();
}

// Stable LSD radix sort of n (key, value) pairs by key bits [0, end_bit):
// the result lands in kout / vout; kalt / valt (n entries) are the other
// buffers of the ping-pong, and kin / vin are left unchanged.
template <typename K>
hipError_t prim_sort_pairs(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, K* kalt,
                           uint32_t* valt, uint32_t n, int end_bit, void* tmp, hipStream_t s) {
  const int passes = (end_bit + sky::kRadixBits - 1) / sky::kRadixBits;
  const uint32_t nt = sky::prim_tiles(n);
  if (n == 0) return hipSuccess;
  uint32_t* hist = static_cast<uint32_t*>(tmp);
  const uint32_t nh = (uint32_t)sky::kRadixBins * nt;
  void* scan_tmp = hist + nh;
  // the last pass writes kout: with an odd count the first does too
  const K* ksrc = kin;
  const uint32_t* vsrc = vin;
  for (int p = 0; p < (passes ? passes : 1); ++p) {
    const bool to_out = ((passes ? passes : 1) - p) % 2 == 1;
    K* kdst = to_out ? kout : kalt;
    uint32_t* vdst = to_out ? vout : valt;
    const int shift = p * sky::kRadixBits;
    hipLaunchKernelGGL(sky::k_rs_hist<K>, dim3(nt), dim3(sky::kPrimThreads), 0, s, ksrc, n,
                       passes ? shift : 0, nt, hist);
    hipError_t e = prim_scan(hist, hist, nh, nullptr, scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sky::k_rs_scatter<K>, dim3(nt), dim3(sky::kPrimThreads), 0, s, ksrc, vsrc, n,
                       passes ? shift : 0, nt, hist, kdst, vdst);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipSuccess;
}

}  // namespace
}  // namespace skyi
