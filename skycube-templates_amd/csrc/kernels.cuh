// kernels.cuh -- HIP kernels of the skycube engine (gfx950, wave64).
//
// Data layout in HBM: points are row-major FP32, D floats per row.  Working
// copies are gathered into score order (skyline stages) or ascending point id
// order (membership stage), projected onto the D columns of the subspace.
//
// Dominance is decided from two D-bit pair masks of a candidate q against a
// point p:  lt = {i : q_i < p_i},  eq = {i : q_i == p_i}  (IEEE ordered
// compares, exactly what DominateLeft_array evaluates,
// src/utilities/dt_avx2.h:550-640).  q strictly dominates p iff lt = all;
// q dominates p in subspace s iff s is a subset of lt|eq and s meets lt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sky {

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr uint32_t kPruners = 1024;  // lowest-score points used as a prefilter
constexpr int kXCap = 2048;       // tie-touched subspaces resolved per round

struct DimList {
  int32_t idx[16];
};

// Order-preserving uint32 key of an FP32 value; -0 and +0 share a key.
__device__ __forceinline__ uint32_t order_key(float f) {
  if (f == 0.0f) f = 0.0f;
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Manhattan norm (ascending-column FP32 sum, the score MommyQ and SkyAlign
// sort by: src/hybrid/mommyq.cpp:25-33) of each candidate over the subspace
// columns, as a sort key.  FP addition is monotone, so a dominator never has
// a larger key.  err bit 0: a candidate id is out of range; bit 1: NaN.
__global__ void k_score(const float* __restrict__ rows, uint32_t n_rows, uint32_t d,
                        const uint32_t* __restrict__ cand, uint32_t c, DimList dims,
                        uint32_t ds, uint32_t* __restrict__ key, uint32_t* __restrict__ pid,
                        uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  const uint32_t r = cand ? cand[i] : i;
  if (r >= n_rows) {
    atomicOr(err, 1u);
    key[i] = 0xFFFFFFFFu;
    pid[i] = 0;
    return;
  }
  const float* row = rows + (size_t)r * d;
  float sum = 0.0f;
  bool nan = false;
  for (uint32_t k = 0; k < ds; ++k) {
    const float v = row[dims.idx[k]];
    nan |= __builtin_isnan(v);
    sum += v;
  }
  if (nan) atomicOr(err, 2u);
  key[i] = order_key(sum);
  pid[i] = r;
}

// out[i] = the subspace columns of row ids[i].
template <int D>
__global__ void k_gather(const float* __restrict__ rows, uint32_t d,
                         const uint32_t* __restrict__ ids, uint32_t c, DimList dims,
                         float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c) return;
  const float* row = rows + (size_t)ids[i] * d;
#pragma unroll
  for (int k = 0; k < D; ++k) out[(size_t)i * D + k] = row[dims.idx[k]];
}

// Gather of score-ordered rows, keys and ids at the positions idx[j].
template <int D>
__global__ void k_gather_idx(const float* __restrict__ srows, const uint32_t* __restrict__ skeys,
                             const uint32_t* __restrict__ spid, const uint32_t* __restrict__ idx,
                             uint32_t m, float* __restrict__ orows, uint32_t* __restrict__ okeys,
                             uint32_t* __restrict__ opid) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t i = idx[j];
#pragma unroll
  for (int k = 0; k < D; ++k) orows[(size_t)j * D + k] = srows[(size_t)i * D + k];
  okeys[j] = skeys[i];
  opid[j] = spid[i];
}

// Prefilter: alive[i] = 0 iff one of the npr lowest-score points strictly
// dominates point i (score order; the role of PQFilter's pruners,
// src/utilities/pq_filter.h:418-505, and of the minimax prune,
// src/utilities/kernels.cu:10-40).  Pruners sit in LDS, read as broadcasts.
template <int D>
__global__ __launch_bounds__(kBlock) void k_prune(const float* __restrict__ srows,
                                                  const uint32_t* __restrict__ skeys, uint32_t c,
                                                  uint32_t npr, uint8_t* __restrict__ alive) {
  __shared__ float tile[kBlock * D];
  __shared__ uint32_t tkey[kBlock];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < c;
  float p[D];
  uint32_t kp = 0;
  if (valid) {
#pragma unroll
    for (int k = 0; k < D; ++k) p[k] = srows[(size_t)i * D + k];
    kp = skeys[i];
  }
  bool dead = false;
  for (uint32_t t0 = 0; t0 < npr; t0 += kBlock) {
    const uint32_t cnt = min((uint32_t)kBlock, npr - t0);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < cnt * D; e += kBlock) tile[e] = srows[(size_t)t0 * D + e];
    if (threadIdx.x < cnt) tkey[threadIdx.x] = skeys[t0 + threadIdx.x];
    __syncthreads();
    if (valid && !dead) {
      for (uint32_t t = 0; t < cnt; ++t) {
        if (tkey[t] > kp) break;
        if (t0 + t == i) continue;
        bool all = true;
#pragma unroll
        for (int k = 0; k < D; ++k) all &= tile[t * D + k] < p[k];
        if (all) {
          dead = true;
          break;
        }
      }
    }
  }
  if (valid) alive[i] = dead ? 0 : 1;
}

// Survivors against each other, in score order: status[i] = 2 if strictly
// dominated, else 1 if dominated (only when need_dom), else 0.  Point i is
// tested against the survivors with key <= its own; q tiles sit in LDS and
// are read as broadcasts by the 256 points of the workgroup, which stops as
// soon as all its points are settled.  One SkyAlign pass of
// total_pruning_moved_data_templated_pitch_two_level_full_warps
// (src/utilities/kernels.cuh:36-411) over the whole candidate set.
template <int D>
__global__ __launch_bounds__(kBlock) void k_self(const float* __restrict__ rows,
                                                 const uint32_t* __restrict__ keys, uint32_t m,
                                                 int need_dom, uint8_t* __restrict__ status) {
  __shared__ float tile[kBlock * D];
  __shared__ uint32_t tkey[kBlock];
  __shared__ int s_active;
  constexpr uint32_t full = (D == 32) ? 0xFFFFFFFFu : ((1u << D) - 1u);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < m;
  float p[D];
  uint32_t kp = 0;
  if (valid) {
#pragma unroll
    for (int k = 0; k < D; ++k) p[k] = rows[(size_t)i * D + k];
    kp = keys[i];
  }
  const uint32_t last = min(m - 1u, blockIdx.x * kBlock + kBlock - 1u);
  const uint32_t kmax = keys[last];
  uint8_t st = 0;
  bool done = !valid;
  for (uint32_t t0 = 0; t0 < m; t0 += kBlock) {
    if (keys[t0] > kmax) break;  // uniform: keys ascend
    const uint32_t cnt = min((uint32_t)kBlock, m - t0);
    __syncthreads();
    if (threadIdx.x == 0) s_active = 0;
    for (uint32_t e = threadIdx.x; e < cnt * D; e += kBlock) tile[e] = rows[(size_t)t0 * D + e];
    if (threadIdx.x < cnt) tkey[threadIdx.x] = keys[t0 + threadIdx.x];
    __syncthreads();
    if (!done) {
      for (uint32_t t = 0; t < cnt; ++t) {
        if (tkey[t] > kp) {
          done = true;
          break;
        }
        if (t0 + t == i) continue;
        uint32_t lt = 0, eq = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const float q = tile[t * D + k];
          lt |= (uint32_t)(q < p[k]) << k;
          eq |= (uint32_t)(q == p[k]) << k;
        }
        if (lt == full) {
          st = 2;
          done = true;
          break;
        }
        if (need_dom && lt != 0 && (lt | eq) == full) st = 1;
      }
    }
    if (!done) s_active = 1;
    __syncthreads();
    if (!s_active) break;
  }
  if (valid) status[i] = st;
}

// flag[pid[j]] = 1 for every survivor j that nothing strictly dominates.
__global__ void k_mark_ext(const uint8_t* __restrict__ status, const uint32_t* __restrict__ pid,
                           uint32_t m, uint8_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m && status[j] != 2) flag[pid[j]] = 1;
}

// ---------------------------------------------------------------------------
// Lattice helpers.  A bitmap over the 2^D subspaces is W = max(1, 2^D/32)
// words; subspace s is bit s&31 of word s>>5.

__device__ __forceinline__ uint32_t low_closure(uint32_t x, int D) {
  // propagate "s dominated => every subset of s dominated" over the 5
  // columns that index bits inside a word
  if (D > 0) x |= (x & 0xAAAAAAAAu) >> 1;
  if (D > 1) x |= (x & 0xCCCCCCCCu) >> 2;
  if (D > 2) x |= (x & 0xF0F0F0F0u) >> 4;
  if (D > 3) x |= (x & 0xFF00FF00u) >> 8;
  if (D > 4) x |= (x & 0xFFFF0000u) >> 16;
  return x;
}

// Bits of word k whose subspace is in scope: s != 0, s < 2^D, popcount(s) <=
// max_d.
__device__ __forceinline__ uint32_t scope_mask(uint32_t k, int D, int max_d) {
  const int rem = max_d - __popc(k);
  uint32_t m = 0;
  if (rem >= 0) {
#pragma unroll
    for (int b = 0; b < 32; ++b) m |= (uint32_t)(__popc(b) <= rem) << b;
  }
  if (D < 5) m &= (1u << (1 << D)) - 1u;
  if (k == 0) m &= ~1u;
  return m;
}

// Per-point cuboid membership, one workgroup per extended-skyline point p
// (the contract of produce_bitmap_iterative_multi_partition_templated_async,
// src/utilities/kernels.cuh:415-908, without its approximations).
//
// Pass 1 over every extended-skyline q builds three lattice bitmaps in LDS:
//   A0: lt|eq of pairs without ties      -> closure D0, exact part
//   A1: lt|eq of every dominating pair   -> closure D_over >= D
//   TE: eq of pairs with ties            -> closure T (tie-touched subspaces)
// For s outside T, D_over[s] is exact: a pair marking s has s not inside its
// eq.  Inside T, s in D0 is dominated and s outside D_over is not; the rest,
// X = T & D_over & ~D0 (usually empty), is decided exactly by a second pass
// over the tie pairs: s is dominated iff some q has s inside lt|eq and s
// meeting lt.  Output: bit s = 1 iff p is not in sky(s); out-of-scope bits 1.
template <int D>
__global__ __launch_bounds__(kBlock) void k_member(const float* __restrict__ erows, uint32_t m,
                                                   int max_d, uint32_t* __restrict__ dom) {
  constexpr int W = D >= 5 ? (1 << (D - 5)) : 1;
  __shared__ uint32_t A0[W];
  __shared__ uint32_t A1[W];
  __shared__ uint32_t TE[W];
  __shared__ uint16_t xl[kXCap];
  __shared__ int s_any;
  __shared__ int s_cnt;
  const uint32_t a = blockIdx.x;
  const int tid = threadIdx.x;
  float p[D];
#pragma unroll
  for (int k = 0; k < D; ++k) p[k] = erows[(size_t)a * D + k];
  for (int k = tid; k < W; k += kBlock) {
    A0[k] = 0;
    A1[k] = 0;
    TE[k] = 0;
  }
  if (tid == 0) s_any = 0;
  __syncthreads();

  for (uint32_t b = tid; b < m; b += kBlock) {
    uint32_t lt = 0, eq = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const float q = erows[(size_t)b * D + k];
      lt |= (uint32_t)(q < p[k]) << k;
      eq |= (uint32_t)(q == p[k]) << k;
    }
    if (lt) {
      const uint32_t le = lt | eq;
      const uint32_t bit = 1u << (le & 31);
      if (!eq) {
        if (!(A0[le >> 5] & bit)) atomicOr(&A0[le >> 5], bit);
      } else {
        if (!(A1[le >> 5] & bit)) atomicOr(&A1[le >> 5], bit);
        const uint32_t ebit = 1u << (eq & 31);
        if (!(TE[eq >> 5] & ebit)) atomicOr(&TE[eq >> 5], ebit);
      }
    }
  }
  __syncthreads();

  for (int k = tid; k < W; k += kBlock) {
    const uint32_t a0 = A0[k];
    A0[k] = low_closure(a0, D);
    A1[k] = low_closure(A1[k] | a0, D);
    TE[k] = low_closure(TE[k], D);
  }
#pragma unroll 1
  for (int j = 5; j < D; ++j) {
    __syncthreads();
    const int sh = j - 5;
    for (int t = tid; t < W / 2; t += kBlock) {
      const int lo = ((t >> sh) << (sh + 1)) | (t & ((1 << sh) - 1));
      const int hi = lo | (1 << sh);
      A0[lo] |= A0[hi];
      A1[lo] |= A1[hi];
      TE[lo] |= TE[hi];
    }
  }
  __syncthreads();

  // X = T & D_over & ~D0 within scope; drop X from D_over (re-added below
  // where exact), keep X in TE.
  for (int k = tid; k < W; k += kBlock) {
    const uint32_t x = TE[k] & A1[k] & ~A0[k] & scope_mask(k, D, max_d);
    TE[k] = x;
    A1[k] &= ~x;
    if (x) s_any = 1;
  }
  __syncthreads();

  if (s_any) {  // uniform
    for (int k = tid; k < W; k += kBlock) A0[k] = 0;  // exact bits of X
#pragma unroll 1
    while (true) {
      __syncthreads();
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      for (int k = tid; k < W; k += kBlock) {
        uint32_t x = TE[k];
        while (x) {
          const int bpos = __ffs(x) - 1;
          const int slot = atomicAdd(&s_cnt, 1);
          if (slot >= kXCap) break;
          xl[slot] = (uint16_t)((k << 5) | bpos);
          x &= x - 1u;
        }
        TE[k] = x;
      }
      __syncthreads();
      const int cnt = min(s_cnt, kXCap);
      if (cnt == 0) break;  // uniform
      for (uint32_t b = tid; b < m; b += kBlock) {
        uint32_t lt = 0, eq = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const float q = erows[(size_t)b * D + k];
          lt |= (uint32_t)(q < p[k]) << k;
          eq |= (uint32_t)(q == p[k]) << k;
        }
        if (lt == 0 || eq == 0) continue;
        const uint32_t le = lt | eq;
        for (int e = 0; e < cnt; ++e) {
          const uint32_t s = xl[e];
          if (!(s & ~le) && (s & lt)) atomicOr(&A0[s >> 5], 1u << (s & 31));
        }
      }
    }
    __syncthreads();
  }

  const int any = s_any;
  for (int k = tid; k < W; k += kBlock) {
    uint32_t w = A1[k];
    if (any) w |= A0[k];
    w |= ~scope_mask(k, D, max_d);
    dom[(size_t)a * W + k] = w;
  }
}

}  // namespace sky
