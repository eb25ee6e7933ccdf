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

#include "common.cuh"

// Everything below has internal linkage: each translation unit of the
// library compiles the kernels it launches.
namespace sky {
namespace {

// Order-preserving uint32 key of an FP32 value; -0 and +0 share a key.
__device__ __forceinline__ uint32_t order_key(float f) {
  if (f == 0.0f) f = 0.0f;
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Sort key of a Manhattan norm.  FP addition is monotone, so a dominator
// never has a larger norm.  A row holding both +inf and -inf has a NaN norm;
// it gets the key of -inf: such a row has a -inf, so do all its dominators,
// whose norms are then -inf or NaN as well -- the order stays safe.
__device__ __forceinline__ uint32_t norm_key(float sum) {
  return sum == sum ? order_key(sum) : order_key(-__builtin_inff());
}

// Manhattan norm (ascending-column FP32 sum, the score MommyQ and SkyAlign
// sort by: src/hybrid/mommyq.cpp:25-33) of each candidate over the subspace
// columns, as a sort key (norm_key).  err bit 0: a candidate id is out of
// range; bit 1: NaN input.
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
  key[i] = norm_key(sum);
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

// Survivors against each other by a plain scan (the batched path, whose
// instances are small), one wave64 per point p.  The lanes walk the
// candidates q with key <= key(p) (score order) 64 at a time; q can
// (weakly) dominate p only if mask(q) is a subset of mask(p) (median masks),
// so only those are kept -- compacted with a ballot into a per-wave LDS
// queue -- and get the exact FP32 test once 64 have gathered.  The wave
// stops at the first strict dominator.  One SkyAlign pass
// (total_pruning_moved_data_templated_pitch_two_level_full_warps,
// src/utilities/kernels.cuh:36-411) over the whole candidate set, with its
// partition containment test (other subset of mine) as the filter.
// Point i over the candidates lo .. hi-1 (score order, i inside): 2 if one
// strictly dominates p, else 1 if one dominates p (only when need_dom),
// else 0.  q is the wave's queue (192 entries of LDS).
template <int D>
__device__ __forceinline__ uint8_t self_scan(const float* __restrict__ rows,
                                             const uint32_t* __restrict__ keys,
                                             const uint32_t* __restrict__ masks, uint32_t lo0,
                                             uint32_t hi0, uint32_t i, int need_dom, uint32_t* q) {
  const int lane = threadIdx.x & 63;
  float p[D];
#pragma unroll
  for (int k = 0; k < D; ++k) p[k] = rows[(size_t)i * D + k];
  const uint32_t kp = keys[i];
  const uint32_t mp = masks[i];
  // ub = first position after i with key > kp (keys ascend)
  uint32_t lo = i + 1, hi = hi0;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] <= kp) lo = mid + 1;
    else hi = mid;
  }
  const uint32_t ub = lo;
  int qn = 0;
  uint8_t st = 0;
  for (uint32_t b0 = lo0; b0 < ub; b0 += 128) {
    const uint32_t t0 = b0 + lane, t1 = t0 + 64;
    const uint32_t m0 = t0 < ub ? masks[t0] : 0xFFFFFFFFu;
    const uint32_t m1 = t1 < ub ? masks[t1] : 0xFFFFFFFFu;
    const bool c0 = t0 < ub && t0 != i && (m0 & ~mp) == 0;
    const bool c1 = t1 < ub && t1 != i && (m1 & ~mp) == 0;
    const unsigned long long b0l = __ballot(c0);
    const unsigned long long b1l = __ballot(c1);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (c0) q[qn + __popcll(b0l & below)] = t0;
    qn += __popcll(b0l);
    if (c1) q[qn + __popcll(b1l & below)] = t1;
    qn += __popcll(b1l);
    __builtin_amdgcn_wave_barrier();
    const bool last = b0 + 128 >= ub;
    if (qn >= 64 || (last && qn > 0)) {
      bool strict = false, dom = false;
      for (int e = lane; e < qn; e += 64) {
        const float* qr = rows + (size_t)q[e] * D;
        // flags, not masks: the compares' lane masks combine as scalars
        bool all_lt = true, any_lt = false, all_le = true;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const float v = qr[k];
          all_lt &= v < p[k];
          any_lt |= v < p[k];
          all_le &= v <= p[k];
        }
        strict |= all_lt;
        dom |= any_lt && all_le;
      }
      qn = 0;
      if (__ballot(strict)) {
        st = 2;
        break;
      }
      if (need_dom && __ballot(dom)) st = 1;
    }
  }
  return st;
}

// Columns k where code_q(k) <= code_p(k), for the bit-plane codes of k_codes
// (a lexicographic compare of the planes, 16 columns at a time): the upper
// halves compare planes 3:2, the lower halves planes 1:0, and the upper
// result decides unless the upper planes are equal.  Written as selects of
// three inputs each, which gfx950 issues as one v_bitop3 apiece (7 VALU
// operations in all); bits above D come out set.
__device__ __forceinline__ uint32_t planes_le(uint2 q, uint2 p) {
  const uint32_t x = q.x ^ p.x;                           // planes 3 | 1 differ
  const uint32_t le2 = (x & p.x) | (~x & (~q.y | p.y));   // planes 3:2 | 1:0: q <= p
  const uint32_t ne = x | (q.y ^ p.y);                    // planes 3:2 | 1:0 differ
  // upper half: planes 3:2 where they differ, else planes 1:0 (moved up)
  return ((ne & le2) | (~ne & (le2 << 16))) >> 16;
}

// ---------------------------------------------------------------------------
// Survivor pass over a mask index.  A weak dominator q of p has q_k <= p_k in
// every column, so its median mask (columns above the median pivot) is a
// subset of p's and every 4-bit code is <= p's.  The survivors are bucketed
// by median mask (stable, so score order holds inside a bucket); point p
// visits only the buckets of the 2^popcount(mask p) subsets of its mask, and
// in each only the prefix with key <= key(p).

// out[j] = j.
__global__ void k_iota(uint32_t* __restrict__ out, uint32_t m) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = j;
}

// boff[s] = first position of the mask-sorted order whose mask is >= s, for
// s in 0 .. nb (so bucket s is boff[s] .. boff[s+1]-1).
__global__ void k_bucket_bounds(const uint32_t* __restrict__ smask, uint32_t m, uint32_t nb,
                                uint32_t* __restrict__ boff) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nb) return;
  uint32_t a = 0, z = m;
  while (a < z) {
    const uint32_t mid = (a + z) >> 1;
    if (smask[mid] < s) a = mid + 1;
    else z = mid;
  }
  boff[s] = a;
}

// Keys (if any) and codes in bucket order.
__global__ void k_bucket_gather(const uint32_t* __restrict__ bidx, const uint32_t* __restrict__ keys,
                                const uint2* __restrict__ codes, uint32_t m,
                                uint32_t* __restrict__ bkey, uint2* __restrict__ bcode) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t i = bidx[j];
  if (keys) bkey[j] = keys[i];
  bcode[j] = codes[i];
}

// head[j] = 1 where a run of equal sorted masks starts.
__global__ void k_bucket_heads(const uint32_t* __restrict__ smask, uint32_t m,
                               uint8_t* __restrict__ head) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) head[j] = j == 0 || smask[j] != smask[j - 1];
}

// The nb non-empty buckets from their head positions: mask bmk[g], first
// position bstart[g], size bcnt[g].
__global__ void k_bucket_list(const uint32_t* __restrict__ smask, const uint32_t* __restrict__ heads,
                              uint32_t nb, uint32_t m, uint32_t* __restrict__ bmk,
                              uint32_t* __restrict__ bstart, uint32_t* __restrict__ bcnt) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nb) return;
  bstart[g] = heads[g];
  bcnt[g] = (g + 1 < nb ? heads[g + 1] : m) - heads[g];
  bmk[g] = smask[heads[g]];
}

// One wave64 per survivor i: status 2 if strictly dominated, else 1 if
// dominated (only when need_dom), else 0.  Per round the lanes take 64 of
// the subsets of mask(i), find their bucket prefixes (key <= key(i)) and
// scan the concatenation 64 entries at a time: code test, then the exact
// FP32 test of the few that pass.  The wave stops at a strict dominator.
// With minkey (the column minima of the candidates) and !need_dom, a point
// that holds a column minimum is answered at once: nothing is strictly
// below it there.  The SkyAlign survivor pass (src/utilities/kernels.cuh:36-411) with its
// partition containment test turned into an index: only candidates that
// pass the containment test are visited.  sLo / sEx: 64 words each of the
// wave's LDS.
template <int D>
__device__ __forceinline__ uint8_t self_idx(const float* __restrict__ rows,
                                            const uint32_t* __restrict__ keys,
                                            const uint2* __restrict__ codes,
                                            const uint32_t* __restrict__ boff,
                                            const uint32_t* __restrict__ bidx,
                                            const uint32_t* __restrict__ bkey,
                                            const uint2* __restrict__ bcode, uint32_t i,
                                            int need_dom, const uint32_t* __restrict__ minkey,
                                            uint32_t* sLo, uint32_t* sEx) {
  constexpr uint32_t FULL = (1u << D) - 1u;
  const int lane = threadIdx.x & 63;
  float p[D];
#pragma unroll
  for (int k = 0; k < D; ++k) p[k] = rows[(size_t)i * D + k];
  if (minkey && !need_dom) {  // uniform: a point holding a column minimum
    bool at_min = false;       // has no strict dominator
#pragma unroll
    for (int k = 0; k < D; ++k) at_min |= order_key(p[k]) == minkey[k];
    if (at_min) return 0;
  }
  const uint32_t kp = keys[i];
  const uint2 pc = codes[i];
  const uint32_t M = pc.x >> 16;  // plane 3: the median mask
  const uint32_t nsub = 1u << __popc(M);
  uint8_t st = 0;
  for (uint32_t s0 = 0; s0 < nsub && st != 2; s0 += 64) {
    const uint32_t t = s0 + (uint32_t)lane;
    uint32_t lo = 0, cnt = 0;
    if (t < nsub) {
      uint32_t S = 0, mm = M, tt = t;  // the bits of t deposited on those of M
      while (mm) {
        const uint32_t b = mm & (0u - mm);
        if (tt & 1u) S |= b;
        tt >>= 1;
        mm &= mm - 1u;
      }
      lo = boff[S];
      uint32_t a = lo, z = boff[S + 1];
      // a small bucket is taken whole (its candidates past the key bound
      // fail the code or the exact test); a large one is cut at the bound
      // by binary search, a chain of dependent loads
      if (z - a > 32) {
        while (a < z) {
          const uint32_t mid = (a + z) >> 1;
          if (bkey[mid] <= kp) a = mid + 1;
          else z = mid;
        }
        cnt = a - lo;
      } else {
        cnt = z - lo;
      }
    }
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(inc, o, 64);
      if (lane >= o) inc += v;
    }
    const uint32_t total = __shfl(inc, 63, 64);
    if (total == 0) continue;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    sLo[lane] = lo;
    sEx[lane] = inc - cnt;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e0 = 0; e0 < total; e0 += 64) {
      const uint32_t e = e0 + (uint32_t)lane;
      bool strict = false, dom = false;
      if (e < total) {
        int l = 0;  // the last lane whose range starts at or before e
#pragma unroll
        for (int step = 32; step; step >>= 1)
          if (sEx[l + step] <= e) l += step;
        const uint32_t j = sLo[l] + (e - sEx[l]);
        if ((planes_le(bcode[j], pc) & FULL) == FULL) {
          const uint32_t qi = bidx[j];
          if (qi != i) {
            const float* qr = rows + (size_t)qi * D;
            bool all_lt = true, any_lt = false, all_le = true;
#pragma unroll
            for (int k = 0; k < D; ++k) {
              const float v = qr[k];
              all_lt &= v < p[k];
              any_lt |= v < p[k];
              all_le &= v <= p[k];
            }
            strict = all_lt;
            dom = any_lt && all_le;
          }
        }
      }
      if (__ballot(strict)) {
        st = 2;
        break;
      }
      if (need_dom && __ballot(dom)) st = 1;
    }
  }
  return st;
}

// One survivor per workgroup of one wave (measured 16-20% faster than four
// per workgroup of four waves: a finished wave's slot is taken at once by
// the next survivor), held to 64 VGPRs (eight waves per SIMD).
template <int D>
__global__ __attribute__((amdgpu_waves_per_eu(8))) __launch_bounds__(64) void k_self_idx(
    const float* __restrict__ rows, const uint32_t* __restrict__ keys,
    const uint2* __restrict__ codes,
    const uint32_t* __restrict__ boff, const uint32_t* __restrict__ bidx,
    const uint32_t* __restrict__ bkey, const uint2* __restrict__ bcode, uint32_t m, int need_dom,
    const uint32_t* __restrict__ minkey, uint8_t* __restrict__ status) {
  __shared__ uint32_t sLo[64];
  __shared__ uint32_t sEx[64];
  const uint32_t i = blockIdx.x;
  if (i >= m) return;
  const uint8_t st = self_idx<D>(rows, keys, codes, boff, bidx, bkey, bcode, i, need_dom,
                                 minkey, sLo, sEx);
  if ((threadIdx.x & 63) == 0) status[i] = st;
}

// flag[pid[j]] = 1 for every survivor j that nothing strictly dominates.
__global__ void k_mark_ext(const uint8_t* __restrict__ status, const uint32_t* __restrict__ pid,
                           uint32_t m, uint8_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m && status[j] != 2) flag[pid[j]] = 1;
}

// flag[j] = 1 for every survivor j that nothing strictly dominates.
__global__ void k_flag_ext(const uint8_t* __restrict__ status, uint32_t m,
                           uint8_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) flag[j] = status[j] != 2;
}

// The extended skyline in score order: rows and ids at survivor positions.
template <int D>
__global__ void k_gather_ext(const float* __restrict__ s1rows, const uint32_t* __restrict__ s1pid,
                             const uint32_t* __restrict__ idx, uint32_t ne,
                             float* __restrict__ erows, uint32_t* __restrict__ epid) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ne) return;
  const uint32_t i = idx[j];
#pragma unroll
  for (int k = 0; k < D; ++k) erows[(size_t)j * D + k] = s1rows[(size_t)i * D + k];
  epid[j] = s1pid[i];
}

// rank[ids[i]] = i (ids ascending point ids of the extended skyline).
__global__ void k_scatter_rank(const uint32_t* __restrict__ ids, uint32_t ne,
                               uint32_t* __restrict__ rank) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) rank[ids[i]] = i;
}

// pos[j] = output row (ascending-id position) of score-order point j.
__global__ void k_pos(const uint32_t* __restrict__ epid, uint32_t ne,
                      const uint32_t* __restrict__ rank, uint32_t* __restrict__ pos) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < ne) pos[j] = rank[epid[j]];
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
  // bits b of a word with popcount(b) <= r, for r = 0..4 (r >= 5: all)
  const int rem = max_d - __popc(k);
  uint32_t m = rem >= 5 ? 0xFFFFFFFFu : 0x7FFFFFFFu;
  m = rem <= 3 ? 0x177F7FFFu : m;
  m = rem <= 2 ? 0x0117177Fu : m;
  m = rem <= 1 ? 0x00010117u : m;
  m = rem <= 0 ? 0x00000001u : m;
  m = rem < 0 ? 0u : m;
  if (D < 5) m &= (1u << (1 << D)) - 1u;
  if (k == 0) m &= ~1u;
  return m;
}

// Bits b of a word (the subsets b of the 5 low columns) with b inside zl,
// and with exactly one column of b outside zl (zl: 5-bit column mask).
__device__ __forceinline__ void low_subsets(uint32_t zl, uint32_t* in, uint32_t* one_out) {
  uint32_t m0 = 1u;
#pragma unroll
  for (int i = 0; i < 5; ++i)
    if ((zl >> i) & 1u) m0 |= m0 << (1u << i);
  uint32_t m1 = 0u;
#pragma unroll
  for (int i = 0; i < 5; ++i)
    if (!((zl >> i) & 1u)) m1 |= m0 << (1u << i);
  *in = m0;
  *one_out = m1;
}

// Column minima of the extended skyline, as order keys, for the boundary
// columns of the membership kernel: a block-wide minimum per column, then
// one atomicMin per block and column.  minkey must start at ~0.
template <int D>
__global__ __launch_bounds__(kBlock) void k_colmin(const float* __restrict__ erows, uint32_t m,
                                                   uint32_t* __restrict__ minkey) {
  __shared__ uint32_t red[D];
  if (threadIdx.x < D) red[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  uint32_t best[D];
#pragma unroll
  for (int k = 0; k < D; ++k) best[k] = 0xFFFFFFFFu;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < D; ++k) best[k] = min(best[k], order_key(erows[(size_t)i * D + k]));
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    uint32_t v = best[k];
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    if ((threadIdx.x & 63) == 0) atomicMin(&red[k], v);
  }
  __syncthreads();
  if (threadIdx.x < D) atomicMin(&minkey[threadIdx.x], red[threadIdx.x]);
}

// Rank of the bits of `sub` inside `zm` (a parallel bit extract): bit j of
// sub, j in zm, goes to bit popcount(zm below j).
__device__ __forceinline__ uint32_t zcompress(uint32_t sub, uint32_t zm) {
  uint32_t idx = 0;
  for (uint32_t t = sub; t; t &= t - 1u) {
    const int bit = __ffs(t) - 1;
    idx |= 1u << __popc(zm & ((1u << bit) - 1u));
  }
  return idx;
}

// Sets bit `s` of an LDS bitmap unless it is already set.
__device__ __forceinline__ void lds_mark(uint32_t* w, uint32_t s) {
  const uint32_t bit = 1u << (s & 31);
  if (!(w[s >> 5] & bit)) atomicOr(&w[s >> 5], bit);
}

// Boundary columns tracked per point (a 2^ZB-word table per point).
template <int D>
constexpr int zbits() {
  return D < 8 ? D : (D >= 15 ? 10 : 8);
}
// At d = 16 the single-instance kernel runs with a table of 2^8 words (and
// 64-entry queues), so that three workgroups fit a CU's LDS; the few points
// with more minimum columns are deferred to a second launch with the full
// table.  (At d = 15 registers, not LDS, set the occupancy.)
template <int D>
constexpr int narrow_zbits() {
  return D == 16 ? 8 : zbits<D>();
}

// Points per workgroup of k_member: as many as keep its LDS near 64 KB,
// and no more than 4 from d = 11 on (measured: the points of a workgroup
// share its candidate groups; more points make more groups live).  The
// single-instance launch at d = 16 with the tie lattices takes one point
// per workgroup instead (launch_member).
template <int D>
constexpr int member_points() {
  return D >= 15 ? 2 : (D >= 11 ? 4 : 8);
}

// The other tile width the batch autotuner may pick: half as many points.
template <int D>
constexpr int member_points_alt() {
  return member_points<D>() > 1 ? member_points<D>() / 2 : 1;
}

// Threads per workgroup of k_member: the LDS of one workgroup (two per CU
// at d >= 13) is shared by eight waves there, four elsewhere.
template <int D>
constexpr int member_threads() {
  return D >= 13 ? 512 : 256;
}

// Quantile pivots of each column, as order keys: kPivotSplit blocks per
// column; a strided sample of up to 2048 values is ranked by counting and
// the values at ranks b*S/16 (b = 1..15) become the pivots.  Any pivots give correct
// codes; good ones make the codes selective.
constexpr int kPivotSplit = 8;  // blocks per column of k_pivots

//
// If dups is given it gains the number of sample values equal to an
// earlier one of their column and above the sample's minimum (how tied the
// columns are away from their minima, where ties are boundary columns).
template <int D>
__global__ __launch_bounds__(kBlock) void k_pivots(const float* __restrict__ erows, uint32_t m,
                                                   uint32_t* __restrict__ piv,
                                                   uint32_t* __restrict__ dups) {
  constexpr int S = 2048;
  __shared__ uint32_t val[S];
  const int col = blockIdx.x / kPivotSplit;
  const int part = blockIdx.x % kPivotSplit;
  const uint32_t cnt = min((uint32_t)S, m);
  for (uint32_t i = threadIdx.x; i < cnt; i += kBlock)
    val[i] = order_key(erows[(size_t)((uint64_t)i * m / cnt) * D + col]);
  __syncthreads();
  // each sample value's rank (ties broken by position, so ranks are
  // distinct); the values of ranks (b+1)*cnt/16 are the pivots
  uint32_t nd = 0;
  for (uint32_t i = part * kBlock + threadIdx.x; i < cnt; i += kBlock * kPivotSplit) {
    const uint32_t v = val[i];
    uint32_t r = 0, e = 0;
    for (uint32_t j = 0; j < cnt; ++j) {
      r += (val[j] < v) || (val[j] == v && j < i);
      e += val[j] == v && j < i;
    }
    nd += e != 0 && r != e;  // r - e values are below v
#pragma unroll 1
    for (int b = 0; b < kCodePivots; ++b)
      if ((uint32_t)((b + 1) * cnt / (kCodePivots + 1)) == r) piv[col * kCodePivots + b] = v;
  }
  if (dups) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nd += __shfl_xor(nd, o, 64);
    if ((threadIdx.x & 63) == 0 && nd) atomicAdd(dups, nd);
  }
}

// 4-bit quantile code of every column (the number of pivots below the
// value, so v <= w implies code(v) <= code(w)), stored as bit planes: bit k
// of plane b is bit b of column k's code.  x = plane3 << 16 | plane1,
// y = plane2 << 16 | plane0.  masks[i] = plane 3: the column is above the
// median pivot.
template <int D>
__global__ void k_codes(const float* __restrict__ erows, uint32_t m,
                        const uint32_t* __restrict__ piv, uint2* __restrict__ codes,
                        uint32_t* __restrict__ masks) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t pl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const uint32_t key = order_key(erows[(size_t)i * D + k]);
    uint32_t c = 0;
#pragma unroll
    for (int b = 0; b < kCodePivots; ++b) c += piv[k * kCodePivots + b] < key;
#pragma unroll
    for (int b = 0; b < 4; ++b) pl[b] |= ((c >> b) & 1u) << k;
  }
  codes[i] = make_uint2((pl[3] << 16) | pl[1], (pl[2] << 16) | pl[0]);
  masks[i] = pl[3];
}

// Marks one (q, p) pair with masks lt, eq into the structures of point pp.
// Without the tie lattices (TIES false), a pair with a residual tie only
// flags its point (tie[pp]), which is then computed again with them.
template <bool TIES, int W, int NU, int WT>
__device__ __forceinline__ void mark_pair(uint32_t (*A0)[W], uint32_t (*A1)[WT],
                                          uint32_t (*TE)[WT], uint32_t (*U)[NU], int* tie, int pp,
                                          uint32_t lt, uint32_t eq, uint32_t z) {
  if (!lt) return;
  const uint32_t le = lt | eq;
  const uint32_t er = eq & ~z;
  if (!er) {
    lds_mark(A0[pp], le);
  } else if constexpr (TIES) {
    lds_mark(A1[pp], le);
    lds_mark(TE[pp], er);
  } else {
    tie[pp] = 1;
    return;
  }
  const uint32_t ez = eq & z;
  if (ez) {
    const uint32_t idx = zcompress(ez, z);
    if ((U[pp][idx] & lt) != lt) atomicOr(&U[pp][idx], lt);
  }
}

template <int D>
__device__ __forceinline__ void pair_masks(const float* __restrict__ q, const float* p,
                                           uint32_t& lt, uint32_t& eq) {
  float qv[D];
#pragma unroll
  for (int k = 0; k < D; ++k) qv[k] = q[k];
  // shifted in from the top column down: no per-column bit constants,
  // which the compiler would otherwise keep in registers
  lt = 0;
  eq = 0;
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    lt = (lt << 1) | (uint32_t)(qv[k] < p[k]);
    eq = (eq << 1) | (uint32_t)(qv[k] == p[k]);
  }
}

// How k_member visits the candidates after the cover set (order):
//   kOrderScan    all of them, in score order
//   kOrderBucket  through median-mask buckets: bucket g holds mask bmk[g] at
//                 positions bstart[g] .. bstart[g] + bcnt[g] - 1 of the
//                 bucket order (codes bcode, candidate ids bidx)
//   kOrderGroup   in groups of 64 consecutive candidates of the code order
//                 (sorted by the bit-plane code key): gcode / gidx are the
//                 codes and candidate ids in that order, gsum[g] the
//                 per-column minimum code of group g as bit planes
//   kOrderSub     the same groups, and ssum[8g + s] the minimum code of
//                 subgroup s of group g (members 8s .. 8s+7): live groups
//                 are narrowed to their live subgroups before any member
//                 is tested
constexpr int kOrderScan = 0, kOrderBucket = 1, kOrderGroup = 2, kOrderSub = 3;
// Default order of the single-instance path (measured: the subgroup steps
// pay off up to d = 14 and at d = 16, where the launch keeps four
// workgroups per CU (one point each with the tie lattices, two without);
// at d = 15 the two orders are level).
template <int D>
constexpr int single_order() {
  return D == 15 ? kOrderGroup : kOrderSub;
}
struct CandIndex {
  uint32_t head = 0;  // cover candidates tested pair by pair (0: an eighth of the cover)
  const uint32_t* bmk = nullptr;
  const uint32_t* bstart = nullptr;
  const uint32_t* bcnt = nullptr;
  uint32_t nbk = 0;
  const uint32_t* bidx = nullptr;
  const uint2* bcode = nullptr;
  const uint2* gsum = nullptr;
  const uint2* ssum = nullptr;  // subgroup g * 8 + s: members 8s .. 8s+7 of group g
  const uint2* gcode = nullptr;
  const uint32_t* gidx = nullptr;
  // the points: pts[0 .. npts-1] if set (else positions 0 .. m-1, through
  // gidx in the group order)
  const uint32_t* pts = nullptr;
  uint32_t npts = 0;
  // if set, a point with more minimum columns than the kernel's boundary
  // table holds is not computed but appended to defer[1 ..] (count defer[0])
  // (with the tie lattices left out, so is a point that meets a residual tie)
  uint32_t* defer = nullptr;
  // if set, such a point is skipped without being appended: k_defer_list
  // listed it before the launch
  bool defer_listed = false;
  // if set, rowcnt[row] gains the number of the row's words that are not
  // all ones (the HashCube export's entries of that point)
  uint32_t* rowcnt = nullptr;
};

// Down-closure of the P bitmaps A (W words each) in place; caller syncs
// before.  Ends with a barrier.
template <int D, int P, int W, int NT>
__device__ __forceinline__ void close_down(uint32_t (*A)[W]) {
  for (int k = threadIdx.x; k < W; k += NT) {
#pragma unroll
    for (int pp = 0; pp < P; ++pp) A[pp][k] = low_closure(A[pp][k], D);
  }
#pragma unroll 1
  for (int j = 5; j < D; ++j) {
    __syncthreads();
    const int sh = j - 5;
    for (int t = threadIdx.x; t < W / 2; t += NT) {
      const int lo = ((t >> sh) << (sh + 1)) | (t & ((1 << sh) - 1));
#pragma unroll
      for (int pp = 0; pp < P; ++pp) A[pp][lo] |= A[pp][lo | (1 << sh)];
    }
  }
  __syncthreads();
}

// Per-point cuboid membership, P extended-skyline points per workgroup (the
// contract of produce_bitmap_iterative_multi_partition_templated_async,
// src/utilities/kernels.cuh:415-908, without its approximations).
//
// Points and candidates are the extended skyline in score order; point a's
// bitmap goes to row pos[a] of dom (ascending point id).
//
// For a point p and a candidate q: lt = {i: q_i < p_i}, eq = {i: q_i ==
// p_i}, L = lt|eq.  q dominates p in s iff s is inside L and s meets lt.
// Boundary columns Z = {j: p_j is the column minimum} (at most zbits<D>())
// admit no q_j < p_j, so for j in Z, j in L iff j in eq.  Write s = sZ + sR
// (sR outside Z) and eR = eq outside Z (the residual ties):
//   sR empty          -> never dominated (no strict column inside s)
//   sR = {j}          -> dominated iff some q has q_j < p_j and ties p on
//                        all of sZ: bit j of the superset-OR transform U over
//                        Z's lattice, U[zc(eq & Z)] |= lt  (exact, ties or not)
//   |sR| >= 2         -> A0 = L of pairs with eR empty (they dominate every s
//                        inside L with sR non-empty); A1 = L of every pair;
//                        TE = eR of the others.  After closures, s outside
//                        T (TE closed down on R, up on Z) follows A1 exactly,
//                        s in T and in A0 is dominated, s outside A1 is not;
//                        the rest, X, is decided by a second pass over the
//                        pairs with residual ties (rare).
//
// Pruning: the `cover` lowest-score candidates are marked first and A0 is
// closed.  Any set bit of A0 is then inside the L of a marked pair with no
// residual tie, which dominates p in every subspace (sR non-empty) that a
// pair with a smaller L could, and carries its U contribution too.  So a
// later q whose code bound on L (columns with code_q <= code_p, a superset
// of L, from the 4-bit bit-plane codes) is set in A0 adds nothing and is
// skipped.  Marks made later only add bits of the true closure, so reading
// A0 while other waves mark it can only skip less.  Survivors of the code
// test go to a per-wave LDS queue and get the exact FP32 test together.
// Output: bit s = 1 iff p is not in sky(s); out-of-scope bits are 1.
//
// member_body processes the points base .. base+P-1 of one instance with NT
// threads; the single-instance kernel and the batched persistent kernel both
// call it.  order (uniform) and ix say how the candidates after the cover
// set are visited (CandIndex).  Candidate ids are queued as id << 3 |
// point, so m < 2^29.
//
// With TIES false the workgroup holds A0 and U only (a third of the LDS at
// d = 16, so more workgroups share a CU): a point whose exact tests meet a
// residual tie (a pair with eR non-empty) is not stored but appended to
// ix.defer, for a launch with TIES true.  Without such pairs A1 = A0 and TE
// is empty, so the result is A0's.
template <int D, int P, int NT, int ZB = zbits<D>(), bool TIES = true>
__device__ __forceinline__ void member_body(const float* __restrict__ erows,
                                            const uint2* __restrict__ codes,
                                            const uint32_t* __restrict__ pos, uint32_t m,
                                            int max_d, const uint32_t* __restrict__ minkey,
                                            uint32_t* __restrict__ dom, uint32_t base,
                                            uint32_t cover, int order, const CandIndex& ix) {
  constexpr int W = D >= 5 ? (1 << (D >= 5 ? D - 5 : 0)) : 1;
  constexpr int NU = 1 << ZB;
  constexpr int QW = D == 16 ? 64 : 128;  // queue entries per wave
  constexpr int QCAP = (NT / 64) * QW;
  constexpr uint32_t FULL = (1u << D) - 1u;
  constexpr int WT = TIES ? W : 1;  // the tie lattices' words
  __shared__ uint32_t A0[P][W];
  __shared__ uint32_t A1[P][WT];
  __shared__ uint32_t TE[P][WT];
  __shared__ uint32_t U[P][NU];
  __shared__ int s_tie[P];  // a residual tie was met (TIES false)
  __shared__ float prow[P][D];
  __shared__ uint32_t queue[QCAP];
  // per-wave lists: the code-order groups' two lists (kList1 + kList2
  // entries), or the bucket order's three arrays of 64
  constexpr int kList1 = 72, kList2 = 64 + 64 * P;
  constexpr int kWList = kList1 + kList2 > 192 ? kList1 + kList2 : 192;
  __shared__ uint32_t wlist[NT / 64][kWList];
  __shared__ int s_any[P];
  __shared__ uint32_t s_zm[P];
  __shared__ uint32_t s_rgt[P];
  __shared__ uint32_t s_pt[P];  // the points: entry base + pp of the score order, the code order or ix.pts
  __shared__ uint2 s_pc[P];     // their codes
  __shared__ uint32_t s_ok;     // bit pp: point pp exists and is computed here
  __shared__ int s_cnt;
  uint16_t* xl = reinterpret_cast<uint16_t*>(queue);  // X list reuses the queue
  constexpr int XCAP = QCAP * 2 < kXCap ? QCAP * 2 : kXCap;
  const int tid = threadIdx.x;

  if (tid < P) {
    // with code-order groups, the workgroup's points are neighbours in the
    // code order, so that the groups they need coincide
    const uint32_t np = ix.pts ? ix.npts : m;
    const uint32_t bt = min(base + tid, np - 1u);
    const uint32_t a =
        ix.pts ? ix.pts[bt] : (order == kOrderGroup || order == kOrderSub ? ix.gidx[bt] : bt);
    bool ok = base + tid < np;
    s_pt[tid] = a;
    uint32_t z = 0, g = 0;
    int nz = 0;
    for (int k = 0; k < D; ++k) {
      const float v = erows[(size_t)a * D + k];
      prow[tid][k] = v;
      const bool at_min = order_key(v) == minkey[k];
      if (at_min && nz < ZB) {
        z |= 1u << k;
        ++nz;
      }
      g |= (uint32_t)(!at_min) << k;
    }
    if ((ix.defer || ix.defer_listed) && ok && __popc(~g & FULL) > ZB) {
      if (!ix.defer_listed) ix.defer[1 + atomicAdd(ix.defer, 1u)] = a;
      ok = false;
    }
    const uint32_t okm = (uint32_t)__ballot(ok);  // the P lanes are in wave 0
    if (tid == 0) s_ok = okm;
    s_pc[tid] = codes[a];
    s_zm[tid] = z;
    s_rgt[tid] = g;
    s_any[tid] = 0;
    s_tie[tid] = 0;
  }
  for (int k = tid; k < W; k += NT) {
#pragma unroll
    for (int pp = 0; pp < P; ++pp) {
      A0[pp][k] = 0;
      if constexpr (TIES) {
        A1[pp][k] = 0;
        TE[pp][k] = 0;
      }
    }
  }
  for (int k = tid; k < NU; k += NT) {
#pragma unroll
    for (int pp = 0; pp < P; ++pp) U[pp][k] = 0;
  }
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  if (s_ok == 0) {  // uniform: no point of this workgroup is computed here
    __syncthreads();
    return;
  }

  // Candidate passes.  Survivors of the code test (rare) are appended to
  // the wave's queue (tested on the spot when it is full), and a wave
  // drains its queue when it holds 64 or more.
  const int lane = tid & 63;
  uint32_t* wq = queue + (tid >> 6) * QW;
  uint32_t nq = 0;  // wave-uniform
  auto exact = [&](uint32_t b, int pp) {
    uint32_t lt, eq;
    pair_masks<D>(erows + (size_t)b * D, prow[pp], lt, eq);
    mark_pair<TIES, W, NU, WT>(A0, A1, TE, U, s_tie, pp, lt, eq, s_zm[pp]);
  };
  // the queue is written and read by different lanes of one wave: LDS
  // operations of a wave complete in order, and the fences keep the
  // compiler from moving queue accesses across the drain
  auto drain = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t e = lane; e < nq; e += 64) {
      const uint32_t v = wq[e];
      exact(v >> 3, (int)(v & 7u));
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    nq = 0;
  };
  // appends the candidate b of this lane for point pp if bit j of hit is set
  auto append = [&](uint32_t hit, int j, int pp, uint32_t b) {
    const uint64_t bj = __ballot((hit >> j) & 1u);
    if (!bj) return;
    if ((hit >> j) & 1u) {
      const uint32_t slot =
          nq + __builtin_amdgcn_mbcnt_hi((uint32_t)(bj >> 32),
                                         __builtin_amdgcn_mbcnt_lo((uint32_t)bj, 0u));
      if (slot < (uint32_t)QW)
        wq[slot] = (b << 3) | (uint32_t)pp;
      else
        exact(b, pp);
    }
    nq = min(nq + (uint32_t)__popcll(bj), (uint32_t)QW);
  };
  uint2 pc[P];
#pragma unroll
  for (int pp = 0; pp < P; ++pp) pc[pp] = codes[s_pt[pp]];

  // Every candidate in lo .. hi-1, in score order: each thread code-tests
  // kPer candidates against the P points per round without branches, the
  // next round's codes already in flight.
  auto scan = [&](uint32_t lo, uint32_t hi) {
    constexpr int kPer = 4;
    static_assert(kPer * P <= 32, "survivor mask is one word");  // also 4 live groups x P
    constexpr uint32_t kStep = (uint32_t)NT * kPer;
    uint2 qn[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) qn[r] = codes[min(lo + r * NT + tid, m - 1u)];
    for (uint32_t c0 = lo; c0 < hi; c0 += kStep) {
      uint2 qc[kPer];
#pragma unroll
      for (int r = 0; r < kPer; ++r) qc[r] = qn[r];
      if (c0 + kStep < hi) {  // uniform
#pragma unroll
        for (int r = 0; r < kPer; ++r) qn[r] = codes[min(c0 + kStep + r * NT + tid, m - 1u)];
      }
      uint32_t hit = 0;
#pragma unroll
      for (int r = 0; r < kPer; ++r) {
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          const uint32_t ub = planes_le(qc[r], pc[pp]) & FULL;
          const uint32_t w = A0[pp][ub >> 5];
          hit |= (~(w >> (ub & 31)) & 1u) << (r * P + pp);
        }
      }
#pragma unroll
      for (int r = 0; r < kPer; ++r)
        if (c0 + r * NT + tid >= hi) hit &= ~(((1u << P) - 1u) << (r * P));
      if (__ballot(hit != 0u) == 0ull) continue;  // uniform
#pragma unroll 1
      for (int j = 0; j < kPer * P; ++j)
        append(hit, j, j % P, c0 + (uint32_t)(j / P) * NT + tid);
      if (nq >= 64) drain();
    }
  };
  auto close_a0 = [&]() { close_down<D, P, W, NT>(A0); };

  // pass 1a: the cover.  Its head (an eighth by default) is tested pair by pair; A0 is
  // closed; the rest of the cover is code-tested against that A0 (in
  // score order, the likely dominators first); A0 is closed again.
  const uint32_t kc = min(m, cover);
  const uint32_t kc0 = min(kc, ix.head ? ix.head : max(64u, cover / 8));
  for (uint32_t e = tid; e < kc0 * P; e += NT) {
    const uint32_t b = e / P;
    const int pp = (int)(e - b * P);
    uint32_t lt, eq;
    pair_masks<D>(erows + (size_t)b * D, prow[pp], lt, eq);
    mark_pair<TIES, W, NU, WT>(A0, A1, TE, U, s_tie, pp, lt, eq, s_zm[pp]);
  }
  __syncthreads();
  close_a0();
  if (kc > kc0) {  // uniform
    scan(kc0, kc);
    drain();
    __syncthreads();
    close_a0();
  }

  // pass 1b: the candidates after the cover.  No workgroup barrier until
  // the end.
  if (order == kOrderGroup) {
    // Groups of 64 candidates of the code order.  For q in group g, every
    // code of q is >= the group's column minimum, so the columns where
    // gsum[g] <= code(p) contain the code bound of every member: if that
    // set is in A0 the whole group is skipped for p.  The waves split the
    // groups; lane l tests group g0 + l for the workgroup's points (which
    // are neighbours in the code order, so they mostly need the same
    // groups), then the wave code-tests the live groups' members, four
    // groups of 64 at a time.
    constexpr int NW = NT / 64;
    constexpr int WPP = NW;
    const int wv = tid >> 6;
    const uint32_t mine = __builtin_amdgcn_readfirstlane(s_ok);
    const uint32_t ng = (m + 63) / 64;
    for (uint32_t g0 = (uint32_t)(wv % WPP) * 64; g0 < ng; g0 += (uint32_t)WPP * 64) {
      const uint32_t g = g0 + (uint32_t)lane;
      uint32_t al = 0;
      if (g < ng) {
        const uint2 gs = ix.gsum[g];
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          if (!((mine >> pp) & 1u)) continue;  // uniform
          const uint32_t ub = planes_le(gs, pc[pp]) & FULL;
          al |= ((~(A0[pp][ub >> 5] >> (ub & 31))) & 1u) << pp;
        }
      }
      uint64_t live = __ballot(al != 0u);
      while (live) {  // uniform; up to kLive live groups per round, loads in flight together
        constexpr int kLive = 4;
        int Ls[kLive];
#pragma unroll
        for (int r = 0; r < kLive; ++r) {
          Ls[r] = live ? __ffsll((long long)live) - 1 : -1;
          live &= live - 1ull;
        }
        uint2 qc[kLive];
        uint32_t gal[kLive], jj[kLive];
#pragma unroll
        for (int r = 0; r < kLive; ++r) {
          const int L = Ls[r] < 0 ? 0 : Ls[r];
          // lane L is the same for every lane: a scalar read, so that the
          // point tests below branch on a scalar mask
          gal[r] = Ls[r] < 0 ? 0u : (uint32_t)__builtin_amdgcn_readlane((int)al, L);
          jj[r] = (g0 + (uint32_t)L) * 64u + (uint32_t)lane;
          qc[r] = ix.gcode[min(jj[r], m - 1u)];
        }
        // set bits of A0 are gathered as they are (bit r * P + pp) and the
        // tested, existing pairs without one become the hits in one step
        uint32_t seen = 0, tested = 0;
#pragma unroll
        for (int r = 0; r < kLive; ++r) {
          tested |= (jj[r] < m ? gal[r] : 0u) << (r * P);
#pragma unroll
          for (int pp = 0; pp < P; ++pp) {
            if (!((gal[r] >> pp) & 1u)) continue;  // uniform
            const uint32_t ub = planes_le(qc[r], pc[pp]) & FULL;
            const uint32_t w = A0[pp][ub >> 5];
            seen |= ((w >> (ub & 31)) & 1u) << (r * P + pp);
          }
        }
        const uint32_t hit = tested & ~seen;
        if (__ballot(hit != 0u) == 0ull) continue;  // uniform
#pragma unroll 1
        for (int r = 0; r < kLive; ++r) {
          uint32_t h = (hit >> (r * P)) & ((1u << P) - 1u);
          if (__ballot(h != 0u) == 0ull) continue;  // uniform
          const uint32_t b = h ? ix.gidx[jj[r]] : 0u;
          if (b < kc) h = 0;  // the cover candidates were marked in pass 1a
#pragma unroll 1
          for (int pp = 0; pp < P; ++pp)
            if ((mine >> pp) & 1u) append(h, pp, pp, b);
          if (nq >= 64) drain();
        }
      }
    }
  } else if (order == kOrderSub) {
    // Groups of 64 candidates of the code order, in three steps.  For q in
    // group g, every code of q is >= the group's column minimum, so the
    // columns where gsum[g] <= code(p) contain the code bound of every
    // member: if that set is in A0 the whole group is skipped for p.  The
    // same holds for each subgroup of 8 with its own minimum (ssum), which
    // is tighter.  Each step feeds a per-wave list in LDS that the next
    // step consumes with every lane busy:
    //   1. lane l tests group g0 + l for the workgroup's points (neighbours
    //      in the code order, so they mostly need the same groups); live
    //      groups go to list 1 (group << 8 | points);
    //   2. eight live groups at a time, lane l tests subgroup l & 7 of
    //      group l >> 3; live (subgroup, point) pairs go to list 2;
    //   3. 64 pairs at a time, lane l code-tests the eight members of its
    //      subgroup (64 contiguous bytes of codes) against its point.
    // The waves split the groups; a step runs when its list holds a full
    // round (or, at the end, whatever is left).
    constexpr int NW = NT / 64;
    const int wv = tid >> 6;
    const uint32_t mine = __builtin_amdgcn_readfirstlane(s_ok);  // uniform: scalar branches
    const uint32_t ng = (m + 63) / 64;
    uint32_t* l1 = wlist[wv];            // < 72 entries
    uint32_t* l2 = wlist[wv] + kList1;   // < 64 + 64 P entries: subgroup << 3 | point
    uint32_t n1 = 0, n2 = 0;             // wave-uniform
    auto sync_wave = [&]() {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    };
    uint32_t g0 = (uint32_t)wv * 64;
    while (true) {  // uniform
      const bool more = g0 < ng;
      if (more) {  // step 1
        const uint32_t g = g0 + (uint32_t)lane;
        uint32_t al = 0;
        if (g < ng) {
          const uint2 gs = ix.gsum[g];
#pragma unroll
          for (int pp = 0; pp < P; ++pp) {
            if (!((mine >> pp) & 1u)) continue;  // uniform
            const uint32_t ub = planes_le(gs, pc[pp]) & FULL;
            al |= ((~(A0[pp][ub >> 5] >> (ub & 31))) & 1u) << pp;
          }
        }
        const uint64_t bl = __ballot(al != 0u);
        if (al)
          l1[n1 + __builtin_amdgcn_mbcnt_hi((uint32_t)(bl >> 32),
                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bl, 0u))] =
              (g << 8) | al;
        n1 += (uint32_t)__popcll(bl);
        g0 += (uint32_t)NW * 64;
      }
#pragma unroll 1
      while (true) {  // uniform
        if (n2 >= 64 || (!more && n1 == 0 && n2 > 0)) {  // step 3: the last 64 pairs
          sync_wave();
          const uint32_t take = min(n2, 64u);
          const bool v = (uint32_t)lane < take;
          const uint32_t ent = v ? l2[n2 - take + (uint32_t)lane] : 0u;
          sync_wave();
          n2 -= take;
          const int pp = (int)(ent & 7u);
          const uint32_t j0 = (ent >> 3) * 8u;  // < m for every pair
          uint2 qc[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) qc[i] = ix.gcode[min(j0 + (uint32_t)i, m - 1u)];
          const uint2 pcv = s_pc[pp];
          uint32_t hit = 0;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const uint32_t ub = planes_le(qc[i], pcv) & FULL;
            hit |= ((~(A0[pp][ub >> 5] >> (ub & 31))) & 1u) << i;
          }
          if (!v) hit = 0;
          if (m - j0 < 8u) hit &= (1u << (m - j0)) - 1u;
          if (__ballot(hit != 0u) == 0ull) continue;  // uniform
#pragma unroll 1
          for (int i = 0; i < 8; ++i) {
            uint32_t h = (hit >> i) & 1u;
            if (__ballot(h != 0u) == 0ull) continue;  // uniform
            const uint32_t b = h ? ix.gidx[j0 + (uint32_t)i] : 0u;
            if (b < kc) h = 0;  // the cover candidates were marked in pass 1a
            append(h, 0, pp, b);
            if (nq >= 64) drain();
          }
          continue;
        }
        if (n1 >= 8 || (!more && n1 > 0)) {  // step 2: the last eight live groups
          sync_wave();
          const uint32_t take = min(n1, 8u);
          const uint32_t k = (uint32_t)lane >> 3;
          const bool v = k < take;
          const uint32_t ent = v ? l1[n1 - take + k] : 0u;
          sync_wave();
          n1 -= take;
          const uint32_t gal = ent & 0xFFu;
          const uint32_t sg = (ent >> 8) * 8u + (uint32_t)(lane & 7);
          const bool sv = v && sg * 8u < m;  // the subgroup has members
          const uint2 ss = ix.ssum[sv ? sg : 0u];
#pragma unroll
          for (int pp = 0; pp < P; ++pp) {
            if (!((mine >> pp) & 1u)) continue;  // uniform
            const uint32_t ub = planes_le(ss, pc[pp]) & FULL;
            const bool lv = sv && ((gal >> pp) & 1u) && !((A0[pp][ub >> 5] >> (ub & 31)) & 1u);
            const uint64_t bl = __ballot(lv);
            if (lv)
              l2[n2 + __builtin_amdgcn_mbcnt_hi((uint32_t)(bl >> 32),
                                                __builtin_amdgcn_mbcnt_lo((uint32_t)bl, 0u))] =
                  (sg << 3) | (uint32_t)pp;
            n2 += (uint32_t)__popcll(bl);
          }
          continue;
        }
        break;
      }
      if (!more) break;
    }
  } else if (order == kOrderBucket) {
    // Candidates grouped by median mask (plane 3 of the codes): bucket g
    // holds the candidates of mask bmk[g] at positions bstart[g] ..
    // bstart[g] + bcnt[g] - 1 of the bucket order.  For q in bucket g, L(q, p) lies
    // inside FULL & ~(bmk[g] & ~mask(p)) (a column where q is above the
    // median and p is not cannot be in L), so if that set is in A0 the
    // whole bucket is skipped for p.  Waves take 64 buckets at a time,
    // keep those alive for some point, and code-test their candidates 64
    // at a time (lane e finds its bucket by a binary search of the prefix
    // sums in LDS).
    uint32_t* sLo = wlist[tid >> 6];
    uint32_t* sEx = wlist[tid >> 6] + 64;
    uint32_t* sAl = wlist[tid >> 6] + 128;
    uint32_t mp[P];
#pragma unroll
    for (int pp = 0; pp < P; ++pp) mp[pp] = pc[pp].x >> 16;
    for (uint32_t g0 = (uint32_t)(tid >> 6) * 64; g0 < ix.nbk; g0 += (uint32_t)NT) {
      const uint32_t g = g0 + (uint32_t)lane;
      uint32_t lo = 0, cnt = 0, al = 0;
      if (g < ix.nbk) {
        const uint32_t mg = ix.bmk[g];
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          const uint32_t ub = FULL & ~(mg & ~mp[pp]);
          al |= ((~(A0[pp][ub >> 5] >> (ub & 31))) & 1u) << pp;
        }
        if (al) {
          lo = ix.bstart[g];
          cnt = ix.bcnt[g];
        }
      }
      uint32_t inc = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
      }
      const uint32_t total = __shfl(inc, 63, 64);
      if (total == 0) continue;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      sLo[lane] = lo;
      sEx[lane] = inc - cnt;
      sAl[lane] = al;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (uint32_t e0 = 0; e0 < total; e0 += 64) {
        const uint32_t e = e0 + (uint32_t)lane;
        uint32_t hit = 0, j = 0;
        if (e < total) {
          int l = 0;  // the last lane whose range starts at or before e
#pragma unroll
          for (int step = 32; step; step >>= 1)
            if (sEx[l + step] <= e) l += step;
          j = sLo[l] + (e - sEx[l]);
          const uint32_t la = sAl[l];
          const uint2 qc = ix.bcode[j];
#pragma unroll
          for (int pp = 0; pp < P; ++pp) {
            const uint32_t ub = planes_le(qc, pc[pp]) & FULL;
            const uint32_t w = A0[pp][ub >> 5];
            hit |= ((~(w >> (ub & 31)) & (la >> pp)) & 1u) << pp;
          }
        }
        if (__ballot(hit != 0u) == 0ull) continue;  // uniform
        // the cover candidates were marked in pass 1a already
        const uint32_t b = hit ? ix.bidx[j] : 0u;
        if (b < kc) hit = 0;
#pragma unroll 1
        for (int pp = 0; pp < P; ++pp) append(hit, pp, pp, b);
        if (nq >= 64) drain();
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  } else {
    scan(kc, m);
  }
  drain();
  __syncthreads();

  // closures: A0 and A1 = A0|A1 down; TE down on R, up on Z; U superset-OR.
  // All four in one sweep of levels, one barrier per level.
  for (int k = tid; k < W; k += NT) {
#pragma unroll
    for (int pp = 0; pp < P; ++pp) {
      const uint32_t a0 = A0[pp][k];
      A0[pp][k] = low_closure(a0, D);
      if constexpr (!TIES) continue;
      A1[pp][k] = low_closure(A1[pp][k] | a0, D);
      uint32_t t = TE[pp][k];
      const uint32_t z = __builtin_amdgcn_readfirstlane(s_zm[pp]);
      if (D > 0) t |= (z & 1u) ? (t & 0x55555555u) << 1 : (t & 0xAAAAAAAAu) >> 1;
      if (D > 1) t |= (z & 2u) ? (t & 0x33333333u) << 2 : (t & 0xCCCCCCCCu) >> 2;
      if (D > 2) t |= (z & 4u) ? (t & 0x0F0F0F0Fu) << 4 : (t & 0xF0F0F0F0u) >> 4;
      if (D > 3) t |= (z & 8u) ? (t & 0x00FF00FFu) << 8 : (t & 0xFF00FF00u) >> 8;
      if (D > 4) t |= (z & 16u) ? (t & 0x0000FFFFu) << 16 : (t & 0xFFFF0000u) >> 16;
      TE[pp][k] = t;
    }
  }
  constexpr int kLevels = (D > 5 ? D - 5 : 0) > ZB ? (D > 5 ? D - 5 : 0) : ZB;
#pragma unroll 1
  for (int lv = 0; lv < kLevels; ++lv) {
    __syncthreads();
    if (lv + 5 < D) {  // uniform: word level j = lv + 5
      const int sh = lv;
      for (int t = tid; t < W / 2; t += NT) {
        const int lo = ((t >> sh) << (sh + 1)) | (t & ((1 << sh) - 1));
        const int hi = lo | (1 << sh);
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          A0[pp][lo] |= A0[pp][hi];
          if constexpr (!TIES) continue;
          A1[pp][lo] |= A1[pp][hi];
          if (__builtin_amdgcn_readfirstlane(s_zm[pp]) >> (lv + 5) & 1u)
            TE[pp][hi] |= TE[pp][lo];
          else
            TE[pp][lo] |= TE[pp][hi];
        }
      }
    }
    if (lv < ZB) {  // uniform: U level lv
      for (int i = tid; i < NU / 2; i += NT) {
        const int lo = ((i >> lv) << (lv + 1)) | (i & ((1 << lv) - 1));
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
          if (lv < __popc(__builtin_amdgcn_readfirstlane(s_zm[pp])))
            U[pp][lo] |= U[pp][lo | (1 << lv)];
      }
    }
  }
  __syncthreads();

  // per point: final words; X bits are parked in TE for the second pass
#pragma unroll 1
  for (int pp = 0; pp < P; ++pp) {
    // the point's values are the same for every lane: scalar registers
    const uint32_t z = __builtin_amdgcn_readfirstlane(s_zm[pp]);
    const uint32_t rg = __builtin_amdgcn_readfirstlane(s_rgt[pp]);
    const uint32_t lowR = ~z & 31u;
    uint32_t m0 = 0, m1 = 0;  // bits b whose low columns outside Z number 0 / 1
    low_subsets(~lowR & 31u, &m0, &m1);
    for (int k = tid; k < W; k += NT) {
      const int ph = __popc(((uint32_t)k << 5) & ~z);
      const uint32_t r0 = ph == 0 ? m0 : 0u;
      const uint32_t r1 = ph == 0 ? m1 : (ph == 1 ? m0 : 0u);
      const uint32_t r2 = ~(r0 | r1);
      const uint32_t a0 = A0[pp][k];
      const uint32_t a1 = TIES ? A1[pp][TIES ? k : 0] : a0;
      const uint32_t t = TIES ? TE[pp][TIES ? k : 0] : 0u;
      const uint32_t sc = scope_mask(k, D, max_d);
      uint32_t w = ((a1 & ~t) | (a0 & t)) & r2;
      const uint32_t x = t & a1 & ~a0 & r2 & sc;
      for (uint32_t bits = r1 & sc; bits; bits &= bits - 1u) {
        const int b = __ffs(bits) - 1;
        const uint32_t s = ((uint32_t)k << 5) | (uint32_t)b;
        const uint32_t sr = s & ~z, sz = s & z;
        const int j = __ffs(sr) - 1;
        const uint32_t v = sz ? (U[pp][zcompress(sz, z)] >> j) & 1u : (rg >> j) & 1u;
        w |= v << b;
      }
      if constexpr (TIES) {
        A1[pp][k] = (w & sc) | ~sc;
        TE[pp][k] = x;
        if (x) s_any[pp] = 1;
      } else {
        A0[pp][k] = (w & sc) | ~sc;  // the same thread read it above
      }
    }
  }
  __syncthreads();

  // second pass for X (residual-tie subspaces): uniform per point.  A
  // candidate can settle an X subspace s only if s lies in L(q, p), inside
  // its code bound ub, and if it ties p in a column outside Z, where their
  // codes are equal too.  So X is closed upwards into A0 (free by now) and a
  // candidate whose ub is not in that closure, or whose codes equal p's in
  // no column outside Z, is skipped before its row is read.
#pragma unroll 1
  for (int pp = 0; pp < (TIES ? P : 0); ++pp) {
    if (!s_any[pp] || !((s_ok >> pp) & 1u)) continue;
    const uint32_t z = s_zm[pp];
    const uint2 pcp = codes[s_pt[pp]];
    for (int k = tid; k < W; k += NT) {
      uint32_t x = TE[pp][k];
      if (D > 0) x |= (x & 0x55555555u) << 1;
      if (D > 1) x |= (x & 0x33333333u) << 2;
      if (D > 2) x |= (x & 0x0F0F0F0Fu) << 4;
      if (D > 3) x |= (x & 0x00FF00FFu) << 8;
      if (D > 4) x |= (x & 0x0000FFFFu) << 16;
      A0[pp][k] = x;
    }
#pragma unroll 1
    for (int j = 5; j < D; ++j) {
      __syncthreads();
      const int sh = j - 5;
      for (int t = tid; t < W / 2; t += NT) {
        const int lo = ((t >> sh) << (sh + 1)) | (t & ((1 << sh) - 1));
        A0[pp][lo | (1 << sh)] |= A0[pp][lo];
      }
    }
#pragma unroll 1
    while (true) {
      __syncthreads();
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      for (int k = tid; k < W; k += NT) {
        uint32_t x = TE[pp][k];
        while (x) {
          const int bpos = __ffs(x) - 1;
          const int slot = atomicAdd(&s_cnt, 1);
          if (slot >= XCAP) break;
          xl[slot] = (uint16_t)((k << 5) | bpos);
          x &= x - 1u;
        }
        TE[pp][k] = x;
      }
      __syncthreads();
      const int cnt = min(s_cnt, XCAP);
      if (cnt == 0) break;
      // candidate b with codes qc: the filters, then the exact test
      auto settle = [&](uint32_t b, uint2 qc) {
        const uint32_t ub = planes_le(qc, pcp) & FULL;
        if (!((A0[pp][ub >> 5] >> (ub & 31)) & 1u)) return;
        const uint32_t ex = ~(qc.x ^ pcp.x), ey = ~(qc.y ^ pcp.y);
        if (((ex >> 16) & ex & (ey >> 16) & ey & FULL & ~z) == 0) return;
        uint32_t lt, eq;
        pair_masks<D>(erows + (size_t)b * D, prow[pp], lt, eq);
        if (lt == 0 || (eq & ~z) == 0) return;
        const uint32_t le = lt | eq;
        for (int e = 0; e < cnt; ++e) {
          const uint32_t s = xl[e];
          if (!(s & ~le) && (s & lt)) lds_mark(A1[pp], s);
        }
      };
      if (order == kOrderGroup || order == kOrderSub) {
        // a group's code bound contains each member's, and the closure of X
        // in A0 is upward closed: a group whose bound is outside it holds
        // no candidate that passes the first filter
        const int wv = tid >> 6;
        const uint32_t ng = (m + 63) / 64;
        for (uint32_t g0 = (uint32_t)wv * 64; g0 < ng; g0 += (uint32_t)NT) {
          const uint32_t g = g0 + (uint32_t)lane;
          bool al = false;
          if (g < ng) {
            const uint32_t ubg = planes_le(ix.gsum[g], pcp) & FULL;
            al = (A0[pp][ubg >> 5] >> (ubg & 31)) & 1u;
          }
          for (uint64_t live = __ballot(al); live; live &= live - 1ull) {  // uniform
            const uint32_t j = (g0 + (uint32_t)(__ffsll((long long)live) - 1)) * 64u +
                               (uint32_t)lane;
            if (j < m) settle(ix.gidx[j], ix.gcode[j]);
          }
        }
      } else {
        for (uint32_t b = tid; b < m; b += NT) settle(b, codes[b]);
      }
    }
  }
  __syncthreads();

  if constexpr (!TIES) {  // points that met a residual tie go to the full kernel
    if (tid < P && s_tie[tid] && ((s_ok >> tid) & 1u))
      ix.defer[1 + atomicAdd(ix.defer, 1u)] = s_pt[tid];
  }
#pragma unroll
  for (int pp = 0; pp < P; ++pp) {
    if (!((s_ok >> pp) & 1u) || (!TIES && s_tie[pp])) continue;
    uint32_t(*R)[W] = TIES ? reinterpret_cast<uint32_t(*)[W]>(&A1[0][0]) : A0;  // the final words
    const uint32_t row = pos[s_pt[pp]];
    // streaming stores: the bitmaps are not read again here, and keeping
    // them out of L2 leaves it to the candidate codes the other workgroups
    // are scanning
    uint32_t cnt = 0;
    for (int k = tid; k < W; k += NT) {
      const uint32_t v = R[pp][k];
      cnt += v != 0xFFFFFFFFu;
      __builtin_nontemporal_store(v, &dom[(size_t)row * W + k]);
    }
    if (ix.rowcnt) {  // uniform
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
      if ((tid & 63) == 0 && cnt) atomicAdd(&ix.rowcnt[row], cnt);
    }
  }
  __syncthreads();  // the LDS may be reused by the caller's next block
}

// Tiles are taken in block order: the workgroups in flight then hold one
// narrow window of the code order, spread over the eight XCDs.  (Giving each
// XCD a contiguous run of tiles instead was measured slower: the cost of a
// tile varies along the code order, so the XCDs finish unevenly.)
//
// At d = 16 (one point per workgroup with the tie lattices, two without
// them) the kernel is held to eight waves per SIMD, four workgroups per CU;
// at d = 13 likewise (measured 14% faster there; no change at d = 10, 12,
// 14, where the registers or the LDS already set the occupancy).
template <int D, bool TIES>
constexpr int member_waves() {
  return D == 16 || D == 13 ? 8 : 1;
}
template <int D, int P, int NT, int ORDER, int ZB = zbits<D>(), bool TIES = true>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(member_waves<D, TIES>())))
void k_member(const float* __restrict__ erows, const uint2* __restrict__ codes,
              const uint32_t* __restrict__ pos, uint32_t m, int max_d,
              const uint32_t* __restrict__ minkey, uint32_t* __restrict__ dom, uint32_t cover,
              CandIndex ix) {
  member_body<D, P, NT, ZB, TIES>(erows, codes, pos, m, max_d, minkey, dom, blockIdx.x * P,
                                  cover, ORDER, ix);
}

// The points (score-order positions, m of them) with more than zb columns at
// the column minimum, appended to list[1 ..] (count list[0], which starts at
// zero): the points a k_member launch with a 2^zb-word boundary table
// leaves to k_member_deferred (the same test as member_body's).
template <int D>
__global__ void k_defer_list(const float* __restrict__ erows, uint32_t m,
                             const uint32_t* __restrict__ minkey, int zb,
                             uint32_t* __restrict__ list) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= m) return;
  int nz = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) nz += order_key(erows[(size_t)a * D + k]) == minkey[k];
  if (nz > zb) list[1 + atomicAdd(list, 1u)] = a;
}

// The points a k_member launch with a narrower boundary table deferred:
// list[0] of them at list[1 ..], P per workgroup, the grid looping over the
// list (its length is known on the device only).
template <int D, int P, int NT, int ORDER>
__global__ __launch_bounds__(NT) void k_member_deferred(
    const float* __restrict__ erows, const uint2* __restrict__ codes,
    const uint32_t* __restrict__ pos, uint32_t m, int max_d, const uint32_t* __restrict__ minkey,
    uint32_t* __restrict__ dom, uint32_t cover, const uint32_t* __restrict__ list, CandIndex ix) {
  const uint32_t n = list[0];
  ix.pts = list + 1;
  ix.npts = n;
  ix.defer = nullptr;
  ix.defer_listed = false;
  for (uint32_t base = blockIdx.x * P; base < n; base += gridDim.x * P)
    member_body<D, P, NT>(erows, codes, pos, m, max_d, minkey, dom, base, cover, ORDER, ix);
}

// ---------------------------------------------------------------------------
// Code order of the candidates (for kOrderGroup).

// 64-bit key of the bit-plane codes, planes 3 to 0 from the top (D bits
// each), and the candidate's position.
template <int D>
__global__ void k_plane_keys(const uint2* __restrict__ codes, uint32_t m,
                             uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint2 c = codes[j];
  key[j] = ((uint64_t)(c.x >> 16) << (3 * D)) | ((uint64_t)(c.y >> 16) << (2 * D)) |
           ((uint64_t)(c.x & 0xFFFFu) << D) | (uint64_t)(c.y & 0xFFFFu);
  val[j] = j;
}

// Group of the code order at positions j0 .. j1-1: the codes in that order
// (gcode[j] = codes[sidx[j]]), the candidate ids minus bias (gidx, if
// given) and the per-column minimum code as bit planes (*gsum).
// Bit planes of per-column codes (4 bits each).
template <int D>
__device__ __forceinline__ uint2 planes_of(const uint32_t* mn) {
  uint32_t pl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < D; ++k)
#pragma unroll
    for (int b = 0; b < 4; ++b) pl[b] |= ((mn[k] >> b) & 1u) << k;
  return make_uint2((pl[3] << 16) | pl[1], (pl[2] << 16) | pl[0]);
}

// ... and the same for each subgroup of 8 (ssum[0 .. 7]; a subgroup past j1
// gets all-15 codes).
template <int D>
__device__ __forceinline__ void group_sum(const uint2* __restrict__ codes,
                                          const uint32_t* __restrict__ sidx, uint32_t j0,
                                          uint32_t j1, uint32_t bias, uint2* __restrict__ gcode,
                                          uint32_t* __restrict__ gidx, uint2* __restrict__ gsum,
                                          uint2* __restrict__ ssum) {
  uint32_t mn[D], sm[D];
#pragma unroll
  for (int k = 0; k < D; ++k) mn[k] = 15;
  for (int s = 0; s < 8; ++s) {
#pragma unroll
    for (int k = 0; k < D; ++k) sm[k] = 15;
    const uint32_t a = j0 + 8u * (uint32_t)s;
    for (uint32_t j = a; j < min(j1, a + 8u); ++j) {
      const uint32_t i = sidx[j];
      const uint2 c = codes[i];
      gcode[j] = c;
      if (gidx) gidx[j] = i - bias;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const uint32_t v = ((c.x >> (16 + k)) & 1u) << 3 | ((c.y >> (16 + k)) & 1u) << 2 |
                           ((c.x >> k) & 1u) << 1 | ((c.y >> k) & 1u);
        sm[k] = min(sm[k], v);
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) mn[k] = min(mn[k], sm[k]);
    ssum[s] = planes_of<D>(sm);
  }
  *gsum = planes_of<D>(mn);
}

// Codes in code order and the summaries of each group of 64 (one thread per
// group): gsum[g], and its subgroups' at gsum[ng + 8g ..].
template <int D>
__global__ void k_group_sum(const uint2* __restrict__ codes, const uint32_t* __restrict__ gidx,
                            uint32_t m, uint2* __restrict__ gcode, uint2* __restrict__ gsum) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ng = (m + 63) / 64;
  if (g >= ng) return;
  group_sum<D>(codes, gidx, g * 64, min(m, g * 64 + 64), 0u, gcode, nullptr, gsum + g,
               gsum + ng + 8u * g);
}

// HashCube export: (word index, word value) -> point ids.  The bitmaps are
// one flat array of ne * W words (point-major); a tile is 4096 consecutive
// words, 16 per thread read as four 16-byte loads, so every wave reads 4 KB
// of contiguous memory.  Entries are the words that are not all ones.

constexpr int kHcItems = 16;
constexpr uint32_t kHcTile = 256 * kHcItems;  // 2^12 words
static_assert(kHcTile == 4096, "per-row offsets assume tiles of 2^12 words");

__device__ __forceinline__ void hc_load(const uint32_t* __restrict__ dom, uint64_t total,
                                        uint64_t k0, uint32_t (&v)[kHcItems]) {
  if (k0 + kHcItems <= total) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* p = reinterpret_cast<const u32x4*>(dom + k0);
#pragma unroll
    for (int q = 0; q < kHcItems / 4; ++q) {
      const u32x4 x = __builtin_nontemporal_load(p + q);
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int r = 0; r < kHcItems; ++r) v[r] = k0 + r < total ? dom[k0 + r] : 0xFFFFFFFFu;
  }
}

// tsum[tile] = entries of the tile.
__global__ __launch_bounds__(256) void k_hc_count(const uint32_t* __restrict__ dom, uint64_t total,
                                                  uint32_t* __restrict__ tsum) {
  __shared__ uint32_t red[4];
  const uint64_t k0 = (uint64_t)blockIdx.x * kHcTile + (uint64_t)threadIdx.x * kHcItems;
  uint32_t v[kHcItems];
  hc_load(dom, total, k0, v);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kHcItems; ++r) s += v[r] != 0xFFFFFFFFu;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Entries (key = j << 32 | value, id of point i) of word k = i * W + j, in
// ascending k from the tile's scanned offset toff[tile << tshift] (W =
// 2^wbits; tshift 0 with per-tile offsets, 12 - wbits with per-row ones: a
// tile holds 2^(12 - wbits) whole rows).
__global__ __launch_bounds__(256) void k_hc_emit(const uint32_t* __restrict__ dom, uint64_t total,
                                                 int wbits, const uint32_t* __restrict__ ids,
                                                 const uint32_t* __restrict__ toff, int tshift,
                                                 uint64_t* __restrict__ key,
                                                 uint32_t* __restrict__ val) {
  __shared__ uint32_t red[4];
  const uint64_t k0 = (uint64_t)blockIdx.x * kHcTile + (uint64_t)threadIdx.x * kHcItems;
  uint32_t v[kHcItems];
  hc_load(dom, total, k0, v);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kHcItems; ++r) s += v[r] != 0xFFFFFFFFu;
  // exclusive prefix of s over the workgroup
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t inc = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) red[wv] = inc;
  __syncthreads();
  uint32_t o = toff[(uint64_t)blockIdx.x << tshift] + inc - s;
  for (int w = 0; w < wv; ++w) o += red[w];
  const uint64_t wmask = (1ull << wbits) - 1ull;
#pragma unroll
  for (int r = 0; r < kHcItems; ++r) {
    if (v[r] == 0xFFFFFFFFu) continue;
    const uint64_t k = k0 + r;
    key[o] = ((k & wmask) << 32) | v[r];
    val[o] = ids[k >> wbits];
    ++o;
  }
}

// Splits run keys into (word, value) and run lengths into 64-bit offsets.
__global__ void k_hc_groups(const uint64_t* __restrict__ ukey, uint64_t ng,
                            uint32_t* __restrict__ gw, uint32_t* __restrict__ gv) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  gw[g] = (uint32_t)(ukey[g] >> 32);
  gv[g] = (uint32_t)ukey[g];
}

}  // namespace
}  // namespace sky
