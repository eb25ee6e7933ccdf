// batch.cuh -- many independent skycube instances in one pass of launches.
//
// Instances of one dimensionality D form a group; a group's points are
// contiguous, instance t owning points off[t] .. off[t+1].  Every stage is
// the single-instance stage applied per segment, with the segment found from
// the offsets; counts that only the device knows (prefilter survivors,
// extended-skyline sizes) stay on the device, so a whole batch enqueues with
// no host synchronisation and can be captured in a HIP graph.  The
// membership stage is one persistent launch per group fed by a device work
// queue of (instance, block of points) tasks.
#pragma once

#include "kernels.cuh"

namespace sky {
namespace {

// Segment of index i: the largest t < T with off[t] <= i (off ascending,
// T + 1 entries); empty segments are skipped over.
__device__ __forceinline__ uint32_t seg_of(const uint32_t* __restrict__ off, uint32_t T,
                                           uint32_t i) {
  uint32_t lo = 0, hi = T;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Graph-safe fills (a captured graph gets no memset nodes).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, uint64_t n, uint8_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// The group's rows, contiguous: point i of the group is row i - off[t] of
// instance t, whose rows start at device address src[t] (instances need not
// share an allocation).
template <int D>
__global__ void k_bcopy(const uint64_t* __restrict__ src, const uint32_t* __restrict__ off,
                        uint32_t T, uint32_t N, float* __restrict__ dst) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t t = seg_of(off, T, i);
  const float* row = reinterpret_cast<const float*>(src[t]) + (size_t)(i - off[t]) * D;
#pragma unroll
  for (int k = 0; k < D; ++k) dst[(size_t)i * D + k] = row[k];
}

// Sort keys: instance in the high word, the Manhattan-norm key in the low
// word (k_score's key), so one radix sort orders every instance by norm.
template <int D>
__global__ void k_bscore(const float* __restrict__ rows, const uint32_t* __restrict__ off,
                         uint32_t T, uint32_t N, uint64_t* __restrict__ key,
                         uint32_t* __restrict__ val, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t t = seg_of(off, T, i);
  float sum = 0.0f;
  bool nan = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const float v = rows[(size_t)i * D + k];
    nan |= __builtin_isnan(v);
    sum += v;
  }
  if (nan) atomicOr(err, 2u);
  key[i] = ((uint64_t)t << 32) | norm_key(sum);
  val[i] = i;
}

template <int D>
__global__ void k_bgather(const float* __restrict__ rows, const uint64_t* __restrict__ skey,
                          const uint32_t* __restrict__ sval, uint32_t N, float* __restrict__ srows,
                          uint32_t* __restrict__ skeys, uint32_t* __restrict__ spid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t g = sval[i];
#pragma unroll
  for (int k = 0; k < D; ++k) srows[(size_t)i * D + k] = rows[(size_t)g * D + k];
  skeys[i] = (uint32_t)skey[i];
  spid[i] = g;
}

// Prefilter, one workgroup per task (instance t, first position): the
// points of the task against the instance's npr_max lowest-norm points.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bprune(const float* __restrict__ srows,
                                                   const uint32_t* __restrict__ skeys,
                                                   const uint32_t* __restrict__ off,
                                                   const uint2* __restrict__ tasks,
                                                   uint32_t npr_max,
                                                   uint8_t* __restrict__ alive) {
  __shared__ float tile[kBlock * D];
  __shared__ uint32_t tkey[kBlock];
  const uint2 task = tasks[blockIdx.x];
  const uint32_t b0 = off[task.x], nt = off[task.x + 1] - b0;
  const uint32_t i = task.y + threadIdx.x;
  const bool valid = i < b0 + nt;
  float p[D];
  uint32_t kp = 0;
  if (valid) {
#pragma unroll
    for (int k = 0; k < D; ++k) p[k] = srows[(size_t)i * D + k];
    kp = skeys[i];
  }
  const uint32_t npr = min(nt, npr_max);
  bool dead = false;
  for (uint32_t t0 = 0; t0 < npr; t0 += kBlock) {
    const uint32_t cnt = min((uint32_t)kBlock, npr - t0);
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < cnt * D; e += kBlock)
      tile[e] = srows[(size_t)(b0 + t0) * D + e];
    if (threadIdx.x < cnt) tkey[threadIdx.x] = skeys[b0 + t0 + threadIdx.x];
    __syncthreads();
    if (valid && !dead) {
      for (uint32_t t = 0; t < cnt; ++t) {
        if (tkey[t] > kp) break;
        if (b0 + t0 + t == i) continue;
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

// Gathers at the positions idx[j], j < *count.
template <int D>
__global__ void k_bgather_idx(const float* __restrict__ srows, const uint32_t* __restrict__ skeys,
                              const uint32_t* __restrict__ spid, const uint32_t* __restrict__ idx,
                              const uint32_t* __restrict__ count, float* __restrict__ orows,
                              uint32_t* __restrict__ okeys, uint32_t* __restrict__ opid) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *count) return;
  const uint32_t i = idx[j];
#pragma unroll
  for (int k = 0; k < D; ++k) orows[(size_t)j * D + k] = srows[(size_t)i * D + k];
  if (okeys) okeys[j] = skeys[i];
  opid[j] = spid[i];
}

// Quantile pivots per instance and column (block = (t, column)): a strided
// sample of up to 2048 of the instance's rows (as many as the
// single-instance path takes), bitonic-sorted in LDS; the values of ranks
// (b+1) * S / 16 become the pivots.  Instances without rows get pivots above
// every value.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bpivots(const float* __restrict__ rows,
                                                    const uint32_t* __restrict__ off, uint32_t T,
                                                    uint32_t* __restrict__ piv) {
  constexpr uint32_t S = 2048;
  __shared__ uint32_t val[S];
  const uint32_t t = blockIdx.x / D;
  const int col = blockIdx.x % D;
  const uint32_t b0 = off[t], m = off[t + 1] - b0;
  uint32_t* out = piv + ((size_t)t * 16 + col) * kCodePivots;
  if (m == 0) {
    if (threadIdx.x < kCodePivots) out[threadIdx.x] = 0xFFFFFFFFu;
    return;
  }
  const uint32_t cnt = min(S, m);
  uint32_t p2 = 1;
  while (p2 < cnt) p2 <<= 1;
  for (uint32_t i = threadIdx.x; i < p2; i += kBlock)
    val[i] = i < cnt ? order_key(rows[(size_t)(b0 + (uint64_t)i * m / cnt) * D + col]) : 0xFFFFFFFFu;
  __syncthreads();
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < p2; i += kBlock) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint32_t a = val[i], b = val[l];
          if ((a > b) == ((i & k) == 0)) {
            val[i] = b;
            val[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x < kCodePivots) out[threadIdx.x] = val[(threadIdx.x + 1) * cnt / (kCodePivots + 1)];
}

template <int D>
__global__ void k_bmasks(const float* __restrict__ rows, const uint32_t* __restrict__ off,
                         uint32_t T, const uint32_t* __restrict__ count,
                         const uint32_t* __restrict__ piv, uint32_t* __restrict__ masks) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *count) return;
  const uint32_t t = seg_of(off, T, j);
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < D; ++k)
    w |= (uint32_t)(order_key(rows[(size_t)j * D + k]) >
                    piv[((size_t)t * 16 + k) * kCodePivots + 7]) << k;
  masks[j] = w;
}

template <int D>
__global__ void k_bcodes(const float* __restrict__ erows, const uint32_t* __restrict__ off,
                         uint32_t T, const uint32_t* __restrict__ count,
                         const uint32_t* __restrict__ piv, uint2* __restrict__ codes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *count) return;
  const uint32_t t = seg_of(off, T, i);
  uint32_t pl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const uint32_t key = order_key(erows[(size_t)i * D + k]);
    const uint32_t* pv = piv + ((size_t)t * 16 + k) * kCodePivots;
    uint32_t c = 0;
#pragma unroll
    for (int b = 0; b < kCodePivots; ++b) c += pv[b] < key;
#pragma unroll
    for (int b = 0; b < 4; ++b) pl[b] |= ((c >> b) & 1u) << k;
  }
  codes[i] = make_uint2((pl[3] << 16) | pl[1], (pl[2] << 16) | pl[0]);
}

// Survivor pass, one wave64 per survivor j < *count, against the survivors
// of its own instance (self_scan, shared with the single-instance kernel);
// need_dom also tells dominated (1) from undominated (0) survivors.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bself(const float* __restrict__ rows,
                                                  const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ masks,
                                                  const uint32_t* __restrict__ off, uint32_t T,
                                                  const uint32_t* __restrict__ count, int need_dom,
                                                  uint8_t* __restrict__ status) {
  __shared__ uint32_t qbuf[kBlock / 64][192];
  const uint32_t j = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (j >= *count) return;  // whole wave; no block barrier in this kernel
  const uint32_t t = seg_of(off, T, j);
  const uint8_t st = self_scan<D>(rows, keys, masks, off[t], off[t + 1], j, need_dom,
                                  qbuf[threadIdx.x >> 6]);
  if ((threadIdx.x & 63) == 0) status[j] = st;
}

// ---------------------------------------------------------------------------
// Mask index of a group, per instance: entries j < *count are keyed
// (instance << D) | median mask (plane 3 of their codes) and radix-sorted
// (stable, so score order holds inside a bucket); entries past *count get
// all-ones keys and sort last.

template <int D>
__global__ void k_bidx_keys(const uint2* __restrict__ codes, const uint32_t* __restrict__ off,
                            uint32_t T, const uint32_t* __restrict__ count, uint32_t N,
                            uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  val[j] = j;
  key[j] = j < *count ? (seg_of(off, T, j) << D) | (codes[j].x >> 16) : 0xFFFFFFFFu;
}

// boff[s] = number of sorted keys (among the first *count) below s, for
// s = 0 .. nb: bucket (t, S) is boff[(t << D) + S] .. boff[(t << D) + S + 1].
__global__ void k_bidx_bounds(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ count,
                              uint32_t nb, uint32_t* __restrict__ boff) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nb) return;
  uint32_t a = 0, z = *count;
  while (a < z) {
    const uint32_t mid = (a + z) >> 1;
    if (skey[mid] < s) a = mid + 1;
    else z = mid;
  }
  boff[s] = a;
}

// In bucket order, j < *count: the keys (if keys), the codes, and (if off)
// each entry's index inside its instance.
__global__ void k_bidx_gather(const uint32_t* __restrict__ sval, const uint32_t* __restrict__ keys,
                              const uint2* __restrict__ codes, const uint32_t* __restrict__ count,
                              const uint32_t* __restrict__ off, uint32_t T,
                              uint32_t* __restrict__ bkey, uint2* __restrict__ bcode,
                              uint32_t* __restrict__ blocal) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *count) return;
  const uint32_t i = sval[j];
  if (keys) bkey[j] = keys[i];
  bcode[j] = codes[i];
  if (off) blocal[j] = i - off[seg_of(off, T, i)];
}

// head[j] = 1 where a run of equal sorted keys starts (j < *count), 0 up to N.
__global__ void k_bidx_heads(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ count,
                             uint32_t N, uint8_t* __restrict__ head) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  head[j] = j < *count && (j == 0 || skey[j] != skey[j - 1]);
}

// The *nb buckets from their run heads: mask bmk[g], first position
// bstart[g], size bcnt[g], instance brun[g].
template <int D>
__global__ void k_bidx_list(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ heads,
                            const uint32_t* __restrict__ nb, const uint32_t* __restrict__ count,
                            uint32_t N, uint32_t* __restrict__ bmk, uint32_t* __restrict__ bstart,
                            uint32_t* __restrict__ bcnt, uint32_t* __restrict__ brun) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = *nb;
  if (g >= n || g >= N) return;
  const uint32_t h = heads[g];
  const uint32_t key = skey[h];
  bstart[g] = h;
  bcnt[g] = (g + 1 < n ? heads[g + 1] : *count) - h;
  bmk[g] = key & ((1u << D) - 1u);
  brun[g] = key >> D;
}

// gfirst[t] = first bucket of instance t, t = 0 .. T (gfirst[T] = *nb).
__global__ void k_bidx_first(const uint32_t* __restrict__ brun, const uint32_t* __restrict__ nb,
                             uint32_t T, uint32_t* __restrict__ gfirst) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > T) return;
  uint32_t a = 0, z = *nb;
  while (a < z) {
    const uint32_t mid = (a + z) >> 1;
    if (brun[mid] < t) a = mid + 1;
    else z = mid;
  }
  gfirst[t] = a;
}

// Survivor pass over the mask index, one wave64 per survivor j < *count,
// against the survivors of its own instance (self_idx, shared with the
// single-instance kernel; boff is the instance-major bucket table).  With
// need_dom, dominated (1) and undominated (0) survivors are told apart and
// minkey is not used.  NT threads per workgroup; the waves stride over the
// survivors when the grid holds fewer.  (The SDSC levels run it with one
// wave per workgroup and per survivor: 12% faster there, where a finished
// wave's slot is taken at once by the next survivor; the batched MDMC pass
// with four waves, its grid then needing a quarter of the workgroups.)
template <int D, int NT = kBlock>
__global__ __launch_bounds__(NT) void k_bself_idx(
    const float* __restrict__ rows, const uint32_t* __restrict__ keys,
    const uint2* __restrict__ codes, const uint32_t* __restrict__ off, uint32_t T,
    const uint32_t* __restrict__ count, const uint32_t* __restrict__ boff,
    const uint32_t* __restrict__ sval, const uint32_t* __restrict__ bkey,
    const uint2* __restrict__ bcode, const uint32_t* __restrict__ minkey, int need_dom,
    uint8_t* __restrict__ status) {
  __shared__ uint32_t sLo[NT / 64][64];
  __shared__ uint32_t sEx[NT / 64][64];
  const int wv = threadIdx.x >> 6;
  const uint32_t cnt = *count;
  const uint32_t stride = gridDim.x * (uint32_t)(NT / 64);
  // whole waves; no block barrier in this kernel
  for (uint32_t j = blockIdx.x * (uint32_t)(NT / 64) + (uint32_t)wv; j < cnt; j += stride) {
    const uint32_t t = seg_of(off, T, j);
    const uint8_t st = self_idx<D>(rows, keys, codes, boff + ((size_t)t << D), sval, bkey, bcode,
                                   j, need_dom, need_dom ? nullptr : minkey + (size_t)t * 16,
                                   sLo[wv], sEx[wv]);
    if ((threadIdx.x & 63) == 0) status[j] = st;
  }
}

// SDSC chunk: the rows of candidate i of instance t (cuboid subs[t]),
// projected onto the D columns of the cuboid, ascending.
template <int D>
__global__ void k_sdsc_rows(const float* __restrict__ rows, uint32_t d,
                            const uint32_t* __restrict__ cand, const uint32_t* __restrict__ subs,
                            const uint32_t* __restrict__ off, uint32_t T, uint32_t N,
                            float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t t = seg_of(off, T, i);
  const float* row = rows + (size_t)cand[i] * d;
  uint32_t sub = subs[t];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int col = __ffs(sub) - 1;
    sub &= sub - 1u;
    out[(size_t)i * D + k] = row[col];
  }
}

// SDSC chunk: per candidate position, keep = in sky or ext (not strictly
// dominated), sky = in the skyline; both cleared up to N first.
__global__ void k_sdsc_flags(const uint8_t* __restrict__ status, const uint32_t* __restrict__ spid,
                             const uint32_t* __restrict__ count, uint8_t* __restrict__ keep,
                             uint8_t* __restrict__ sky) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *count) return;
  const uint8_t st = status[j];
  keep[spid[j]] = st != 2;
  sky[spid[j]] = st == 0;
}

// Group point ids of the extended skyline as row numbers inside their
// instance.
__global__ void k_blocal(const uint32_t* __restrict__ ids, uint32_t cnt,
                         const uint32_t* __restrict__ off, uint32_t T, uint32_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  const uint32_t i = ids[j];
  out[j] = i - off[seg_of(off, T, i)];
}

// Code order per instance (for kOrderGroup): keys of the entries j < *count
// (all ones past it) with the instance above the top kbits bits of the
// 4D-bit plane key, so that one sort orders every instance segment.
template <int D>
__global__ void k_bplane_keys(const uint2* __restrict__ codes, const uint32_t* __restrict__ eoff,
                              uint32_t T, const uint32_t* __restrict__ count, uint32_t N, int kbits,
                              uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  val[j] = j;
  if (j >= *count) {
    key[j] = ~0ull;
    return;
  }
  const uint2 c = codes[j];
  const uint64_t pk = ((uint64_t)(c.x >> 16) << (3 * D)) | ((uint64_t)(c.y >> 16) << (2 * D)) |
                      ((uint64_t)(c.x & 0xFFFFu) << D) | (uint64_t)(c.y & 0xFFFFu);
  const uint64_t t = seg_of(eoff, T, j);
  key[j] = (kbits < 64 ? t << kbits : 0ull) | (pk >> (4 * D - kbits));
}

// Groups of 64 per instance: cnt[t] = ceil(m_t / 64), cnt[T] = 0 (for the
// exclusive scan into group offsets).
__global__ void k_bgroup_counts(const uint32_t* __restrict__ eoff, uint32_t T,
                                uint32_t* __restrict__ cnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > T) return;
  cnt[t] = t < T ? (eoff[t + 1] - eoff[t] + 63) / 64 : 0u;
}

// Group g (of all instances, goff[T] in total) of instance t: codes in code
// order, candidate ids local to the instance, per-column minimum codes of
// the group (gsum[g]) and of its subgroups of 8 (ssum[8g ..]).
template <int D>
__global__ void k_bgroup_sum(const uint2* __restrict__ codes, const uint32_t* __restrict__ sval,
                             const uint32_t* __restrict__ eoff, uint32_t T,
                             const uint32_t* __restrict__ goff, uint2* __restrict__ gcode,
                             uint32_t* __restrict__ glocal, uint2* __restrict__ gsum,
                             uint2* __restrict__ ssum) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= goff[T]) return;
  const uint32_t t = seg_of(goff, T, g);
  const uint32_t e0 = eoff[t], e1 = eoff[t + 1];
  const uint32_t j0 = e0 + (g - goff[t]) * 64;
  group_sum<D>(codes, sval, j0, min(e1, j0 + 64), e0, gcode, glocal, gsum + g, ssum + 8u * g);
}

// flag_pt[spid[j]] = 1 and flag_s1[j] = 1 for survivors nothing strictly
// dominates; flag_s1 is cleared up to N.
__global__ void k_bext_flags(const uint8_t* __restrict__ status, const uint32_t* __restrict__ spid,
                             const uint32_t* __restrict__ count, uint32_t N,
                             uint8_t* __restrict__ flag_pt, uint8_t* __restrict__ flag_s1) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const bool ext = j < *count && status[j] != 2;
  flag_s1[j] = ext;
  if (ext) flag_pt[spid[j]] = 1;
}

__global__ void k_bscatter_rank(const uint32_t* __restrict__ ids,
                                const uint32_t* __restrict__ count, uint32_t* __restrict__ rank) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < *count) rank[ids[i]] = i;
}

__global__ void k_bpos(const uint32_t* __restrict__ epid, const uint32_t* __restrict__ count,
                       const uint32_t* __restrict__ rank, uint32_t* __restrict__ pos) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < *count) pos[j] = rank[epid[j]];
}

// Column minima per instance as order keys, minkey[t * 16 + k] (columns
// k < D; the entries must start at ~0).  Workgroup (t, y) of a T x Y grid
// takes every Y-th stretch of kBlock rows of instance t, so that a large
// instance is spread over Y workgroups.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bcolmin(const float* __restrict__ erows,
                                                    const uint32_t* __restrict__ off,
                                                    uint32_t* __restrict__ minkey) {
  __shared__ uint32_t red[D];
  const uint32_t t = blockIdx.x;
  if (threadIdx.x < D) red[threadIdx.x] = 0xFFFFFFFFu;
  __syncthreads();
  uint32_t best[D];
#pragma unroll
  for (int k = 0; k < D; ++k) best[k] = 0xFFFFFFFFu;
  const uint32_t step = kBlock * gridDim.y;
  for (uint32_t i = off[t] + blockIdx.y * kBlock + threadIdx.x; i < off[t + 1]; i += step) {
#pragma unroll
    for (int k = 0; k < D; ++k) best[k] = min(best[k], order_key(erows[(size_t)i * D + k]));
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    uint32_t v = best[k];
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    if ((threadIdx.x & 63) == 0) atomicMin(&red[k], v);
  }
  __syncthreads();
  if (threadIdx.x < D && red[threadIdx.x] != 0xFFFFFFFFu)
    atomicMin(&minkey[(size_t)t * 16 + threadIdx.x], red[threadIdx.x]);
}

// Per-instance task counts (blocks of P points) with a zero past the end,
// for the exclusive scan that numbers the membership tasks.
__global__ void k_btasks(const uint32_t* __restrict__ off, uint32_t T, uint32_t P,
                         uint32_t* __restrict__ cnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > T) return;
  cnt[t] = t < T ? (off[t + 1] - off[t] + P - 1) / P : 0u;
}

// Membership, persistent: each workgroup takes (instance, block) tasks from
// a device counter until the group's tasks (toff[T]) are done.  With goff,
// every instance visits its candidates in code-order groups (instance t's
// groups from goff[t] on, their subgroups from ssum + 8 goff[t] on; gcode /
// glocal in the group code order); else with gfirst,
// instance t's buckets are gfirst[t] .. gfirst[t+1]-1 of the bucket list
// (bmk, bstart, bcnt over the bucket order bcode / blocal), used when they
// hold min_bucket candidates or more on average; cover as in member_body
// (0: auto_cover of each instance's extended skyline).
// ORD = kOrderGroup or kOrderSub compiles that code-order visit only (goff
// must be set); ORD = 0 picks buckets or the scan per instance;
// with workgroups of four waves it also holds the kernel to 64 VGPRs, so that
// eight waves run per SIMD (with a grid of 2048 workgroups); the few values
// that then live in scratch cost less than the doubled occupancy gains
// (bmember_waves; not from d = 11 on, where the spills grow past what the
// occupancy gains, nor for the subgroup steps with eight points per
// workgroup from d = 8 on, which need more registers).  With eight waves per
// workgroup the kernel is held to 64 VGPRs at d = 13 (eight waves per SIMD;
// a few values spill, and a large instance still runs 8-9% faster) and to
// 80 at d = 14, six waves per SIMD, as many as its LDS allows.
template <int D, int P, int NT, int ORD>
constexpr int bmember_waves() {
  return ORD != 0 && NT == 256 && D <= 10 && !(ORD == kOrderSub && P == 8 && D >= 8) ? 8
         : (ORD != 0 && NT == 512 && D == 13 ? 8 : (ORD != 0 && NT == 512 && D <= 14 ? 6 : 1));
}

template <int D, int P, int NT, int ORD = 0>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(bmember_waves<D, P, NT, ORD>())))
void k_bmember(const float* __restrict__ erows,
                                                const uint2* __restrict__ codes,
                                                const uint32_t* __restrict__ pos,
                                                const uint32_t* __restrict__ eoff, uint32_t T,
                                                const uint32_t* __restrict__ toff,
                                                const uint32_t* __restrict__ maxd,
                                                const uint32_t* __restrict__ minkey,
                                                uint32_t* __restrict__ dom,
                                                uint32_t* __restrict__ work, uint32_t cover,
                                                uint32_t head, uint32_t min_bucket,
                                                const uint32_t* __restrict__ goff,
                                                const uint2* __restrict__ gsum,
                                                const uint2* __restrict__ ssum,
                                                const uint2* __restrict__ gcode,
                                                const uint32_t* __restrict__ glocal,
                                                const uint32_t* __restrict__ gfirst,
                                                const uint32_t* __restrict__ bmk,
                                                const uint32_t* __restrict__ bstart,
                                                const uint32_t* __restrict__ bcnt,
                                                const uint32_t* __restrict__ blocal,
                                                const uint2* __restrict__ bcode) {
  __shared__ uint32_t s_task;
  const uint32_t total = toff[T];
  while (true) {
    if (threadIdx.x == 0) s_task = atomicAdd(work, 1u);
    __syncthreads();
    // the same for every lane: scalar, so that the task's instance, offsets
    // and pointers below live in scalar registers
    const uint32_t task = __builtin_amdgcn_readfirstlane(s_task);
    __syncthreads();
    if (task >= total) break;  // uniform
    const uint32_t t = seg_of(toff, T, task);
    const uint32_t e0 = eoff[t], m = eoff[t + 1] - e0;
    uint32_t g0 = 0, nb = 0;
    constexpr bool GROUPS = ORD != 0;
    if (gfirst && !GROUPS) {
      g0 = gfirst[t];
      nb = gfirst[t + 1] - g0;
    }
    const bool use_idx = !GROUPS && gfirst && (uint64_t)m >= (uint64_t)nb * min_bucket;
    CandIndex ix;
    ix.head = head;
    int order = use_idx ? kOrderBucket : kOrderScan;
    if (GROUPS) {
      order = ORD;
      ix.gsum = gsum + goff[t];
      ix.ssum = ssum + 8u * goff[t];
      ix.gcode = gcode + e0;
      ix.gidx = glocal + e0;
    } else if (use_idx) {
      ix.bmk = bmk + g0;
      ix.bstart = bstart + g0;
      ix.bcnt = bcnt + g0;
      ix.nbk = nb;
      ix.bidx = blocal;
      ix.bcode = bcode;
    }
    member_body<D, P, NT>(erows + (size_t)e0 * D, codes + e0, pos + e0, m, (int)maxd[t],
                          minkey + (size_t)t * 16, dom, (task - toff[t]) * P,
                          cover ? cover : auto_cover(m), order, ix);
  }
}

}  // namespace
}  // namespace sky
