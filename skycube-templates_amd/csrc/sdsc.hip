// sdsc.hip -- the SDSC lattice walk on the device (sky_sdsc, sky_sdsc_fetch).
//
// The reference's SDSC<DIMS>::Execute (src/sdsc/SDSC.cu:248-335) computes
// the full space over every point, then walks the lattice top-down, one
// level at a time with a barrier per level (levels above max_d skipped);
// cuboid s is a SkyAlign over the sky and ext lists of its parent with the
// fewest, or of the full space when its first parent (the largest id) has
// an empty skyline (SDSC.cu:81-104, findMinParent :365-376).
//
// Here a whole level is one batched pass over all its cuboids: each cuboid
// is an instance of the batched split (stage_batch.hip, sdsc_chunk_d) whose
// candidates are its parent's list, gathered on the device and projected
// onto the cuboid's columns.  Every cuboid's result stays on the device as
// one ascending id list (sky and ext together) plus a sky flag per id, in a
// pool that grows level by level; a level's parent choice reads the pool's
// per-cuboid counts on the device.  The host synchronises a few times per
// level (to size the level and each chunk), never per cuboid.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch.cuh"
#include "internal.hpp"
#include "prims.cuh"

using sky::kBlock;

namespace sky {
namespace {

// Parent of each cuboid of a level: src[t] and its list length cnt[t].
// Parents s | 1 << j are visited in descending id order (the order of the
// reference lattice's parents lists); index 0 of the lattice arrays is the
// list of all points, the full space's parent.
__global__ void k_sdsc_parent(const uint32_t* __restrict__ subs, uint32_t T, uint32_t d,
                              const uint32_t* __restrict__ lat_cnt,
                              const uint32_t* __restrict__ lat_sky, uint32_t* __restrict__ src,
                              uint32_t* __restrict__ cnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const uint32_t s = subs[t], full = (1u << d) - 1u;
  uint32_t best = 0;
  if (s != full) {
    uint32_t first = 0;
    for (int j = (int)d - 1; j >= 0; --j)
      if (!((s >> j) & 1u)) {
        first = s | (1u << j);
        break;
      }
    if (lat_sky[first] == 0) {
      best = full;
    } else {
      best = first;
      for (int j = (int)d - 1; j >= 0; --j) {
        if ((s >> j) & 1u) continue;
        const uint32_t p = s | (1u << j);
        if (lat_cnt[p] < lat_cnt[best]) best = p;
      }
    }
  }
  src[t] = best;
  cnt[t] = lat_cnt[best];
}

// Candidates of a chunk: cuboid t's are its parent's list.
__global__ void k_sdsc_cand(const uint32_t* __restrict__ pool, const uint64_t* __restrict__ lat_off,
                            const uint32_t* __restrict__ src, const uint32_t* __restrict__ off,
                            uint32_t T, uint32_t N, uint32_t* __restrict__ cand) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t t = seg_of(off, T, i);
  cand[i] = pool[lat_off[src[t]] + (i - off[t])];
}

// Kept candidates (sky or ext) appended to the pool with their sky flags.
__global__ void k_sdsc_append(const uint32_t* __restrict__ sel, const uint32_t* __restrict__ count,
                              const uint32_t* __restrict__ cand, const uint8_t* __restrict__ sky,
                              uint32_t* __restrict__ pool, uint8_t* __restrict__ flag) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= *count) return;
  const uint32_t i = sel[j];
  pool[j] = cand[i];
  flag[j] = sky[i];
}

// Per cuboid of the chunk: where its list starts in the pool, its length
// and its skyline size.
__global__ void k_sdsc_record(const uint32_t* __restrict__ subs, uint32_t T, uint64_t base,
                              const uint32_t* __restrict__ koff, const uint32_t* __restrict__ kcnt,
                              const uint32_t* __restrict__ scnt, uint64_t* __restrict__ lat_off,
                              uint32_t* __restrict__ lat_cnt, uint32_t* __restrict__ lat_sky) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const uint32_t s = subs[t];
  lat_off[s] = base + koff[t];
  lat_cnt[s] = kcnt[t];
  lat_sky[s] = scnt[t];
}

}  // namespace
}  // namespace sky

namespace skyi {
namespace {

// Candidates per chunk of a level (the batched split's buffers scale with it).
constexpr uint64_t kSdscChunk = 1ull << 25;

// Grows b to `bytes`, keeping its first `used` bytes.
int grow_keep(sky_ctx* c, DevBuf& b, size_t bytes, size_t used) {
  if (b.cap >= bytes) return SKY_OK;
  DevBuf nb;
  TRY(ensure(c, nb, bytes + bytes / 2));
  if (used) HIPCHK(hipMemcpyAsync(nb.p, b.p, used, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_buf(c, b);
  b = nb;
  return SKY_OK;
}

int read_u32(sky_ctx* c, const uint32_t* dev, uint32_t* host) {
  HIPCHK(hipMemcpyAsync(host, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

// One chunk: cuboids subs[t0 .. t1) of level L with host offsets hoff.
int sdsc_chunk(sky_ctx* c, const float* rows_dev, uint32_t d, uint32_t L, uint32_t t0,
               uint32_t t1, const std::vector<uint64_t>& hoff, uint64_t* pool_used) {
  const uint32_t T = t1 - t0;
  const uint64_t N64 = hoff[t1] - hoff[t0];
  const uint32_t N = (uint32_t)N64;
  std::vector<uint32_t> off(T + 1);
  std::vector<uint2> tasks;
  for (uint32_t k = 0; k <= T; ++k) off[k] = (uint32_t)(hoff[t0 + k] - hoff[t0]);
  for (uint32_t k = 0; k < T; ++k)
    for (uint32_t s0 = off[k]; s0 < off[k + 1]; s0 += kBlock) tasks.push_back(make_uint2(k, s0));
  TRY(ensure(c, c->sd_coff, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, c->sd_tasks, sizeof(uint2) * std::max<size_t>(tasks.size(), 1)));
  TRY(ensure(c, c->sd_cand, sizeof(uint32_t) * std::max<uint32_t>(N, 1)));
  TRY(ensure(c, c->sd_keep, std::max<uint32_t>(N, 1)));
  TRY(ensure(c, c->sd_skyf, std::max<uint32_t>(N, 1)));
  TRY(ensure(c, c->sd_sel, sizeof(uint32_t) * std::max<uint32_t>(N, 1)));
  TRY(ensure(c, c->sd_kcnt, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, c->sd_koff, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, c->sd_scnt, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, c->sd_soff, sizeof(uint32_t) * (T + 1)));
  TRY(batch_buffers(c, N, T, batch_self_index((int)L, T) ? ((uint64_t)T << L) + 1 : 1));
  HIPCHK(hipMemcpyAsync(c->sd_coff.p, off.data(), sizeof(uint32_t) * (T + 1),
                        hipMemcpyHostToDevice, c->stream));
  if (!tasks.empty())
    HIPCHK(hipMemcpyAsync(c->sd_tasks.p, tasks.data(), sizeof(uint2) * tasks.size(),
                          hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(sky::k_sdsc_cand, dim3(grid_for(N)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->sd_pool), P<uint64_t>(c->sd_lat_off),
                     P<uint32_t>(c->sd_src) + t0, P<uint32_t>(c->sd_coff), T, N,
                     P<uint32_t>(c->sd_cand));
  LAUNCHCHK();
  SdscChunk k;
  k.rows = rows_dev;
  k.d = d;
  k.subs = P<uint32_t>(c->sd_subs) + t0;
  k.cand = P<uint32_t>(c->sd_cand);
  k.off = P<uint32_t>(c->sd_coff);
  k.tasks = P<uint2>(c->sd_tasks);
  k.ntasks = (uint32_t)tasks.size();
  k.T = T;
  k.N = N;
  for (uint32_t q = 0; q < T; ++q) k.max_n = std::max(k.max_n, off[q + 1] - off[q]);
  k.keep = P<uint8_t>(c->sd_keep);
  k.sky = P<uint8_t>(c->sd_skyf);
  k.cnt = P<uint32_t>(c->sd_counters);
  int rc = SKY_ERR_ARG;
  SKY_DISPATCH_D(L, rc = sdsc_chunk_d<D>(c, k));
  TRY(rc);
  // the kept candidates, in candidate order (ascending ids per cuboid)
  uint32_t* kept_dev = P<uint32_t>(c->sd_counters) + 8;
  TRY(select_dev(c, k.keep, N, P<uint32_t>(c->sd_sel), kept_dev));
  TRY(seg_count_scan(c, k.keep, k.off, T, P<uint32_t>(c->sd_kcnt), P<uint32_t>(c->sd_koff),
                     k.max_n));
  TRY(seg_count_scan(c, k.sky, k.off, T, P<uint32_t>(c->sd_scnt), P<uint32_t>(c->sd_soff),
                     k.max_n));
  uint32_t kept = 0;
  TRY(read_u32(c, kept_dev, &kept));
  uint32_t err = 0;
  TRY(read_u32(c, k.cnt, &err));
  if (err & 2u) return fail(c, SKY_ERR_NAN, "input holds NaN");
  TRY(grow_keep(c, c->sd_pool, sizeof(uint32_t) * (*pool_used + kept), sizeof(uint32_t) * *pool_used));
  TRY(grow_keep(c, c->sd_flag, *pool_used + kept, *pool_used));
  hipLaunchKernelGGL(sky::k_sdsc_append, dim3(grid_for(std::max<uint32_t>(kept, 1))), dim3(kBlock), 0,
                     c->stream, P<uint32_t>(c->sd_sel), kept_dev, P<uint32_t>(c->sd_cand),
                     P<uint8_t>(c->sd_skyf), P<uint32_t>(c->sd_pool) + *pool_used,
                     P<uint8_t>(c->sd_flag) + *pool_used);
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_sdsc_record, dim3(grid_for(T)), dim3(kBlock), 0, c->stream,
                     k.subs, T, *pool_used, P<uint32_t>(c->sd_koff), P<uint32_t>(c->sd_kcnt),
                     P<uint32_t>(c->sd_scnt), P<uint64_t>(c->sd_lat_off),
                     P<uint32_t>(c->sd_lat_cnt), P<uint32_t>(c->sd_lat_sky));
  LAUNCHCHK();
  *pool_used += kept;
  return SKY_OK;
}

// One level: every cuboid of popcount L (the full space when L = d).
int sdsc_level(sky_ctx* c, const float* rows_dev, uint32_t d, uint32_t L, uint64_t* pool_used) {
  std::vector<uint32_t> subs;
  for (uint32_t s = 1; s < (1u << d); ++s)
    if ((uint32_t)__builtin_popcount(s) == L) subs.push_back(s);
  const uint32_t T = (uint32_t)subs.size();
  TRY(ensure(c, c->sd_subs, sizeof(uint32_t) * T));
  TRY(ensure(c, c->sd_src, sizeof(uint32_t) * T));
  TRY(ensure(c, c->sd_pcnt, sizeof(uint32_t) * T));
  HIPCHK(hipMemcpyAsync(c->sd_subs.p, subs.data(), sizeof(uint32_t) * T, hipMemcpyHostToDevice,
                        c->stream));
  hipLaunchKernelGGL(sky::k_sdsc_parent, dim3(grid_for(T)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->sd_subs), T, d, P<uint32_t>(c->sd_lat_cnt),
                     P<uint32_t>(c->sd_lat_sky), P<uint32_t>(c->sd_src), P<uint32_t>(c->sd_pcnt));
  LAUNCHCHK();
  std::vector<uint32_t> cnt(T);
  HIPCHK(hipMemcpyAsync(cnt.data(), c->sd_pcnt.p, sizeof(uint32_t) * T, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<uint64_t> hoff(T + 1, 0);
  for (uint32_t t = 0; t < T; ++t) hoff[t + 1] = hoff[t] + cnt[t];
  // chunks of whole cuboids, each within kSdscChunk candidates (a cuboid
  // larger than that is a chunk of its own)
  uint32_t t0 = 0;
  while (t0 < T) {
    uint32_t t1 = t0 + 1;
    while (t1 < T && hoff[t1 + 1] - hoff[t0] <= kSdscChunk) ++t1;
    if (hoff[t1] - hoff[t0] >= (1ull << 32))
      return fail(c, SKY_ERR_CAPACITY, "a cuboid of more than 2^32 candidates");
    if (hoff[t1] > hoff[t0]) TRY(sdsc_chunk(c, rows_dev, d, L, t0, t1, hoff, pool_used));
    t0 = t1;
  }
  return SKY_OK;
}

}  // namespace

// The lattice of n x d rows on the device: pool lists and per-cuboid
// offsets / counts / skyline sizes in c->sd_*.
int sdsc_walk(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d) {
  const uint32_t S = 1u << d;
  c->has_sdsc = false;
  c->has_batch = false;  // the batched buffers are reused here
  c->batch.ready = false;
  c->batch.groups.clear();
  if (c->batch.exec) (void)hipGraphExecDestroy(c->batch.exec);
  if (c->batch.graph) (void)hipGraphDestroy(c->batch.graph);
  c->batch.exec = nullptr;
  c->batch.graph = nullptr;
  TRY(ensure(c, c->sd_lat_off, sizeof(uint64_t) * S));
  TRY(ensure(c, c->sd_lat_cnt, sizeof(uint32_t) * S));
  TRY(ensure(c, c->sd_lat_sky, sizeof(uint32_t) * S));
  TRY(ensure(c, c->sd_counters, sizeof(uint32_t) * 16));
  TRY(fill32(c, c->sd_lat_off.p, 2ull * S, 0u));
  TRY(fill32(c, c->sd_lat_cnt.p, S, 0u));
  TRY(fill32(c, c->sd_lat_sky.p, S, 0u));
  // entry 0: every point, the candidates of the full space
  TRY(grow_keep(c, c->sd_pool, sizeof(uint32_t) * std::max<uint32_t>(n, 1), 0));
  TRY(grow_keep(c, c->sd_flag, std::max<uint32_t>(n, 1), 0));
  if (n) {
    hipLaunchKernelGGL(sky::k_iota, dim3(grid_for(n)), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(c->sd_pool), n);
    LAUNCHCHK();
    TRY(fill32(c, c->sd_lat_cnt.p, 1, n));
    TRY(fill32(c, c->sd_lat_sky.p, 1, n));
  }
  uint64_t used = n;
  if (n) {
    TRY(sdsc_level(c, rows_dev, d, d, &used));
    for (uint32_t L = d - 1; L >= 1; --L)
      if (L <= max_d) TRY(sdsc_level(c, rows_dev, d, L, &used));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->sd_used = used;
  c->sd_d = d;
  c->has_sdsc = true;
  return SKY_OK;
}

// The last lattice as CSR arrays in cuboid order (see sky_sdsc).
int sdsc_copy_out(sky_ctx* c, uint64_t* sky_off, uint32_t* sky_ids, uint64_t sky_cap,
                  uint64_t* ext_off, uint32_t* ext_ids, uint64_t ext_cap) {
  const uint32_t d = c->sd_d, S = 1u << d;
  std::vector<uint64_t> loff(S);
  std::vector<uint32_t> lcnt(S), lsky(S);
  HIPCHK(hipMemcpyAsync(loff.data(), c->sd_lat_off.p, sizeof(uint64_t) * S, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(lcnt.data(), c->sd_lat_cnt.p, sizeof(uint32_t) * S, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(lsky.data(), c->sd_lat_sky.p, sizeof(uint32_t) * S, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  lcnt[0] = lsky[0] = 0;  // entry 0 is the list of all points, not a cuboid
  uint64_t os = 0, oe = 0;
  for (uint32_t s = 0; s < S; ++s) {
    sky_off[s] = os;
    ext_off[s] = oe;
    os += lsky[s];
    oe += lcnt[s] - lsky[s];
  }
  sky_off[S] = os;
  ext_off[S] = oe;
  if (!sky_ids || !ext_ids || os > sky_cap || oe > ext_cap)
    return fail(c, SKY_ERR_CAPACITY, "sdsc needs %llu sky ids and %llu ext ids",
                (unsigned long long)os, (unsigned long long)oe);
  const uint64_t used = c->sd_used;
  std::vector<uint32_t> ids(used);
  std::vector<uint8_t> flag(used);
  if (used) {
    HIPCHK(hipMemcpyAsync(ids.data(), c->sd_pool.p, sizeof(uint32_t) * used,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(flag.data(), c->sd_flag.p, used, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  for (uint32_t s = 1; s < S; ++s) {
    uint64_t ps = sky_off[s], pe = ext_off[s];
    for (uint64_t e = loff[s]; e < loff[s] + lcnt[s]; ++e) {
      if (flag[e]) sky_ids[ps++] = ids[e];
      else ext_ids[pe++] = ids[e];
    }
  }
  return SKY_OK;
}

}  // namespace skyi
