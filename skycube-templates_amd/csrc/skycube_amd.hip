// skycube_amd.hip -- C ABI of the MI355X skycube engine (see
// include/skycube_amd.h) and the host-side orchestration of its kernels.
//
// One context = one device + one stream + the device buffers of its calls,
// grown on demand within an HBM budget.  Every HIP return code is checked
// and turned into a SKY_ERR_* status with a message.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/skycube_amd.h"
#include "batch.cuh"
#include "kernels.cuh"

using sky::DimList;
using sky::kBlock;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// One dimensionality class of a batch: its instances and device state.
struct BatchGroup {
  int D = 0;
  std::vector<uint32_t> inst;  // caller's instance ids, in group order
  std::vector<uint32_t> off;   // point offsets of the instances (T + 1)
  uint32_t T = 0, N = 0, ntasks = 0;
  DevBuf d_off, d_src_foff, d_tasks, d_maxd;  // descriptors
  DevBuf ext_pids, eoff, dom, cnt;            // results; cnt: counters
  mutable std::vector<uint32_t> eo_host;       // eoff read back by the last fetch
};

struct BatchState {
  std::vector<uint32_t> sig_n, sig_d, sig_m;
  const float* sig_rows = nullptr;
  BatchGroup pool[17];              // per dimensionality, buffers kept across batches
  std::vector<BatchGroup*> groups;  // the groups of the current batch, ascending d
  uint32_t n_inst = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ready = false;
  uint64_t sig_tune = 0;
  // shared scratch, sized for the largest group
  DevBuf rows, key_a, key_b, val_a, val_b, srows, skeys, spid, alive, s1idx, s1rows, s1keys,
      s1pid, masks, status, flag_pt, flag_s1, eidx, erows, epid, pos, codes, rank, piv, minkey,
      seg_a, seg_b, tmp;
  // mask index: sort keys / values, bucket order, bucket list, bucket table
  DevBuf ikey, ikey2, ival, isort, ibkey, ibcode, iblocal, ihead, ibmk, ibstart, ibcnt, ibrun,
      igfirst, iboff;
  // code-order groups: per-instance group counts / offsets, summaries
  DevBuf gcnt, goff, gsum;
  DevBuf sminkey;  // column minima of the survivors per instance
  size_t tmp_bytes = 0;
};

// Launch settings of one dimensionality (sky_ctx_set_option, autotuner).
struct Tuning {
  uint32_t cover = 0;  // candidates covered first; 0: auto_cover of each extended skyline
  uint32_t pruners = sky::kPruners;       // prefilter points
  uint32_t min_bucket = sky::kMinBucket;  // bucket-index threshold
  uint32_t grid = 2048;                   // workgroups of the persistent kernel
  uint32_t tile = 0;                      // 0: member_points, 1: half of it
  uint32_t order = 2;                     // batched membership: 2 code-order groups, 0 scan / buckets
  uint32_t head = 0;                      // cover head tested pair by pair (0: cover / 4)
};

}  // namespace

struct sky_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::string err;
  // device buffers
  DevBuf rows, key_a, key_b, pid_a, pid_b, srows, alive, s1idx, s1rows, s1keys, s1pid,
      status, flag, ext_pids, erows, dom, cub_tmp, counter, cand, eidx, epid, pos, codes, piv,
      masks, hc_cnt, hc_off, hc_key_a, hc_key_b, hc_val_a, hc_val_b, hc_ukey, hc_runs, hc_gw,
      hc_gv, hc_goff, bmask, biota, bidx, bkey, bcode, boff, bcnt, gsum, defer;
  // the last batch
  BatchState batch;
  // membership candidate order: -1 by bucket size, 0 scan, 1 bucket index
  int opt_member_index = -1;
  Tuning tune[17];       // per dimensionality
  uint64_t tune_ver = 0;  // bumped on every change (cached graphs are rebuilt)
  bool has_batch = false;
  // grouping of the last HashCube export
  bool has_hc = false;
  uint64_t hc_groups = 0, hc_entries = 0;
  // stage timing (HIP events on `stream`)
  bool timing = false;
  hipEvent_t ev[SKY_STAGE_COUNT + 1] = {};
  float stage_ms[SKY_STAGE_COUNT] = {};
  // last MDMC result
  bool has_mdmc = false;
  uint32_t res_n_ext = 0, res_d = 0, res_W = 0;
};

namespace {

int fail(sky_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(c, SKY_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,                  \
                  hipGetErrorString(e_), __FILE__, __LINE__);                         \
  } while (0)

#define LAUNCHCHK()                                                                   \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return fail(c, SKY_ERR_DEVICE, "kernel launch failed: %s (%s:%d)",              \
                  hipGetErrorString(e_), __FILE__, __LINE__);                         \
  } while (0)

#define TRY(expr)                 \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != SKY_OK) return rc_; \
  } while (0)

int ensure(sky_ctx* c, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return SKY_OK;
  size_t want = bytes + bytes / 8;  // headroom against regrowth
  if (b.p) {
    HIPCHK(hipFree(b.p));
    c->in_use -= b.cap;
    b.p = nullptr;
    b.cap = 0;
  }
  if (c->in_use + want > c->budget) {
    want = bytes;
    if (c->in_use + want > c->budget)
      return fail(c, SKY_ERR_OOM_BUDGET,
                  "device buffer of %zu bytes exceeds the HBM budget (%zu of %zu in use)",
                  bytes, c->in_use, c->budget);
  }
  HIPCHK(hipMalloc(&b.p, want));
  b.cap = want;
  c->in_use += want;
  return SKY_OK;
}

template <typename T>
T* P(DevBuf& b) {
  return static_cast<T*>(b.p);
}

// Records the event that opens stage `k` (stage k-1 ends there).
int mark(sky_ctx* c, int k) {
  if (!c->timing) return SKY_OK;
  if (!c->ev[k]) HIPCHK(hipEventCreate(&c->ev[k]));
  HIPCHK(hipEventRecord(c->ev[k], c->stream));
  return SKY_OK;
}

int read_marks(sky_ctx* c) {
  if (!c->timing) return SKY_OK;
  for (int k = 0; k < SKY_STAGE_COUNT; ++k) {
    c->stage_ms[k] = 0.0f;
    if (c->ev[k] && c->ev[k + 1]) HIPCHK(hipEventElapsedTime(&c->stage_ms[k], c->ev[k], c->ev[k + 1]));
  }
  return SKY_OK;
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

#define SKY_DISPATCH_D(DV, ...)                          \
  switch (DV) {                                          \
    case 1: { constexpr int D = 1; __VA_ARGS__; } break;   \
    case 2: { constexpr int D = 2; __VA_ARGS__; } break;   \
    case 3: { constexpr int D = 3; __VA_ARGS__; } break;   \
    case 4: { constexpr int D = 4; __VA_ARGS__; } break;   \
    case 5: { constexpr int D = 5; __VA_ARGS__; } break;   \
    case 6: { constexpr int D = 6; __VA_ARGS__; } break;   \
    case 7: { constexpr int D = 7; __VA_ARGS__; } break;   \
    case 8: { constexpr int D = 8; __VA_ARGS__; } break;   \
    case 9: { constexpr int D = 9; __VA_ARGS__; } break;   \
    case 10: { constexpr int D = 10; __VA_ARGS__; } break; \
    case 11: { constexpr int D = 11; __VA_ARGS__; } break; \
    case 12: { constexpr int D = 12; __VA_ARGS__; } break; \
    case 13: { constexpr int D = 13; __VA_ARGS__; } break; \
    case 14: { constexpr int D = 14; __VA_ARGS__; } break; \
    case 15: { constexpr int D = 15; __VA_ARGS__; } break; \
    case 16: { constexpr int D = 16; __VA_ARGS__; } break; \
    default: break;                                      \
  }

DimList dims_of(uint32_t sub, uint32_t* ds) {
  DimList L{};
  uint32_t k = 0;
  for (int j = 0; j < 16; ++j)
    if (sub >> j & 1u) L.idx[k++] = j;
  *ds = k;
  return L;
}

int copy_count(sky_ctx* c, uint32_t* host) {
  HIPCHK(hipMemcpyAsync(host, P<uint32_t>(c->counter), sizeof(uint32_t), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

// Indices i in [0, n) with flags[i] != 0, ascending, into out; count into
// the context counter.
int select_flagged(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out) {
  hipcub::CountingInputIterator<uint32_t> it(0);
  size_t bytes = 0;
  HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flags, out, P<uint32_t>(c->counter),
                                        n, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(hipcub::DeviceSelect::Flagged(c->cub_tmp.p, bytes, it, flags, out,
                                        P<uint32_t>(c->counter), n, c->stream));
  return SKY_OK;
}

// The sky/ext split of a candidate set over subspace `sub` (ds columns).
// Leaves, in score order, s1pid[0..m) and status[0..m) (0 sky, 1 ext, 2
// strictly dominated) of the prefilter survivors in the context; points
// removed by the prefilter are strictly dominated.  need_dom = 0 computes
// only the strict part (status 0 or 2).
// Buckets the m survivors by median mask (ds bits) for k_self_idx: bidx =
// survivor positions in (mask, score) order, boff = bucket bounds, bkey /
// bcode = their keys and codes in that order.
int self_index(sky_ctx* c, uint32_t ds, uint32_t m) {
  const uint32_t nb = 1u << ds;
  TRY(ensure(c, c->bmask, sizeof(uint32_t) * m));
  TRY(ensure(c, c->biota, sizeof(uint32_t) * m));
  TRY(ensure(c, c->bidx, sizeof(uint32_t) * m));
  TRY(ensure(c, c->bkey, sizeof(uint32_t) * m));
  TRY(ensure(c, c->bcode, sizeof(uint2) * m));
  TRY(ensure(c, c->boff, sizeof(uint32_t) * (nb + 1)));
  hipLaunchKernelGGL(sky::k_iota, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->biota), m);
  LAUNCHCHK();
  size_t bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, P<uint32_t>(c->masks),
                                            P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), m, 0, (int)ds, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, bytes, P<uint32_t>(c->masks),
                                            P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), m, 0, (int)ds, c->stream));
  hipLaunchKernelGGL(sky::k_bucket_bounds, dim3(grid_for(nb + 1)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bmask), m, nb, P<uint32_t>(c->boff));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bucket_gather, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bidx), P<uint32_t>(c->s1keys), P<uint2>(c->codes), m,
                     P<uint32_t>(c->bkey), P<uint2>(c->bcode));
  LAUNCHCHK();
  return SKY_OK;
}

int run_split(sky_ctx* c, const float* rows_dev, uint32_t n_rows, uint32_t d, uint32_t sub,
              const uint32_t* cand_dev, uint32_t n_cand, int need_dom, uint32_t* m_out) {
  uint32_t ds = 0;
  const DimList dims = dims_of(sub, &ds);
  const uint32_t nc = n_cand;
  *m_out = 0;
  if (nc == 0) return SKY_OK;
  TRY(ensure(c, c->key_a, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->key_b, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->pid_a, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->pid_b, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->srows, sizeof(float) * (size_t)nc * ds));
  TRY(ensure(c, c->alive, nc));
  TRY(ensure(c, c->s1idx, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->s1rows, sizeof(float) * (size_t)nc * ds));
  TRY(ensure(c, c->s1keys, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->s1pid, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->status, nc));
  TRY(ensure(c, c->masks, sizeof(uint32_t) * nc));
  TRY(ensure(c, c->codes, sizeof(uint2) * nc));
  TRY(ensure(c, c->piv, sizeof(uint32_t) * 16 * sky::kCodePivots));
  TRY(ensure(c, c->counter, 64 * sizeof(uint32_t)));

  TRY(mark(c, SKY_STAGE_SCORE_SORT));
  HIPCHK(hipMemsetAsync(c->counter.p, 0, 4 * sizeof(uint32_t), c->stream));
  uint32_t* err_dev = P<uint32_t>(c->counter) + 1;
  hipLaunchKernelGGL(sky::k_score, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, rows_dev,
                     n_rows, d, cand_dev, nc, dims, ds, P<uint32_t>(c->key_a),
                     P<uint32_t>(c->pid_a), err_dev);
  LAUNCHCHK();
  uint32_t errv = 0;
  HIPCHK(hipMemcpyAsync(&errv, err_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (errv & 1u) return fail(c, SKY_ERR_ARG, "candidate id out of range (n_rows=%u)", n_rows);
  if (errv & 2u) return fail(c, SKY_ERR_NAN, "input holds NaN");

  size_t bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, P<uint32_t>(c->key_a),
                                            P<uint32_t>(c->key_b), P<uint32_t>(c->pid_a),
                                            P<uint32_t>(c->pid_b), nc, 0, 32, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, bytes, P<uint32_t>(c->key_a),
                                            P<uint32_t>(c->key_b), P<uint32_t>(c->pid_a),
                                            P<uint32_t>(c->pid_b), nc, 0, 32, c->stream));

  const uint32_t npr = std::min<uint32_t>(nc, c->tune[ds].pruners);
  SKY_DISPATCH_D(ds, {
    hipLaunchKernelGGL(sky::k_gather<D>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream,
                       rows_dev, d, P<uint32_t>(c->pid_b), nc, dims, P<float>(c->srows));
    LAUNCHCHK();
    TRY(mark(c, SKY_STAGE_PRUNE));
    hipLaunchKernelGGL(sky::k_prune<D>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream,
                       P<float>(c->srows), P<uint32_t>(c->key_b), nc, npr, P<uint8_t>(c->alive));
    LAUNCHCHK();
  });
  TRY(select_flagged(c, P<uint8_t>(c->alive), nc, P<uint32_t>(c->s1idx)));
  uint32_t m = 0;
  TRY(copy_count(c, &m));
  if (m == 0) return fail(c, SKY_ERR_DEVICE, "prefilter removed every point");
  SKY_DISPATCH_D(ds, {
    hipLaunchKernelGGL(sky::k_gather_idx<D>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                       P<float>(c->srows), P<uint32_t>(c->key_b), P<uint32_t>(c->pid_b),
                       P<uint32_t>(c->s1idx), m, P<float>(c->s1rows), P<uint32_t>(c->s1keys),
                       P<uint32_t>(c->s1pid));
    LAUNCHCHK();
    TRY(mark(c, SKY_STAGE_SELF));
    hipLaunchKernelGGL(sky::k_pivots<D>, dim3(D * sky::kPivotSplit), dim3(kBlock), 0, c->stream,
                       P<float>(c->s1rows), m, P<uint32_t>(c->piv));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_codes<D>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                       P<float>(c->s1rows), m, P<uint32_t>(c->piv), P<uint2>(c->codes),
                       P<uint32_t>(c->masks));
    LAUNCHCHK();
  });
  TRY(self_index(c, ds, m));
  // column minima of the survivors (a point holding one has no strict dominator)
  uint32_t* sminkey = P<uint32_t>(c->counter) + 32;
  if (!need_dom) HIPCHK(hipMemsetAsync(sminkey, 0xFF, 16 * sizeof(uint32_t), c->stream));
  SKY_DISPATCH_D(ds, {
    if (!need_dom) {
      hipLaunchKernelGGL(sky::k_colmin<D>, dim3(std::min<unsigned>(grid_for(m), 1024u)),
                         dim3(kBlock), 0, c->stream, P<float>(c->s1rows), m, sminkey);
      LAUNCHCHK();
    }
    hipLaunchKernelGGL(sky::k_self_idx<D>, dim3((m + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(c->s1rows), P<uint32_t>(c->s1keys),
                       P<uint2>(c->codes), P<uint32_t>(c->boff), P<uint32_t>(c->bidx),
                       P<uint32_t>(c->bkey), P<uint2>(c->bcode), m, need_dom,
                       need_dom ? nullptr : sminkey, P<uint8_t>(c->status));
    LAUNCHCHK();
  });
  *m_out = m;
  return SKY_OK;
}

int check_ctx(sky_ctx* c) {
  if (!c) return SKY_ERR_ARG;
  c->err.clear();
  HIPCHK(hipSetDevice(c->device));
  return SKY_OK;
}

// Buckets the ne extended-skyline points (score order, median masks in
// c->masks) by median mask for k_member: bidx = their positions in (mask,
// score) order, bcode = their codes in that order, and the nbk non-empty
// buckets: mask bkey[g], positions boff[g] .. boff[g] + bcnt[g] - 1.
int member_index(sky_ctx* c, uint32_t d, uint32_t ne, uint32_t* nbk) {
  TRY(ensure(c, c->bmask, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->biota, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bidx, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bcode, sizeof(uint2) * ne));
  TRY(ensure(c, c->bkey, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->boff, sizeof(uint32_t) * (ne + 1)));
  TRY(ensure(c, c->bcnt, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->alive, ne));
  hipLaunchKernelGGL(sky::k_iota, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->biota), ne);
  LAUNCHCHK();
  size_t bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, P<uint32_t>(c->masks),
                                            P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), ne, 0, (int)d, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, bytes, P<uint32_t>(c->masks),
                                            P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), ne, 0, (int)d, c->stream));
  hipLaunchKernelGGL(sky::k_bucket_heads, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bmask), ne, P<uint8_t>(c->alive));
  LAUNCHCHK();
  // head positions into biota (free after the sort)
  TRY(select_flagged(c, P<uint8_t>(c->alive), ne, P<uint32_t>(c->biota)));
  TRY(copy_count(c, nbk));
  hipLaunchKernelGGL(sky::k_bucket_list, dim3(grid_for(*nbk)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bmask), P<uint32_t>(c->biota), *nbk, ne, P<uint32_t>(c->bkey),
                     P<uint32_t>(c->boff), P<uint32_t>(c->bcnt));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bucket_gather, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bidx), nullptr, P<uint2>(c->codes), ne, nullptr,
                     P<uint2>(c->bcode));
  LAUNCHCHK();
  return SKY_OK;
}

// Code order of the ne extended-skyline points for kOrderGroup: bidx =
// positions in order of the bit-plane code key, bcode = their codes, gsum =
// per-column minimum code of every group of 64.
template <int D>
int member_groups(sky_ctx* c, uint32_t ne) {
  const uint32_t ng = (ne + 63) / 64;
  TRY(ensure(c, c->hc_key_a, sizeof(uint64_t) * ne));
  TRY(ensure(c, c->hc_key_b, sizeof(uint64_t) * ne));
  TRY(ensure(c, c->biota, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bidx, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bcode, sizeof(uint2) * ne));
  TRY(ensure(c, c->gsum, sizeof(uint2) * ng));
  hipLaunchKernelGGL(sky::k_plane_keys<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint2>(c->codes), ne, P<uint64_t>(c->hc_key_a), P<uint32_t>(c->biota));
  LAUNCHCHK();
  size_t bytes = 0;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, P<uint64_t>(c->hc_key_a),
                                            P<uint64_t>(c->hc_key_b), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), ne, 0, 4 * D, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, bytes, P<uint64_t>(c->hc_key_a),
                                            P<uint64_t>(c->hc_key_b), P<uint32_t>(c->biota),
                                            P<uint32_t>(c->bidx), ne, 0, 4 * D, c->stream));
  hipLaunchKernelGGL(sky::k_group_sum<D>, dim3(grid_for(ng)), dim3(kBlock), 0, c->stream,
                     P<uint2>(c->codes), P<uint32_t>(c->bidx), ne, P<uint2>(c->bcode),
                     P<uint2>(c->gsum));
  LAUNCHCHK();
  return SKY_OK;
}

// The membership launch of the single-instance path.  Where the kernel runs
// with a narrower boundary table (d >= 15), the points it defers follow in a
// second launch with the full table.
template <int D, int PP, int NT, int O>
int launch_member(sky_ctx* c, uint32_t ne, uint32_t max_d, const uint32_t* minkey,
                  uint32_t cover, sky::CandIndex ix) {
  const dim3 grid((ne + PP - 1) / PP), blk(NT);
  constexpr int ZN = sky::narrow_zbits<D>();
  if constexpr (ZN < sky::zbits<D>()) {
    TRY(ensure(c, c->defer, sizeof(uint32_t) * ((size_t)ne + 1)));
    HIPCHK(hipMemsetAsync(c->defer.p, 0, sizeof(uint32_t), c->stream));
    ix.defer = P<uint32_t>(c->defer);
    hipLaunchKernelGGL((sky::k_member<D, PP, NT, O, ZN>), grid, blk, 0, c->stream,
                       P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                       (int)max_d, minkey, P<uint32_t>(c->dom), cover, ix);
    LAUNCHCHK();
    ix.defer = nullptr;
    // one point per workgroup: the deferred points are few, and their
    // workgroups are the tail of the pass
    const dim3 grid2(std::min<uint32_t>(ne, 512u));
    hipLaunchKernelGGL((sky::k_member_deferred<D, 1, NT, O>), grid2, blk, 0, c->stream,
                       P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                       (int)max_d, minkey, P<uint32_t>(c->dom), cover, P<uint32_t>(c->defer), ix);
    LAUNCHCHK();
  } else {
    hipLaunchKernelGGL((sky::k_member<D, PP, NT, O>), grid, blk, 0, c->stream,
                       P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                       (int)max_d, minkey, P<uint32_t>(c->dom), cover, ix);
    LAUNCHCHK();
  }
  return SKY_OK;
}

int run_mdmc(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d,
             uint32_t* n_ext) {
  c->has_mdmc = false;
  c->has_hc = false;
  const uint32_t W = sky_words_per_point(d);
  if (max_d == 0 || max_d > d) max_d = d;
  *n_ext = 0;
  if (n == 0) {
    c->has_mdmc = true;
    c->res_n_ext = 0;
    c->res_d = d;
    c->res_W = W;
    return SKY_OK;
  }
  uint32_t m = 0;
  TRY(run_split(c, rows_dev, n, d, (1u << d) - 1u, nullptr, n, 0, &m));
  TRY(mark(c, SKY_STAGE_COMPACT));
  TRY(ensure(c, c->flag, n));
  TRY(ensure(c, c->ext_pids, sizeof(uint32_t) * n));
  HIPCHK(hipMemsetAsync(c->flag.p, 0, n, c->stream));
  hipLaunchKernelGGL(sky::k_mark_ext, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<uint8_t>(c->status), P<uint32_t>(c->s1pid), m, P<uint8_t>(c->flag));
  LAUNCHCHK();
  TRY(select_flagged(c, P<uint8_t>(c->flag), n, P<uint32_t>(c->ext_pids)));
  uint32_t ne = 0;
  TRY(copy_count(c, &ne));
  if (ne >= (1u << 29))
    return fail(c, SKY_ERR_CAPACITY, "extended skyline of %u points (limit 2^29 - 1)", ne);
  TRY(ensure(c, c->erows, sizeof(float) * (size_t)ne * d));
  TRY(ensure(c, c->dom, sizeof(uint32_t) * (size_t)ne * W));
  TRY(ensure(c, c->eidx, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->epid, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->pos, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->codes, sizeof(uint2) * (size_t)ne));
  TRY(ensure(c, c->piv, sizeof(uint32_t) * 16 * sky::kCodePivots));
  uint32_t* minkey = P<uint32_t>(c->counter) + 16;
  HIPCHK(hipMemsetAsync(minkey, 0xFF, 16 * sizeof(uint32_t), c->stream));
  if (ne) {
    // the extended skyline in score order (candidates and points of the
    // membership kernel) and each point's output row in ascending-id order
    hipLaunchKernelGGL(sky::k_flag_ext, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                       P<uint8_t>(c->status), m, P<uint8_t>(c->alive));
    LAUNCHCHK();
    TRY(select_flagged(c, P<uint8_t>(c->alive), m, P<uint32_t>(c->eidx)));
    hipLaunchKernelGGL(sky::k_scatter_rank, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(c->ext_pids), ne, P<uint32_t>(c->key_a));
    LAUNCHCHK();
    SKY_DISPATCH_D(d, {
      hipLaunchKernelGGL(sky::k_gather_ext<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                         P<float>(c->s1rows), P<uint32_t>(c->s1pid), P<uint32_t>(c->eidx), ne,
                         P<float>(c->erows), P<uint32_t>(c->epid));
      LAUNCHCHK();
      hipLaunchKernelGGL(sky::k_pos, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                         P<uint32_t>(c->epid), ne, P<uint32_t>(c->key_a), P<uint32_t>(c->pos));
      LAUNCHCHK();
      hipLaunchKernelGGL(sky::k_colmin<D>, dim3(std::min<unsigned>(grid_for(ne), 1024u)),
                         dim3(kBlock), 0, c->stream, P<float>(c->erows), ne, minkey);
      LAUNCHCHK();
      hipLaunchKernelGGL(sky::k_pivots<D>, dim3(D * sky::kPivotSplit), dim3(kBlock), 0, c->stream,
                         P<float>(c->erows), ne, P<uint32_t>(c->piv));
      LAUNCHCHK();
      hipLaunchKernelGGL(sky::k_codes<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                         P<float>(c->erows), ne, P<uint32_t>(c->piv), P<uint2>(c->codes),
                         P<uint32_t>(c->masks));
      LAUNCHCHK();
      TRY(mark(c, SKY_STAGE_MEMBER));
      uint32_t nbk = 0;
      if (c->opt_member_index == sky::kOrderBucket) TRY(member_index(c, d, ne, &nbk));
      constexpr int PP = sky::member_points<D>();
      constexpr int NT = sky::member_threads<D>();
      const Tuning& tu = c->tune[d];
      int order = c->opt_member_index;
      if (order < 0) order = sky::kOrderGroup;
      const uint32_t cover = tu.cover ? tu.cover : sky::auto_cover(ne);
      sky::CandIndex ix;
      ix.head = tu.head;
      if (order == sky::kOrderBucket) {
        ix.bmk = P<uint32_t>(c->bkey);
        ix.bstart = P<uint32_t>(c->boff);
        ix.bcnt = P<uint32_t>(c->bcnt);
        ix.nbk = nbk;
        ix.bidx = P<uint32_t>(c->bidx);
        ix.bcode = P<uint2>(c->bcode);
      } else if (order == sky::kOrderGroup) {
        TRY(member_groups<D>(c, ne));
        ix.gsum = P<uint2>(c->gsum);
        ix.gcode = P<uint2>(c->bcode);
        ix.gidx = P<uint32_t>(c->bidx);
      }
      if (order == sky::kOrderGroup)
        TRY((launch_member<D, PP, NT, sky::kOrderGroup>(c, ne, max_d, minkey, cover, ix)));
      else if (order == sky::kOrderBucket)
        TRY((launch_member<D, PP, NT, sky::kOrderBucket>(c, ne, max_d, minkey, cover, ix)));
      else
        TRY((launch_member<D, PP, NT, sky::kOrderScan>(c, ne, max_d, minkey, cover, ix)));
    });
  } else {
    TRY(mark(c, SKY_STAGE_MEMBER));
  }
  TRY(mark(c, SKY_STAGE_COUNT));
  HIPCHK(hipStreamSynchronize(c->stream));
  TRY(read_marks(c));
  c->has_mdmc = true;
  c->res_n_ext = ne;
  c->res_d = d;
  c->res_W = W;
  *n_ext = ne;
  return SKY_OK;
}

int upload_rows(sky_ctx* c, const float* rows, uint32_t n, uint32_t d) {
  TRY(ensure(c, c->rows, sizeof(float) * (size_t)n * d));
  if (n)
    HIPCHK(hipMemcpyAsync(c->rows.p, rows, sizeof(float) * (size_t)n * d, hipMemcpyHostToDevice,
                          c->stream));
  return SKY_OK;
}

int check_dims(sky_ctx* c, uint32_t d) {
  if (d < 1 || d > 16) return fail(c, SKY_ERR_ARG, "d=%u outside [1,16]", d);
  return SKY_OK;
}

// sky(s) / ext(s) of a candidate set already on the device, as ascending
// row-id lists on the host.
int skyalign_dev(sky_ctx* c, const float* rows_dev, uint32_t n_rows, uint32_t d, uint32_t sub,
                 const uint32_t* cand_dev, uint32_t n_cand, std::vector<uint32_t>& sky,
                 std::vector<uint32_t>& ext) {
  sky.clear();
  ext.clear();
  uint32_t m = 0;
  TRY(run_split(c, rows_dev, n_rows, d, sub, cand_dev, n_cand, 1, &m));
  if (m == 0) return SKY_OK;
  std::vector<uint8_t> st(m);
  std::vector<uint32_t> pid(m);
  HIPCHK(hipMemcpyAsync(st.data(), c->status.p, m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(pid.data(), c->s1pid.p, sizeof(uint32_t) * m, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (uint32_t j = 0; j < m; ++j) {
    if (st[j] == 0) sky.push_back(pid[j]);
    else if (st[j] == 1) ext.push_back(pid[j]);
  }
  std::sort(sky.begin(), sky.end());
  std::sort(ext.begin(), ext.end());
  return SKY_OK;
}

// Groups the last MDMC bitmaps into HashCube entries on the device.
int build_hashcube(sky_ctx* c) {
  const uint32_t ne = c->res_n_ext, W = c->res_W;
  const uint32_t extra = c->res_d >= 5 ? 1u : 0u;
  c->hc_groups = c->hc_entries = 0;
  if (ne == 0) {
    c->has_hc = true;
    return SKY_OK;
  }
  TRY(ensure(c, c->hc_cnt, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->hc_off, sizeof(uint64_t) * (ne + 1)));
  hipLaunchKernelGGL(sky::k_hc_count, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->dom), ne, W, extra, P<uint32_t>(c->hc_cnt));
  LAUNCHCHK();
  // 64-bit exclusive scan of the counts (ne + 1 entries: the last is the total)
  TRY(ensure(c, c->hc_runs, sizeof(uint64_t) * (ne + 1)));
  uint64_t* cnt64 = P<uint64_t>(c->hc_runs);
  HIPCHK(hipMemsetAsync(cnt64 + ne, 0, sizeof(uint64_t), c->stream));
  {
    hipcub::TransformInputIterator<uint64_t, hipcub::CastOp<uint64_t>, const uint32_t*> it(
        P<uint32_t>(c->hc_cnt), hipcub::CastOp<uint64_t>());
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, P<uint64_t>(c->hc_off), ne,
                                            c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, bytes, it, P<uint64_t>(c->hc_off), ne,
                                            c->stream));
  }
  uint64_t last_off = 0;
  uint32_t last_cnt = 0;
  HIPCHK(hipMemcpyAsync(&last_off, P<uint64_t>(c->hc_off) + (ne - 1), sizeof(uint64_t),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&last_cnt, P<uint32_t>(c->hc_cnt) + (ne - 1), sizeof(uint32_t),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint64_t total = last_off + last_cnt;
  if (total > 0xFFFFFFFFull)
    return fail(c, SKY_ERR_CAPACITY, "HashCube export of %llu entries exceeds 2^32",
                (unsigned long long)total);
  TRY(ensure(c, c->hc_key_a, sizeof(uint64_t) * total));
  TRY(ensure(c, c->hc_key_b, sizeof(uint64_t) * total));
  TRY(ensure(c, c->hc_val_a, sizeof(uint32_t) * total));
  TRY(ensure(c, c->hc_val_b, sizeof(uint32_t) * total));
  hipLaunchKernelGGL(sky::k_hc_emit, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->dom), P<uint32_t>(c->ext_pids), ne, W, extra,
                     P<uint64_t>(c->hc_off), P<uint64_t>(c->hc_key_a), P<uint32_t>(c->hc_val_a));
  LAUNCHCHK();
  int end_bit = 32;
  while ((1ull << (end_bit - 32)) <= (uint64_t)W) ++end_bit;  // word index bits
  {
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, P<uint64_t>(c->hc_key_a),
                                              P<uint64_t>(c->hc_key_b), P<uint32_t>(c->hc_val_a),
                                              P<uint32_t>(c->hc_val_b), (uint32_t)total, 0,
                                              end_bit, c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(c->cub_tmp.p, bytes, P<uint64_t>(c->hc_key_a),
                                              P<uint64_t>(c->hc_key_b), P<uint32_t>(c->hc_val_a),
                                              P<uint32_t>(c->hc_val_b), (uint32_t)total, 0,
                                              end_bit, c->stream));
  }
  // runs of equal keys = groups
  TRY(ensure(c, c->hc_ukey, sizeof(uint64_t) * total));
  TRY(ensure(c, c->hc_gw, sizeof(uint32_t) * total));
  TRY(ensure(c, c->hc_goff, sizeof(uint64_t) * (total + 1)));
  uint32_t* run_len = P<uint32_t>(c->hc_val_a);  // reuse: the unsorted values are done
  {
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceRunLengthEncode::Encode(nullptr, bytes, P<uint64_t>(c->hc_key_b),
                                                 P<uint64_t>(c->hc_ukey), run_len,
                                                 P<uint32_t>(c->counter), (uint32_t)total,
                                                 c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(hipcub::DeviceRunLengthEncode::Encode(c->cub_tmp.p, bytes, P<uint64_t>(c->hc_key_b),
                                                 P<uint64_t>(c->hc_ukey), run_len,
                                                 P<uint32_t>(c->counter), (uint32_t)total,
                                                 c->stream));
  }
  uint32_t ng = 0;
  TRY(copy_count(c, &ng));
  TRY(ensure(c, c->hc_gv, sizeof(uint32_t) * (ng + 1)));
  {
    hipcub::TransformInputIterator<uint64_t, hipcub::CastOp<uint64_t>, const uint32_t*> it(
        run_len, hipcub::CastOp<uint64_t>());
    size_t bytes = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, it, P<uint64_t>(c->hc_goff), ng,
                                            c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp.p, bytes, it, P<uint64_t>(c->hc_goff), ng,
                                            c->stream));
  }
  HIPCHK(hipMemcpyAsync(P<uint64_t>(c->hc_goff) + ng, &total, sizeof(uint64_t),
                        hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(sky::k_hc_groups, dim3(grid_for(ng)), dim3(kBlock), 0, c->stream,
                     P<uint64_t>(c->hc_ukey), (uint64_t)ng, P<uint32_t>(c->hc_gw),
                     P<uint32_t>(c->hc_gv));
  LAUNCHCHK();
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hc_groups = ng;
  c->hc_entries = total;
  c->has_hc = true;
  return SKY_OK;
}

void free_buf(sky_ctx* c, DevBuf& b) {
  if (b.p) {
    (void)hipFree(b.p);
    c->in_use -= b.cap;
  }
  b.p = nullptr;
  b.cap = 0;
}

// Forgets the current batch (graph included); the buffers stay for the next.
void free_batch(sky_ctx* c) {
  BatchState& B = c->batch;
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
  B.groups.clear();
  B.ready = false;
}

// Frees every device buffer of the batch path.
void free_batch_buffers(sky_ctx* c) {
  free_batch(c);
  BatchState& B = c->batch;
  for (BatchGroup& g : B.pool) {
    DevBuf* gb[] = {&g.d_off, &g.d_src_foff, &g.d_tasks, &g.d_maxd, &g.ext_pids, &g.eoff, &g.dom,
                    &g.cnt};
    for (DevBuf* b : gb) free_buf(c, *b);
  }
  DevBuf* sb[] = {&B.rows,  &B.key_a,  &B.key_b,  &B.val_a, &B.val_b,  &B.srows,  &B.skeys,
                  &B.spid,  &B.alive,  &B.s1idx,  &B.s1rows, &B.s1keys, &B.s1pid, &B.masks,
                  &B.status, &B.flag_pt, &B.flag_s1, &B.eidx, &B.erows, &B.epid,  &B.pos,
                  &B.codes, &B.rank,   &B.piv,    &B.minkey, &B.seg_a, &B.seg_b, &B.tmp,
                  &B.ikey,  &B.ikey2,  &B.ival,   &B.isort,  &B.ibkey, &B.ibcode, &B.iblocal,
                  &B.ihead, &B.ibmk,   &B.ibstart, &B.ibcnt, &B.ibrun, &B.igfirst, &B.iboff,
                  &B.gcnt,  &B.goff,   &B.gsum,   &B.sminkey};
  for (DevBuf* b : sb) free_buf(c, *b);
  B.tmp_bytes = 0;
}

constexpr int kGroupCounters = 8;  // err, survivors, ext (id order), ext (score order), work

int fill32(sky_ctx* c, void* p, uint64_t n, uint32_t v) {
  if (!n) return SKY_OK;
  hipLaunchKernelGGL(sky::k_fill_u32, dim3((unsigned)std::min<uint64_t>(grid_for(n), 4096)),
                     dim3(kBlock), 0, c->stream, static_cast<uint32_t*>(p), n, v);
  LAUNCHCHK();
  return SKY_OK;
}

int fill8(sky_ctx* c, void* p, uint64_t n, uint8_t v) {
  if (!n) return SKY_OK;
  hipLaunchKernelGGL(sky::k_fill_u8, dim3((unsigned)std::min<uint64_t>(grid_for(n), 4096)),
                     dim3(kBlock), 0, c->stream, static_cast<uint8_t*>(p), n, v);
  LAUNCHCHK();
  return SKY_OK;
}

// The batched survivor pass uses the mask index when its bucket table (one
// entry per instance and median mask) stays within 2^26 entries.
bool batch_self_index(int D, uint32_t T) { return ((uint64_t)T << D) <= (1ull << 26); }

// Temp bytes of every rocPRIM call of a group of N points, T instances.
int batch_tmp_bytes(sky_ctx* c, uint32_t N, uint32_t T, size_t* out) {
  size_t mx = 0, b = 0;
  hipcub::CountingInputIterator<uint32_t> it(0);
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                            (uint32_t*)nullptr, (uint32_t*)nullptr, N, 0, 64,
                                            c->stream));
  mx = std::max(mx, b);
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                            (uint32_t*)nullptr, (uint32_t*)nullptr, N, 0, 32,
                                            c->stream));
  mx = std::max(mx, b);
  HIPCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(
      nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, N,
      T, (const uint32_t*)nullptr, (const uint32_t*)nullptr, 0, 64, c->stream));
  mx = std::max(mx, b);
  HIPCHK(hipcub::DeviceSelect::Flagged(nullptr, b, it, (uint8_t*)nullptr, (uint32_t*)nullptr,
                                        (uint32_t*)nullptr, N, c->stream));
  mx = std::max(mx, b);
  hipcub::TransformInputIterator<uint32_t, hipcub::CastOp<uint32_t>, const uint8_t*> cast(
      nullptr, hipcub::CastOp<uint32_t>());
  HIPCHK(hipcub::DeviceSegmentedReduce::Sum(nullptr, b, cast, (uint32_t*)nullptr, T,
                                            (const uint32_t*)nullptr, (const uint32_t*)nullptr,
                                            c->stream));
  mx = std::max(mx, b);
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                          T + 1, c->stream));
  mx = std::max(mx, b);
  *out = mx;
  return SKY_OK;
}

// Builds the groups of a batch and every device buffer it needs.
int batch_setup(sky_ctx* c, const float* rows_dev, uint32_t n_inst, const uint32_t* n,
                const uint32_t* d, const uint32_t* max_d) {
  free_batch(c);
  BatchState& B = c->batch;
  B.n_inst = n_inst;
  std::vector<uint64_t> foff(n_inst + 1, 0);
  for (uint32_t t = 0; t < n_inst; ++t) foff[t + 1] = foff[t] + (uint64_t)n[t] * d[t];
  uint32_t maxN = 0, maxT = 0;
  for (int D = 1; D <= 16; ++D) {
    BatchGroup& g = B.pool[D];
    g.D = D;
    g.inst.clear();
    for (uint32_t t = 0; t < n_inst; ++t)
      if ((int)d[t] == D && n[t] > 0) g.inst.push_back(t);
    if (g.inst.empty()) continue;
    g.T = (uint32_t)g.inst.size();
    g.off.assign(g.T + 1, 0);
    std::vector<uint64_t> sf(g.T);
    std::vector<uint32_t> md(g.T);
    std::vector<uint2> tasks;
    for (uint32_t k = 0; k < g.T; ++k) {
      const uint32_t t = g.inst[k];
      if ((uint64_t)g.off[k] + n[t] > 0xFFFFFFFFull)
        return fail(c, SKY_ERR_ARG, "batch group of d=%d exceeds 2^32 points", D);
      if (n[t] >= (1u << 29))
        return fail(c, SKY_ERR_CAPACITY, "instance %u: %u points (limit 2^29 - 1)", t, n[t]);
      g.off[k + 1] = g.off[k] + n[t];
      sf[k] = foff[t];
      md[k] = (max_d && max_d[t] && max_d[t] < (uint32_t)D) ? max_d[t] : (uint32_t)D;
      for (uint32_t s0 = 0; s0 < n[t]; s0 += kBlock) tasks.push_back(make_uint2(k, g.off[k] + s0));
    }
    g.N = g.off[g.T];
    g.ntasks = (uint32_t)tasks.size();
    TRY(ensure(c, g.d_off, sizeof(uint32_t) * (g.T + 1)));
    TRY(ensure(c, g.d_src_foff, sizeof(uint64_t) * g.T));
    TRY(ensure(c, g.d_tasks, sizeof(uint2) * tasks.size()));
    TRY(ensure(c, g.d_maxd, sizeof(uint32_t) * g.T));
    TRY(ensure(c, g.ext_pids, sizeof(uint32_t) * g.N));
    TRY(ensure(c, g.eoff, sizeof(uint32_t) * (g.T + 1)));
    TRY(ensure(c, g.dom, sizeof(uint32_t) * (size_t)g.N * sky_words_per_point(D)));
    TRY(ensure(c, g.cnt, sizeof(uint32_t) * kGroupCounters));
    HIPCHK(hipMemcpyAsync(g.d_off.p, g.off.data(), sizeof(uint32_t) * (g.T + 1),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(g.d_src_foff.p, sf.data(), sizeof(uint64_t) * g.T,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(g.d_tasks.p, tasks.data(), sizeof(uint2) * tasks.size(),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(g.d_maxd.p, md.data(), sizeof(uint32_t) * g.T, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));  // host vectors go out of scope
    maxN = std::max(maxN, g.N);
    maxT = std::max(maxT, g.T);
    B.groups.push_back(&g);
  }
  const size_t N = maxN, T = maxT;
  TRY(ensure(c, B.rows, sizeof(float) * N * 16));
  TRY(ensure(c, B.key_a, sizeof(uint64_t) * N));
  TRY(ensure(c, B.key_b, sizeof(uint64_t) * N));
  TRY(ensure(c, B.val_a, sizeof(uint32_t) * N));
  TRY(ensure(c, B.val_b, sizeof(uint32_t) * N));
  TRY(ensure(c, B.srows, sizeof(float) * N * 16));
  TRY(ensure(c, B.skeys, sizeof(uint32_t) * N));
  TRY(ensure(c, B.spid, sizeof(uint32_t) * N));
  TRY(ensure(c, B.alive, N));
  TRY(ensure(c, B.s1idx, sizeof(uint32_t) * N));
  TRY(ensure(c, B.s1rows, sizeof(float) * N * 16));
  TRY(ensure(c, B.s1keys, sizeof(uint32_t) * N));
  TRY(ensure(c, B.s1pid, sizeof(uint32_t) * N));
  TRY(ensure(c, B.masks, sizeof(uint32_t) * N));
  TRY(ensure(c, B.status, N));
  TRY(ensure(c, B.flag_pt, N));
  TRY(ensure(c, B.flag_s1, N));
  TRY(ensure(c, B.eidx, sizeof(uint32_t) * N));
  TRY(ensure(c, B.erows, sizeof(float) * N * 16));
  TRY(ensure(c, B.epid, sizeof(uint32_t) * N));
  TRY(ensure(c, B.pos, sizeof(uint32_t) * N));
  TRY(ensure(c, B.codes, sizeof(uint2) * N));
  TRY(ensure(c, B.rank, sizeof(uint32_t) * N));
  TRY(ensure(c, B.piv, sizeof(uint32_t) * T * 16 * sky::kCodePivots));
  TRY(ensure(c, B.minkey, sizeof(uint32_t) * T * 16));
  TRY(ensure(c, B.sminkey, sizeof(uint32_t) * T * 16));
  TRY(ensure(c, B.seg_a, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, B.seg_b, sizeof(uint32_t) * (T + 1)));
  DevBuf* ib[] = {&B.ikey, &B.ikey2, &B.ival, &B.isort, &B.ibkey, &B.iblocal, &B.ibmk, &B.ibstart,
                  &B.ibcnt, &B.ibrun};
  for (DevBuf* b : ib) TRY(ensure(c, *b, sizeof(uint32_t) * N));
  TRY(ensure(c, B.ibcode, sizeof(uint2) * N));
  TRY(ensure(c, B.ihead, N));
  TRY(ensure(c, B.igfirst, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, B.gcnt, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, B.goff, sizeof(uint32_t) * (T + 1)));
  TRY(ensure(c, B.gsum, sizeof(uint2) * (N / 64 + T + 1)));
  uint64_t tab = 1;
  for (const BatchGroup* g : B.groups)
    if (batch_self_index(g->D, g->T)) tab = std::max<uint64_t>(tab, ((uint64_t)g->T << g->D) + 1);
  TRY(ensure(c, B.iboff, sizeof(uint32_t) * tab));
  size_t tb = 0;
  for (const BatchGroup* g : B.groups) {
    size_t b = 0;
    TRY(batch_tmp_bytes(c, g->N, g->T, &b));
    tb = std::max(tb, b);
  }
  TRY(ensure(c, B.tmp, tb));
  B.tmp_bytes = tb;
  B.sig_rows = rows_dev;
  B.sig_tune = c->tune_ver;
  B.sig_n.assign(n, n + n_inst);
  B.sig_d.assign(d, d + n_inst);
  B.sig_m.assign(n_inst, 0);
  if (max_d)
    for (uint32_t t = 0; t < n_inst; ++t) B.sig_m[t] = max_d[t];
  B.ready = true;
  return SKY_OK;
}

// Segmented sum of 0/1 flags per instance into cnt[0..T) with cnt[T] = 0,
// then an exclusive scan into off[0..T].
int seg_count_scan(sky_ctx* c, const uint8_t* flags, const uint32_t* seg, uint32_t T,
                   uint32_t* cnt, uint32_t* off) {
  BatchState& B = c->batch;
  TRY(fill32(c, cnt, T + 1, 0));
  hipcub::TransformInputIterator<uint32_t, hipcub::CastOp<uint32_t>, const uint8_t*> cast(
      flags, hipcub::CastOp<uint32_t>());
  size_t b = B.tmp_bytes;
  HIPCHK(hipcub::DeviceSegmentedReduce::Sum(B.tmp.p, b, cast, cnt, T, seg, seg + 1, c->stream));
  b = B.tmp_bytes;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(B.tmp.p, b, cnt, off, T + 1, c->stream));
  return SKY_OK;
}

int select_dev(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count) {
  hipcub::CountingInputIterator<uint32_t> it(0);
  size_t b = c->batch.tmp_bytes;
  HIPCHK(hipcub::DeviceSelect::Flagged(c->batch.tmp.p, b, it, flags, out, count, n, c->stream));
  return SKY_OK;
}

// Keys (instance << D | median mask) of the entries j < *count of a group
// (segments off) sorted with their positions: ikey2 / isort.
template <int D>
int batch_mask_sort(sky_ctx* c, const uint32_t* off, uint32_t T, uint32_t N,
                    const uint32_t* count) {
  BatchState& B = c->batch;
  hipLaunchKernelGGL(sky::k_bidx_keys<D>, dim3(grid_for(N)), dim3(kBlock), 0, c->stream,
                     P<uint2>(B.codes), off, T, count, N, P<uint32_t>(B.ikey), P<uint32_t>(B.ival));
  LAUNCHCHK();
  int tbits = 0;
  while ((1u << tbits) < T) ++tbits;
  size_t b = B.tmp_bytes;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(B.tmp.p, b, P<uint32_t>(B.ikey), P<uint32_t>(B.ikey2),
                                            P<uint32_t>(B.ival), P<uint32_t>(B.isort), N, 0,
                                            std::min(32, tbits + D), c->stream));
  return SKY_OK;
}

// The membership stage of a group: task numbering for tiles of PP points,
// then the persistent kernel.
template <int D, int PP>
int batch_member(sky_ctx* c, BatchGroup& g, const uint32_t* eoff, const Tuning& tu) {
  BatchState& B = c->batch;
  const uint32_t T = g.T;
  uint32_t* cnt = P<uint32_t>(g.cnt);
  hipLaunchKernelGGL(sky::k_btasks, dim3(grid_for(T + 1)), dim3(kBlock), 0, c->stream, eoff, T,
                     (uint32_t)PP, P<uint32_t>(B.seg_a));
  LAUNCHCHK();
  {
    size_t bb = B.tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(B.tmp.p, bb, P<uint32_t>(B.seg_a),
                                            P<uint32_t>(B.seg_b), T + 1, c->stream));
  }
  constexpr int NT = sky::member_threads<D>();
  if (tu.order == 2)  // code-order groups for every instance
    hipLaunchKernelGGL((sky::k_bmember<D, PP, NT, true>), dim3(tu.grid), dim3(NT), 0, c->stream,
                       P<float>(B.erows), P<uint2>(B.codes), P<uint32_t>(B.pos), eoff, T,
                       P<uint32_t>(B.seg_b), P<uint32_t>(g.d_maxd), P<uint32_t>(B.minkey),
                       P<uint32_t>(g.dom), cnt + 4, tu.cover, tu.head,
                       tu.min_bucket, P<uint32_t>(B.goff), P<uint2>(B.gsum), P<uint2>(B.ibcode),
                       P<uint32_t>(B.iblocal), nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
  else
    hipLaunchKernelGGL((sky::k_bmember<D, PP, NT>), dim3(tu.grid), dim3(NT), 0, c->stream,
                       P<float>(B.erows), P<uint2>(B.codes), P<uint32_t>(B.pos), eoff, T,
                       P<uint32_t>(B.seg_b), P<uint32_t>(g.d_maxd), P<uint32_t>(B.minkey),
                       P<uint32_t>(g.dom), cnt + 4, tu.cover, tu.head,
                       tu.min_bucket, nullptr, P<uint2>(B.gsum), P<uint2>(B.ibcode),
                       P<uint32_t>(B.iblocal), P<uint32_t>(B.igfirst), P<uint32_t>(B.ibmk),
                       P<uint32_t>(B.ibstart), P<uint32_t>(B.ibcnt), P<uint32_t>(B.iblocal),
                       P<uint2>(B.ibcode));
  LAUNCHCHK();
  return SKY_OK;
}

// Enqueues one group's whole pass on the context stream (no host sync, no
// allocation, no memset: graph-capturable).
template <int D>
int batch_group(sky_ctx* c, BatchGroup& g, const float* rows_dev) {
  BatchState& B = c->batch;
  const uint32_t N = g.N, T = g.T;
  uint32_t* cnt = P<uint32_t>(g.cnt);  // [0] err [1] survivors [2] ext ids [3] ext score [4] work
  const uint32_t* off = P<uint32_t>(g.d_off);
  const unsigned gN = grid_for(N);
  TRY(fill32(c, cnt, kGroupCounters, 0));
  TRY(fill8(c, B.flag_pt.p, N, 0));
  hipLaunchKernelGGL(sky::k_bcopy<D>, dim3(gN), dim3(kBlock), 0, c->stream, rows_dev,
                     P<uint64_t>(g.d_src_foff), off, T, N, P<float>(B.rows));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bscore<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.rows), off,
                     T, N, P<uint64_t>(B.key_a), P<uint32_t>(B.val_a), cnt);
  LAUNCHCHK();
  int tbits = 0;
  while ((1u << tbits) < T) ++tbits;
  size_t b = B.tmp_bytes;
  HIPCHK(hipcub::DeviceRadixSort::SortPairs(B.tmp.p, b, P<uint64_t>(B.key_a), P<uint64_t>(B.key_b),
                                            P<uint32_t>(B.val_a), P<uint32_t>(B.val_b), N, 0,
                                            32 + tbits, c->stream));
  hipLaunchKernelGGL(sky::k_bgather<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.rows),
                     P<uint64_t>(B.key_b), P<uint32_t>(B.val_b), N, P<float>(B.srows),
                     P<uint32_t>(B.skeys), P<uint32_t>(B.spid));
  LAUNCHCHK();
  const Tuning& tu = c->tune[D];
  hipLaunchKernelGGL(sky::k_bprune<D>, dim3(g.ntasks), dim3(kBlock), 0, c->stream,
                     P<float>(B.srows), P<uint32_t>(B.skeys), off, P<uint2>(g.d_tasks),
                     tu.pruners, P<uint8_t>(B.alive));
  LAUNCHCHK();
  TRY(select_dev(c, P<uint8_t>(B.alive), N, P<uint32_t>(B.s1idx), cnt + 1));
  uint32_t* s1off = P<uint32_t>(B.seg_b);
  TRY(seg_count_scan(c, P<uint8_t>(B.alive), off, T, P<uint32_t>(B.seg_a), s1off));
  hipLaunchKernelGGL(sky::k_bgather_idx<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<float>(B.srows), P<uint32_t>(B.skeys), P<uint32_t>(B.spid),
                     P<uint32_t>(B.s1idx), cnt + 1, P<float>(B.s1rows), P<uint32_t>(B.s1keys),
                     P<uint32_t>(B.s1pid));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpivots<D>, dim3(T * D), dim3(kBlock), 0, c->stream,
                     P<float>(B.s1rows), s1off, T, P<uint32_t>(B.piv));
  LAUNCHCHK();
  if (batch_self_index(D, T)) {
    // survivors' codes, then the instance-major mask index and its table
    hipLaunchKernelGGL(sky::k_bcodes<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, T, cnt + 1, P<uint32_t>(B.piv),
                       P<uint2>(B.codes));
    LAUNCHCHK();
    TRY(batch_mask_sort<D>(c, s1off, T, N, cnt + 1));
    hipLaunchKernelGGL(sky::k_bidx_bounds, dim3(grid_for(((uint64_t)T << D) + 1)), dim3(kBlock), 0,
                       c->stream, P<uint32_t>(B.ikey2), cnt + 1, T << D, P<uint32_t>(B.iboff));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bidx_gather, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(B.isort), P<uint32_t>(B.s1keys), P<uint2>(B.codes), cnt + 1,
                       nullptr, T, P<uint32_t>(B.ibkey), P<uint2>(B.ibcode), nullptr);
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bcolmin<D>, dim3(T), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, P<uint32_t>(B.sminkey));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bself_idx<D>, dim3((N + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint2>(B.codes), s1off, T,
                       cnt + 1, P<uint32_t>(B.iboff), P<uint32_t>(B.isort), P<uint32_t>(B.ibkey),
                       P<uint2>(B.ibcode), P<uint32_t>(B.sminkey), P<uint8_t>(B.status));
    LAUNCHCHK();
  } else {
    hipLaunchKernelGGL(sky::k_bmasks<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, T, cnt + 1, P<uint32_t>(B.piv),
                       P<uint32_t>(B.masks));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bself<D>, dim3((N + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint32_t>(B.masks), s1off, T,
                       cnt + 1, P<uint8_t>(B.status));
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(sky::k_bext_flags, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<uint8_t>(B.status), P<uint32_t>(B.s1pid), cnt + 1, N,
                     P<uint8_t>(B.flag_pt), P<uint8_t>(B.flag_s1));
  LAUNCHCHK();
  // extended skyline in id order (the result) and in score order
  TRY(select_dev(c, P<uint8_t>(B.flag_pt), N, P<uint32_t>(g.ext_pids), cnt + 2));
  uint32_t* eoff = P<uint32_t>(g.eoff);
  TRY(seg_count_scan(c, P<uint8_t>(B.flag_pt), off, T, P<uint32_t>(B.seg_a), eoff));
  TRY(select_dev(c, P<uint8_t>(B.flag_s1), N, P<uint32_t>(B.eidx), cnt + 3));
  hipLaunchKernelGGL(sky::k_bgather_idx<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<float>(B.s1rows), nullptr, P<uint32_t>(B.s1pid), P<uint32_t>(B.eidx),
                     cnt + 3, P<float>(B.erows), nullptr, P<uint32_t>(B.epid));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bscatter_rank, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(g.ext_pids), cnt + 2, P<uint32_t>(B.rank));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpos, dim3(gN), dim3(kBlock), 0, c->stream, P<uint32_t>(B.epid),
                     cnt + 3, P<uint32_t>(B.rank), P<uint32_t>(B.pos));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bcolmin<D>, dim3(T), dim3(kBlock), 0, c->stream, P<float>(B.erows),
                     eoff, P<uint32_t>(B.minkey));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpivots<D>, dim3(T * D), dim3(kBlock), 0, c->stream,
                     P<float>(B.erows), eoff, T, P<uint32_t>(B.piv));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bcodes<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.erows),
                     eoff, T, cnt + 3, P<uint32_t>(B.piv), P<uint2>(B.codes));
  LAUNCHCHK();
  if (tu.order == 2) {
    // code order per instance, groups of 64 with their summaries
    hipLaunchKernelGGL(sky::k_bplane_keys<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint2>(B.codes), cnt + 3, N, P<uint64_t>(B.key_a), P<uint32_t>(B.ival));
    LAUNCHCHK();
    size_t bb = B.tmp_bytes;
    HIPCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(
        B.tmp.p, bb, P<uint64_t>(B.key_a), P<uint64_t>(B.key_b), P<uint32_t>(B.ival),
        P<uint32_t>(B.isort), N, T, eoff, eoff + 1, 0, 4 * D, c->stream));
    hipLaunchKernelGGL(sky::k_bgroup_counts, dim3(grid_for(T + 1)), dim3(kBlock), 0, c->stream,
                       eoff, T, P<uint32_t>(B.gcnt));
    LAUNCHCHK();
    bb = B.tmp_bytes;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(B.tmp.p, bb, P<uint32_t>(B.gcnt), P<uint32_t>(B.goff),
                                            T + 1, c->stream));
    hipLaunchKernelGGL(sky::k_bgroup_sum<D>, dim3(grid_for(N / 64 + T + 1)), dim3(kBlock), 0,
                       c->stream, P<uint2>(B.codes), P<uint32_t>(B.isort), eoff, T,
                       P<uint32_t>(B.goff), P<uint2>(B.ibcode), P<uint32_t>(B.iblocal),
                       P<uint2>(B.gsum));
    LAUNCHCHK();
  } else {
    // bucket list of the extended skyline per instance (count in cnt[5])
    TRY(batch_mask_sort<D>(c, eoff, T, N, cnt + 3));
    hipLaunchKernelGGL(sky::k_bidx_heads, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(B.ikey2), cnt + 3, N, P<uint8_t>(B.ihead));
    LAUNCHCHK();
    TRY(select_dev(c, P<uint8_t>(B.ihead), N, P<uint32_t>(B.ival), cnt + 5));
    hipLaunchKernelGGL(sky::k_bidx_list<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(B.ikey2), P<uint32_t>(B.ival), cnt + 5, cnt + 3, N,
                       P<uint32_t>(B.ibmk), P<uint32_t>(B.ibstart), P<uint32_t>(B.ibcnt),
                       P<uint32_t>(B.ibrun));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bidx_first, dim3(grid_for(T + 1)), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(B.ibrun), cnt + 5, T, P<uint32_t>(B.igfirst));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bidx_gather, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(B.isort), nullptr, P<uint2>(B.codes), cnt + 3, eoff, T,
                       nullptr, P<uint2>(B.ibcode), P<uint32_t>(B.iblocal));
    LAUNCHCHK();
  }
  if (tu.tile == 0)
    TRY((batch_member<D, sky::member_points<D>()>(c, g, eoff, tu)));
  else
    TRY((batch_member<D, sky::member_points_alt<D>()>(c, g, eoff, tu)));
  return SKY_OK;
}

int batch_enqueue(sky_ctx* c, const float* rows_dev) {
  for (BatchGroup* g : c->batch.groups) {
    int rc = SKY_OK;
    SKY_DISPATCH_D(g->D, { rc = batch_group<D>(c, *g, rows_dev); });
    if (rc != SKY_OK) return rc;
  }
  return SKY_OK;
}

// Extended-skyline size and dimensionality of every instance of the last
// batch (group result offsets read back from the device).
int batch_sizes(sky_ctx* c, std::vector<uint32_t>& ne, std::vector<uint32_t>& dim) {
  BatchState& B = c->batch;
  ne.assign(B.n_inst, 0);
  dim.assign(B.sig_d.begin(), B.sig_d.end());
  std::vector<std::vector<uint32_t>> eo(B.groups.size());
  for (size_t gi = 0; gi < B.groups.size(); ++gi) {
    eo[gi].resize(B.groups[gi]->T + 1);
    HIPCHK(hipMemcpyAsync(eo[gi].data(), B.groups[gi]->eoff.p,
                          sizeof(uint32_t) * (B.groups[gi]->T + 1), hipMemcpyDeviceToHost,
                          c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t gi = 0; gi < B.groups.size(); ++gi) {
    const BatchGroup& g = *B.groups[gi];
    g.eo_host.assign(eo[gi].begin(), eo[gi].end());
    for (uint32_t k = 0; k < g.T; ++k) ne[g.inst[k]] = eo[gi][k + 1] - eo[gi][k];
  }
  return SKY_OK;
}

// Copies instance t's extended-skyline ids (local row numbers) to ext_dst[t]
// and its bitmaps to dom_dst[t] (null pointers skip an instance), straight
// from the group buffers; ids are made local on the device first.
int batch_copy_out(sky_ctx* c, uint32_t* const* ext_dst, uint32_t* const* dom_dst) {
  BatchState& B = c->batch;
  std::vector<uint32_t> ne, dim;
  TRY(batch_sizes(c, ne, dim));
  for (BatchGroup* gp : B.groups) {
    BatchGroup& g = *gp;
    const uint32_t tot = g.eo_host[g.T];
    if (!tot) continue;
    const uint32_t W = sky_words_per_point(g.D);
    uint32_t* loc = P<uint32_t>(B.isort);  // scratch of the last pass
    hipLaunchKernelGGL(sky::k_blocal, dim3(grid_for(tot)), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(g.ext_pids), tot, P<uint32_t>(g.d_off), g.T, loc);
    LAUNCHCHK();
    for (uint32_t k = 0; k < g.T; ++k) {
      const uint32_t t = g.inst[k], e0 = g.eo_host[k], n_e = g.eo_host[k + 1] - e0;
      if (!n_e) continue;
      if (ext_dst[t])
        HIPCHK(hipMemcpyAsync(ext_dst[t], loc + e0, sizeof(uint32_t) * n_e,
                              hipMemcpyDeviceToHost, c->stream));
      if (dom_dst[t])
        HIPCHK(hipMemcpyAsync(dom_dst[t], P<uint32_t>(g.dom) + (size_t)e0 * W,
                              sizeof(uint32_t) * (size_t)n_e * W, hipMemcpyDeviceToHost,
                              c->stream));
    }
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

bool batch_same(const BatchState& B, const float* rows, uint32_t n_inst, const uint32_t* n,
                const uint32_t* d, const uint32_t* max_d) {
  if (!B.ready || B.sig_rows != rows || B.n_inst != n_inst) return false;
  for (uint32_t t = 0; t < n_inst; ++t)
    if (B.sig_n[t] != n[t] || B.sig_d[t] != d[t] || B.sig_m[t] != (max_d ? max_d[t] : 0u))
      return false;
  return true;
}

}  // namespace

extern "C" {

uint32_t sky_words_per_point(uint32_t d) { return d >= 5 ? (1u << (d - 5)) : 1u; }

int sky_ctx_create(int device, size_t hbm_budget_bytes, sky_ctx** out) {
  if (!out) return SKY_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return SKY_ERR_DEVICE;
  sky_ctx* c = new sky_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SKY_ERR_DEVICE;
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : free_b;
  *out = c;
  return SKY_OK;
}

}  // extern "C"

namespace {

// Frees every device buffer the context holds; results are gone afterwards.
void free_all(sky_ctx* c) {
  DevBuf* bufs[] = {&c->rows,   &c->key_a,  &c->key_b,    &c->pid_a,  &c->pid_b,
                    &c->srows,  &c->alive,  &c->s1idx,    &c->s1rows, &c->s1keys,
                    &c->s1pid,  &c->status, &c->flag,     &c->ext_pids, &c->erows,
                    &c->dom,    &c->cub_tmp, &c->counter, &c->cand,  &c->eidx,
                    &c->epid,   &c->pos,    &c->codes,    &c->piv,   &c->masks,
                    &c->hc_cnt, &c->hc_off, &c->hc_key_a, &c->hc_key_b, &c->hc_val_a,
                    &c->hc_val_b, &c->hc_ukey, &c->hc_runs, &c->hc_gw, &c->hc_gv,
                    &c->hc_goff, &c->bmask,  &c->biota,  &c->bidx,   &c->bkey,   &c->bcode,
                    &c->boff,    &c->bcnt,   &c->gsum,   &c->defer};
  for (DevBuf* b : bufs) free_buf(c, *b);
  free_batch_buffers(c);
  c->has_mdmc = c->has_hc = c->has_batch = false;
}

}  // namespace

extern "C" {

void sky_ctx_destroy(sky_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int sky_ctx_set_option(sky_ctx* c, int option, int64_t value) {
  if (!c) return SKY_ERR_ARG;
  const int opt = option & 0xFF, dd = option >> 8;
  if (dd < 0 || dd > 16) return fail(c, SKY_ERR_ARG, "option %d: dimensionality %d", option, dd);
  if (opt == SKY_OPT_MEMBER_INDEX) {
    if (value < -1 || value > 2 || dd) return fail(c, SKY_ERR_ARG, "bad member index option");
    c->opt_member_index = (int)value;
    return SKY_OK;
  }
  struct Range {
    int opt;
    int64_t lo, hi;
    uint32_t Tuning::*field;
  };
  static const Range ranges[] = {
      {SKY_OPT_COVER, 0, 1 << 16, &Tuning::cover},
      {SKY_OPT_PRUNERS, 64, 1 << 13, &Tuning::pruners},
      {SKY_OPT_MIN_BUCKET, 1, 1 << 30, &Tuning::min_bucket},
      {SKY_OPT_MEMBER_GRID, 64, 1 << 13, &Tuning::grid},
      {SKY_OPT_TILE, 0, 1, &Tuning::tile},
      {SKY_OPT_BATCH_ORDER, 0, 2, &Tuning::order},
      {SKY_OPT_COVER_HEAD, 0, 1 << 16, &Tuning::head},
  };
  for (const Range& r : ranges) {
    if (r.opt != opt) continue;
    if (value < r.lo || value > r.hi || (opt == SKY_OPT_COVER && value > 0 && value < 64))
      return fail(c, SKY_ERR_ARG, "option %d: value %lld outside [%lld, %lld]", opt,
                  (long long)value, (long long)r.lo, (long long)r.hi);
    for (int k = dd ? dd : 1; k <= (dd ? dd : 16); ++k) c->tune[k].*(r.field) = (uint32_t)value;
    c->tune_ver++;
    return SKY_OK;
  }
  return fail(c, SKY_ERR_ARG, "unknown option %d", option);
}

int sky_ctx_get_option(sky_ctx* c, int option, int64_t* value) {
  if (!c || !value) return SKY_ERR_ARG;
  const int opt = option & 0xFF, dd = option >> 8;
  if (dd < 0 || dd > 16) return SKY_ERR_ARG;
  const Tuning& t = c->tune[dd ? dd : 1];
  switch (opt) {
    case SKY_OPT_MEMBER_INDEX: *value = c->opt_member_index; return SKY_OK;
    case SKY_OPT_COVER: *value = t.cover; return SKY_OK;
    case SKY_OPT_PRUNERS: *value = t.pruners; return SKY_OK;
    case SKY_OPT_MIN_BUCKET: *value = t.min_bucket; return SKY_OK;
    case SKY_OPT_MEMBER_GRID: *value = t.grid; return SKY_OK;
    case SKY_OPT_TILE: *value = t.tile; return SKY_OK;
    case SKY_OPT_BATCH_ORDER: *value = t.order; return SKY_OK;
    case SKY_OPT_COVER_HEAD: *value = t.head; return SKY_OK;
    default: return fail(c, SKY_ERR_ARG, "unknown option %d", option);
  }
}

int sky_batch_autotune(sky_ctx* c, const float* rows, int rows_on_device, uint32_t n_inst,
                       const uint32_t* n, const uint32_t* d, const uint32_t* max_d, char* report,
                       size_t report_cap) {
  TRY(check_ctx(c));
  if (!n || !d || (n_inst && !rows)) return fail(c, SKY_ERR_ARG, "null argument");
  uint64_t tot = 0;
  // a first pass builds the batch and loads every kernel
  TRY(sky_batch_mdmc(c, rows, rows_on_device, n_inst, n, d, max_d, 0, &tot));
  const float* rows_dev = rows_on_device ? rows : P<float>(c->rows);
  BatchState& B = c->batch;
  const size_t G = B.groups.size();
  std::vector<hipEvent_t> ev(2 * G, nullptr);
  for (hipEvent_t& e : ev) HIPCHK(hipEventCreate(&e));
  struct Knob {
    const char* name;
    uint32_t Tuning::*field;
    std::vector<uint32_t> values;
  };
  const Knob knobs[] = {
      {"order", &Tuning::order, {2, 0}},
      {"tile", &Tuning::tile, {0, 1}},
      {"cover", &Tuning::cover, {0, 512, 1024, 2048, 4096}},  // 0: auto_cover
      {"pruners", &Tuning::pruners, {128, 256, 512, 1024}},
      {"min_bucket", &Tuning::min_bucket, {4, 16, 64, 1u << 30}},
      {"grid", &Tuning::grid, {256, 512, 1024, 2048}},
  };
  std::vector<float> best_ms(G, 0.0f);
  int rc = SKY_OK;
  for (const Knob& k : knobs) {
    std::vector<uint32_t> best(G);
    std::vector<float> bms(G, 1e30f);
    for (uint32_t v : k.values) {
      for (BatchGroup* g : B.groups) c->tune[g->D].*(k.field) = v;
      for (size_t gi = 0; gi < G && rc == SKY_OK; ++gi) {
        BatchGroup& g = *B.groups[gi];
        if (hipEventRecord(ev[2 * gi], c->stream) != hipSuccess) rc = SKY_ERR_DEVICE;
        if (rc == SKY_OK) SKY_DISPATCH_D(g.D, { rc = batch_group<D>(c, g, rows_dev); });
        if (rc == SKY_OK && hipEventRecord(ev[2 * gi + 1], c->stream) != hipSuccess)
          rc = SKY_ERR_DEVICE;
      }
      if (rc == SKY_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = SKY_ERR_DEVICE;
      if (rc != SKY_OK) break;
      for (size_t gi = 0; gi < G; ++gi) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ev[2 * gi], ev[2 * gi + 1]) != hipSuccess) rc = SKY_ERR_DEVICE;
        if (ms < bms[gi]) {
          bms[gi] = ms;
          best[gi] = v;
        }
      }
    }
    if (rc != SKY_OK) break;
    for (size_t gi = 0; gi < G; ++gi) {
      c->tune[B.groups[gi]->D].*(k.field) = best[gi];
      best_ms[gi] = bms[gi];
    }
  }
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  c->tune_ver++;
  c->has_batch = false;  // the tuning passes overwrote the results
  if (rc != SKY_OK) return fail(c, rc, "autotuning pass failed");
  if (report && report_cap) {
    std::string out = "{";
    for (size_t gi = 0; gi < G; ++gi) {
      const int D = B.groups[gi]->D;
      const Tuning& t = c->tune[D];
      char buf[256];
      std::snprintf(buf, sizeof(buf),
                    "%s\"%d\": {\"instances\": %u, \"order\": %u, \"tile\": %u, "
                    "\"cover\": %u, \"pruners\": %u, \"min_bucket\": %u, \"grid\": %u, "
                    "\"ms\": %.3f}",
                    gi ? ", " : "", D, B.groups[gi]->T, t.order, t.tile, t.cover, t.pruners,
                    t.min_bucket, t.grid, best_ms[gi]);
      out += buf;
    }
    out += "}";
    std::snprintf(report, report_cap, "%s", out.c_str());
  }
  return SKY_OK;
}

int sky_ctx_release(sky_ctx* c) {
  TRY(check_ctx(c));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_all(c);
  return SKY_OK;
}

const char* sky_last_error(const sky_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t sky_ctx_bytes_in_use(const sky_ctx* c) { return c ? c->in_use : 0; }

int sky_ctx_set_timing(sky_ctx* c, int enable) {
  if (!c) return SKY_ERR_ARG;
  c->timing = enable != 0;
  return SKY_OK;
}

int sky_ctx_stage_times(const sky_ctx* c, float* ms, int count) {
  if (!c || !ms || count < 0) return SKY_ERR_ARG;
  for (int k = 0; k < count && k < SKY_STAGE_COUNT; ++k) ms[k] = c->stage_ms[k];
  return SKY_OK;
}

int sky_mdmc_device(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d,
                    uint32_t* n_ext) {
  TRY(check_ctx(c));
  TRY(check_dims(c, d));
  if (!n_ext || (n && !rows_dev)) return fail(c, SKY_ERR_ARG, "null argument");
  return run_mdmc(c, rows_dev, n, d, max_d, n_ext);
}

int sky_mdmc(sky_ctx* c, const float* rows, uint32_t n, uint32_t d, uint32_t max_d,
             uint32_t* ext_pids, uint32_t* n_ext, uint32_t* dom_words, uint64_t dom_cap_words) {
  TRY(check_ctx(c));
  TRY(check_dims(c, d));
  if (!n_ext || !ext_pids || (n && !rows)) return fail(c, SKY_ERR_ARG, "null argument");
  TRY(upload_rows(c, rows, n, d));
  TRY(run_mdmc(c, P<float>(c->rows), n, d, max_d, n_ext));
  const uint32_t ne = *n_ext;
  if (ne)
    HIPCHK(hipMemcpyAsync(ext_pids, c->ext_pids.p, sizeof(uint32_t) * ne, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!dom_words) return SKY_OK;
  return sky_mdmc_fetch(c, nullptr, dom_words, dom_cap_words);
}

int sky_mdmc_fetch(sky_ctx* c, uint32_t* ext_pids, uint32_t* dom_words, uint64_t dom_cap_words) {
  TRY(check_ctx(c));
  if (!c->has_mdmc) return fail(c, SKY_ERR_STATE, "no MDMC result is held");
  const uint64_t need = (uint64_t)c->res_n_ext * c->res_W;
  if (dom_words && dom_cap_words < need)
    return fail(c, SKY_ERR_CAPACITY, "dom_words holds %llu words, %llu needed",
                (unsigned long long)dom_cap_words, (unsigned long long)need);
  if (ext_pids && c->res_n_ext)
    HIPCHK(hipMemcpyAsync(ext_pids, c->ext_pids.p, sizeof(uint32_t) * c->res_n_ext,
                          hipMemcpyDeviceToHost, c->stream));
  if (dom_words && need)
    HIPCHK(hipMemcpyAsync(dom_words, c->dom.p, sizeof(uint32_t) * need, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

int sky_mdmc_device_results(sky_ctx* c, const uint32_t** ext_pids_dev,
                            const uint32_t** dom_words_dev) {
  TRY(check_ctx(c));
  if (!c->has_mdmc) return fail(c, SKY_ERR_STATE, "no MDMC result is held");
  if (ext_pids_dev) *ext_pids_dev = P<uint32_t>(c->ext_pids);
  if (dom_words_dev) *dom_words_dev = P<uint32_t>(c->dom);
  return SKY_OK;
}

int sky_mdmc_hashcube(sky_ctx* c, uint64_t* n_groups, uint64_t* n_entries, uint32_t* group_word,
                      uint32_t* group_value, uint64_t* group_off, uint32_t* pids,
                      uint64_t cap_groups, uint64_t cap_entries) {
  TRY(check_ctx(c));
  if (!n_groups || !n_entries) return fail(c, SKY_ERR_ARG, "null argument");
  if (!c->has_mdmc) return fail(c, SKY_ERR_STATE, "no MDMC result is held");
  if (!c->has_hc) TRY(build_hashcube(c));
  *n_groups = c->hc_groups;
  *n_entries = c->hc_entries;
  if (!group_word || !group_value || !group_off || !pids) return SKY_OK;
  if (cap_groups < c->hc_groups || cap_entries < c->hc_entries)
    return fail(c, SKY_ERR_CAPACITY, "HashCube needs %llu groups and %llu entries",
                (unsigned long long)c->hc_groups, (unsigned long long)c->hc_entries);
  const uint64_t ng = c->hc_groups, nt = c->hc_entries;
  if (ng) {
    HIPCHK(hipMemcpyAsync(group_word, c->hc_gw.p, sizeof(uint32_t) * ng, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(group_value, c->hc_gv.p, sizeof(uint32_t) * ng,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(group_off, c->hc_goff.p, sizeof(uint64_t) * (ng + 1),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pids, c->hc_val_b.p, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost,
                          c->stream));
  } else {
    group_off[0] = 0;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

int sky_batch_mdmc(sky_ctx* c, const float* rows, int rows_on_device, uint32_t n_inst,
                   const uint32_t* n, const uint32_t* d, const uint32_t* max_d, int use_graph,
                   uint64_t* total_ext) {
  TRY(check_ctx(c));
  if (!n || !d || !total_ext || (n_inst && !rows)) return fail(c, SKY_ERR_ARG, "null argument");
  uint64_t total_f = 0;
  for (uint32_t t = 0; t < n_inst; ++t) {
    if (d[t] < 1 || d[t] > 16) return fail(c, SKY_ERR_ARG, "instance %u: d=%u outside [1,16]", t, d[t]);
    total_f += (uint64_t)n[t] * d[t];
  }
  c->has_mdmc = false;
  c->has_hc = false;
  c->has_batch = false;
  const float* rows_dev = rows;
  if (!rows_on_device) {
    TRY(ensure(c, c->rows, sizeof(float) * std::max<uint64_t>(total_f, 1)));
    if (total_f)
      HIPCHK(hipMemcpyAsync(c->rows.p, rows, sizeof(float) * total_f, hipMemcpyHostToDevice,
                            c->stream));
    rows_dev = P<float>(c->rows);
  }
  BatchState& B = c->batch;
  if (!batch_same(B, rows_dev, n_inst, n, d, max_d) || B.sig_tune != c->tune_ver)
    TRY(batch_setup(c, rows_dev, n_inst, n, d, max_d));
  if (use_graph) {
    if (!B.exec) {
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = batch_enqueue(c, rows_dev);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (rc != SKY_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess)
        return fail(c, SKY_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
      B.graph = graph;
      HIPCHK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(B.exec, c->stream));
  } else {
    TRY(batch_enqueue(c, rows_dev));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  uint64_t tot = 0;
  std::vector<uint32_t> h(kGroupCounters * B.groups.size());
  for (size_t gi = 0; gi < B.groups.size(); ++gi)
    HIPCHK(hipMemcpyAsync(&h[gi * kGroupCounters], B.groups[gi]->cnt.p,
                          sizeof(uint32_t) * kGroupCounters, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t gi = 0; gi < B.groups.size(); ++gi) {
    if (h[gi * kGroupCounters] & 2u) return fail(c, SKY_ERR_NAN, "input holds NaN");
    tot += h[gi * kGroupCounters + 2];
  }
  *total_ext = tot;
  c->has_batch = true;
  return SKY_OK;
}

int sky_batch_fetch(sky_ctx* c, uint64_t* ext_off, uint32_t* ext_pids, uint64_t* dom_off,
                    uint32_t* dom_words, uint64_t cap_ext, uint64_t cap_words) {
  TRY(check_ctx(c));
  if (!ext_off || !dom_off) return fail(c, SKY_ERR_ARG, "null argument");
  if (!c->has_batch) return fail(c, SKY_ERR_STATE, "no batch result is held");
  const uint32_t NI = c->batch.n_inst;
  std::vector<uint32_t> ne, dim;
  TRY(batch_sizes(c, ne, dim));
  ext_off[0] = dom_off[0] = 0;
  for (uint32_t t = 0; t < NI; ++t) {
    ext_off[t + 1] = ext_off[t] + ne[t];
    dom_off[t + 1] = dom_off[t] + (uint64_t)ne[t] * sky_words_per_point(dim[t]);
  }
  if (!ext_pids || !dom_words) return SKY_OK;
  if (cap_ext < ext_off[NI] || cap_words < dom_off[NI])
    return fail(c, SKY_ERR_CAPACITY, "batch needs %llu ids and %llu words",
                (unsigned long long)ext_off[NI], (unsigned long long)dom_off[NI]);
  std::vector<uint32_t*> ed(NI), dd(NI);
  for (uint32_t t = 0; t < NI; ++t) {
    ed[t] = ext_pids + ext_off[t];
    dd[t] = dom_words + dom_off[t];
  }
  return batch_copy_out(c, ed.data(), dd.data());
}

int sky_batch_fetch_into(sky_ctx* c, uint32_t* const* ext_dst, uint32_t* const* dom_dst) {
  TRY(check_ctx(c));
  if (!ext_dst || !dom_dst) return fail(c, SKY_ERR_ARG, "null argument");
  if (!c->has_batch) return fail(c, SKY_ERR_STATE, "no batch result is held");
  return batch_copy_out(c, ext_dst, dom_dst);
}

int sky_skyalign(sky_ctx* c, const float* rows, uint32_t n_rows, uint32_t d, uint32_t subspace,
                 const uint32_t* cand, uint32_t n_cand, uint32_t* sky_out, uint32_t* n_sky,
                 uint32_t* ext_out, uint32_t* n_ext) {
  TRY(check_ctx(c));
  TRY(check_dims(c, d));
  if (subspace == 0 || (subspace >> d) != 0)
    return fail(c, SKY_ERR_ARG, "subspace 0x%x invalid for d=%u", subspace, d);
  if (!sky_out || !n_sky || !ext_out || !n_ext || (n_rows && !rows))
    return fail(c, SKY_ERR_ARG, "null argument");
  c->has_mdmc = false;
  TRY(upload_rows(c, rows, n_rows, d));
  const uint32_t nc = cand ? n_cand : n_rows;
  const uint32_t* cand_dev = nullptr;
  if (cand) {
    TRY(ensure(c, c->cand, sizeof(uint32_t) * (size_t)n_cand));
    if (n_cand)
      HIPCHK(hipMemcpyAsync(c->cand.p, cand, sizeof(uint32_t) * n_cand, hipMemcpyHostToDevice,
                            c->stream));
    cand_dev = P<uint32_t>(c->cand);
  }
  std::vector<uint32_t> sky, ext;
  TRY(skyalign_dev(c, P<float>(c->rows), n_rows, d, subspace, cand_dev, nc, sky, ext));
  std::copy(sky.begin(), sky.end(), sky_out);
  std::copy(ext.begin(), ext.end(), ext_out);
  *n_sky = (uint32_t)sky.size();
  *n_ext = (uint32_t)ext.size();
  return SKY_OK;
}

int sky_sdsc(sky_ctx* c, const float* rows, uint32_t n, uint32_t d, uint32_t max_d,
             uint64_t* sky_off, uint32_t* sky_ids, uint64_t sky_cap, uint64_t* ext_off,
             uint32_t* ext_ids, uint64_t ext_cap) {
  TRY(check_ctx(c));
  TRY(check_dims(c, d));
  if (!sky_off || !ext_off || (n && !rows)) return fail(c, SKY_ERR_ARG, "null argument");
  if (max_d == 0 || max_d > d) max_d = d;
  c->has_mdmc = false;
  const uint32_t S = 1u << d, full = S - 1u;
  std::vector<std::vector<uint32_t>> lsky(S), lext(S);
  TRY(upload_rows(c, rows, n, d));
  const float* rows_dev = P<float>(c->rows);
  // full space over every point (SDSC.cu:274)
  TRY(skyalign_dev(c, rows_dev, n, d, full, nullptr, n, lsky[full], lext[full]));
  std::vector<uint32_t> cand;
  // levels top-down, skipping popcount > max_d (SDSC.cu:291-294)
  for (int lev = (int)d - 1; lev >= 1; --lev) {
    if ((uint32_t)lev > max_d) continue;
    for (uint32_t s = 1; s < full; ++s) {
      if (__builtin_popcount(s) != lev) continue;
      // parents in generate_lattice order (descending id); the smallest
      // sky+ext wins, the full space stands in when parents were skipped
      // (src/sdsc/SDSC.cu:81-104,365-376)
      uint32_t src = full;
      uint32_t first = 0;
      for (int j = (int)d - 1; j >= 0; --j)
        if (!(s >> j & 1u)) {
          first = s | (1u << j);
          break;
        }
      if (!lsky[first].empty()) {
        size_t best = (size_t)-1;
        for (int j = (int)d - 1; j >= 0; --j) {
          if (s >> j & 1u) continue;
          const uint32_t par = s | (1u << j);
          const size_t sz = lsky[par].size() + lext[par].size();
          if (sz < best) {
            best = sz;
            src = par;
          }
        }
      }
      cand.assign(lsky[src].begin(), lsky[src].end());
      cand.insert(cand.end(), lext[src].begin(), lext[src].end());
      TRY(ensure(c, c->cand, sizeof(uint32_t) * std::max<size_t>(cand.size(), 1)));
      if (!cand.empty())
        HIPCHK(hipMemcpyAsync(c->cand.p, cand.data(), sizeof(uint32_t) * cand.size(),
                              hipMemcpyHostToDevice, c->stream));
      TRY(skyalign_dev(c, rows_dev, n, d, s, P<uint32_t>(c->cand), (uint32_t)cand.size(),
                       lsky[s], lext[s]));
    }
  }
  uint64_t ts = 0, te = 0;
  sky_off[0] = 0;
  ext_off[0] = 0;
  for (uint32_t s = 1; s < S; ++s) {
    ts += lsky[s].size();
    te += lext[s].size();
  }
  sky_off[0] = ext_off[0] = 0;
  uint64_t os = 0, oe = 0;
  for (uint32_t s = 0; s < S; ++s) {
    sky_off[s] = os;
    ext_off[s] = oe;
    if (s == 0) continue;
    if (sky_ids && os + lsky[s].size() <= sky_cap)
      std::copy(lsky[s].begin(), lsky[s].end(), sky_ids + os);
    if (ext_ids && oe + lext[s].size() <= ext_cap)
      std::copy(lext[s].begin(), lext[s].end(), ext_ids + oe);
    os += lsky[s].size();
    oe += lext[s].size();
  }
  sky_off[S] = os;
  ext_off[S] = oe;
  if (!sky_ids || !ext_ids || ts > sky_cap || te > ext_cap)
    return fail(c, SKY_ERR_CAPACITY, "sdsc needs %llu sky ids and %llu ext ids",
                (unsigned long long)ts, (unsigned long long)te);
  return SKY_OK;
}

}  // extern "C"
