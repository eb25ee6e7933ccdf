// internal.hpp -- state and helpers shared by the translation units of
// libskycube_amd.so (not part of the C ABI).
//
// The library is built from several translation units so that they compile
// in parallel: core.hip (the C ABI and every step that does not depend on the
// dimensionality), and stage_single.hip / stage_batch.hip, each compiled once
// per dimensionality D = 1..16 (-DSKY_D=D) with the stages templated on D.
// Every HIP return code is checked and turned into a SKY_ERR_* status with a
// message on the context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/skycube_amd.h"
#include "common.cuh"

namespace skyi {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// One dimensionality class of a batch: its instances and device state.
struct BatchGroup {
  int D = 0;
  std::vector<uint32_t> inst;  // caller's instance ids, in group order
  std::vector<uint32_t> off;   // point offsets of the instances (T + 1)
  uint32_t T = 0, N = 0, ntasks = 0, max_n = 0;  // max_n: the largest instance
  DevBuf d_off, d_src, d_tasks, d_maxd;       // descriptors (d_src: row address per instance)
  DevBuf ext_pids, eoff, dom, cnt;            // results; cnt: counters
  mutable std::vector<uint32_t> eo_host;       // eoff read back by the last fetch
};

struct BatchState {
  std::vector<uint32_t> sig_n, sig_d, sig_m;
  std::vector<uint64_t> sig_src;    // row address of every instance
  uint8_t* desc_host = nullptr;     // pinned staging of the descriptors
  size_t desc_cap = 0;
  hipEvent_t desc_done = nullptr;   // the staging's last copy
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
  DevBuf gcnt, goff, gsum;  // gsum: gsum_cap group summaries, then 8 subgroup summaries each
  DevBuf sminkey;         // column minima of the survivors per instance
  DevBuf rs_key, rs_val;  // the other half of the radix sorts' ping-pong
  size_t tmp_bytes = 0;
};

// Launch settings of one dimensionality (sky_ctx_set_option, autotuner).
struct Tuning {
  uint32_t cover = 0;  // candidates covered first; 0: auto_cover of each extended skyline
  uint32_t pruners = sky::kPruners;       // prefilter points
  uint32_t min_bucket = sky::kMinBucket;  // bucket-index threshold
  uint32_t grid = 2048;                   // workgroups of the persistent kernel
  uint32_t tile = 0;                      // 0: member_points, 1: half of it
  uint32_t order = 3;                     // batched membership: 3 subgroups, 2 groups, 0 scan / buckets
  uint32_t head = 0;                      // cover head tested pair by pair (0: cover / 8)
};

// Bump allocator over the allocation-free path's scratch array.
struct Carve {
  uint8_t* base;
  size_t cap, top;
};

// Bytes of the scratch array of the allocation-free MDMC (one per device).
constexpr uint64_t kNoallocScratchBytes = 64ull << 20;

}  // namespace skyi

struct sky_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::string err;
  // device buffers
  skyi::DevBuf rows, key_a, key_b, pid_a, pid_b, srows, alive, s1idx, s1rows, s1keys, s1pid,
      status, flag, ext_pids, erows, dom, cub_tmp, counter, cand, eidx, epid, pos, codes, piv,
      masks, hc_cnt, hc_off, hc_key_a, hc_key_b, hc_val_a, hc_val_b, hc_ukey, hc_runs, hc_gw,
      hc_gv, hc_goff, bmask, biota, bidx, bkey, bcode, boff, bcnt, gsum, defer;
  // the last batch
  skyi::BatchState batch;
  // membership candidate order: -1 by bucket size, 0 scan, 1 bucket index
  int opt_member_index = -1;
  skyi::Tuning tune[17];  // per dimensionality
  uint64_t tune_ver = 0;  // bumped on every change (cached graphs are rebuilt)
  bool has_batch = false;
  // grouping of the last HashCube export
  bool has_hc = false;
  uint64_t hc_groups = 0, hc_entries = 0, hc_sorted = 0;
  // a second stream (created on first use, side_stream): the export's early
  // copy of the ids, the membership's deferred points
  hipStream_t copy_stream = nullptr;
  hipEvent_t sorted_ev = nullptr, side_ev = nullptr;
  // stage timing (HIP events on `stream`)
  bool timing = false;
  hipEvent_t ev[SKY_STAGE_COUNT + 1] = {};
  float stage_ms[SKY_STAGE_COUNT] = {};
  // last MDMC result
  bool has_mdmc = false;
  uint32_t res_n_ext = 0, res_d = 0, res_W = 0;
  bool res_rowcnt = false;  // hc_off holds the result's per-row entry counts
  // set during sky_mdmc_device_noalloc: buffers come from the scratch array
  skyi::Carve* carve = nullptr;
  // the last SDSC lattice (sdsc.hip): per cuboid s an ascending id list (sky
  // and ext together) at sd_pool[sd_lat_off[s] ..] of sd_lat_cnt[s] ids,
  // sd_flag 1 for sky; sd_lat_sky[s] skyline sizes; index 0 = all points
  skyi::DevBuf sd_pool, sd_flag, sd_lat_off, sd_lat_cnt, sd_lat_sky, sd_subs, sd_src, sd_pcnt,
      sd_coff, sd_tasks, sd_cand, sd_keep, sd_skyf, sd_sel, sd_kcnt, sd_koff, sd_scnt, sd_soff,
      sd_counters;
  bool has_sdsc = false;
  uint64_t sd_used = 0;
  uint32_t sd_d = 0;
};

namespace skyi {

// Records `code` and a formatted message on the context; returns code.
int fail(sky_ctx* c, int code, const char* fmt, ...);

#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::skyi::fail(c, SKY_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,          \
                          hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

#define LAUNCHCHK()                                                                   \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return ::skyi::fail(c, SKY_ERR_DEVICE, "kernel launch failed: %s (%s:%d)",      \
                          hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

#define TRY(expr)                  \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != SKY_OK) return rc_; \
  } while (0)

// Grows b to at least `bytes` (1/8 headroom) within the context's budget;
// while c->carve is set, carves it from the scratch array instead.
int ensure(sky_ctx* c, DevBuf& b, size_t bytes);

// The device buffers of the single-instance path (every DevBuf of sky_ctx
// outside the batch state).
constexpr int kSingleBufs = 45;
inline void single_bufs(sky_ctx* c, DevBuf** out) {
  DevBuf* b[kSingleBufs] = {
      &c->rows,     &c->key_a,    &c->key_b,    &c->pid_a,    &c->pid_b,    &c->srows,
      &c->alive,    &c->s1idx,    &c->s1rows,   &c->s1keys,   &c->s1pid,    &c->status,
      &c->flag,     &c->ext_pids, &c->erows,    &c->dom,      &c->cub_tmp,  &c->counter,
      &c->cand,     &c->eidx,     &c->epid,     &c->pos,      &c->codes,    &c->piv,
      &c->masks,    &c->hc_cnt,   &c->hc_off,   &c->hc_key_a, &c->hc_key_b, &c->hc_val_a,
      &c->hc_val_b, &c->hc_ukey,  &c->hc_runs,  &c->hc_gw,    &c->hc_gv,    &c->hc_goff,
      &c->bmask,    &c->biota,    &c->bidx,     &c->bkey,     &c->bcode,    &c->boff,
      &c->bcnt,     &c->gsum,     &c->defer};
  for (int i = 0; i < kSingleBufs; ++i) out[i] = b[i];
}

// The single-instance MDMC on rows in device memory (core.hip).
int run_mdmc(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d,
             uint32_t* n_ext);

// Looks up the device address of the allocation-free path's scratch array
// on `device` (current) once; noalloc.hip.
int noalloc_init_device(int device);
// Upper bound of the scratch bytes of an allocation-free MDMC of n x d rows.
uint64_t noalloc_bytes(uint64_t n, uint32_t d);
void free_buf(sky_ctx* c, DevBuf& b);

template <typename T>
T* P(DevBuf& b) {
  return static_cast<T*>(b.p);
}

// Records the event that opens stage `k` (stage k-1 ends there).
int mark(sky_ctx* c, int k);
int read_marks(sky_ctx* c);

inline unsigned grid_for(uint64_t n) {
  return (unsigned)((n + sky::kBlock - 1) / sky::kBlock);
}

// The counter word of the context's stream, copied to the host (synchronous).
int copy_count(sky_ctx* c, uint32_t* host);

// Indices i in [0, n) with flags[i] != 0, ascending, into out; their count
// into the context counter.
int select_flagged(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out);

// Radix sorts of (key, value) pairs over key bits [0, end_bit) with rocPRIM
// (core.hip); the temporary storage is the context's cub_tmp buffer.
int sort_pairs(sky_ctx* c, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
               uint32_t* vout, uint32_t n, int end_bit);
int sort_pairs(sky_ctx* c, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
               uint32_t* vout, uint32_t n, int end_bit);

// Fills on the context stream by kernels (no memset nodes in a captured graph).
int fill32(sky_ctx* c, void* p, uint64_t n, uint32_t v);
// Creates the context's second stream and its events if not yet there.
int side_stream(sky_ctx* c);
int fill8(sky_ctx* c, void* p, uint64_t n, uint8_t v);

inline sky::DimList dims_of(uint32_t sub, uint32_t* ds) {
  sky::DimList L{};
  uint32_t k = 0;
  for (int j = 0; j < 16; ++j)
    if (sub >> j & 1u) L.idx[k++] = j;
  *ds = k;
  return L;
}

// The batched survivor pass uses the mask index when its bucket table (one
// entry per instance and median mask) stays within 2^26 entries.
// Code-order groups of a batch group of N points in T instances, at most
// (the subgroup summaries follow that many group summaries in gsum).
inline size_t gsum_cap(uint32_t N, uint32_t T) { return (size_t)N / 64 + T + 1; }
inline bool batch_self_index(int D, uint32_t T) { return ((uint64_t)T << D) <= (1ull << 26); }

constexpr int kGroupCounters = 8;  // err, survivors, ext (id order), ext (score order), work

// Batch primitives without memset (core.hip): segmented counts of 0/1 flags
// per instance and their exclusive scan; flagged select with a device count.
// cnt[t] = flags set in segment t, off = their exclusive scan (T + 1
// entries); max_len bounds the segments' lengths (it sets the parallelism).
int seg_count_scan(sky_ctx* c, const uint8_t* flags, const uint32_t* seg, uint32_t T,
                   uint32_t* cnt, uint32_t* off, uint32_t max_len);
int select_dev(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count);
int exclusive_scan_dev(sky_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n);

// The SDSC lattice of n x d rows in device memory, kept on the context
// (sdsc.hip), and its copy into CSR arrays in cuboid order.
int sdsc_walk(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d);
int sdsc_copy_out(sky_ctx* c, uint64_t* sky_off, uint32_t* sky_ids, uint64_t sky_cap,
                  uint64_t* ext_off, uint32_t* ext_ids, uint64_t ext_cap);

// The shared buffers of the batched path for groups of up to N points and T
// instances, with a median-mask bucket table of `tab` entries (core.hip).
int batch_buffers(sky_ctx* c, size_t N, size_t T, uint64_t tab);

// ---- stages templated on the dimensionality (stage_single.hip, stage_batch.hip)

// Extended-skyline split of a candidate set over a subspace of D columns:
// leaves, in score order, s1pid[0..m) and status[0..m) (0 sky, 1 ext, 2
// strictly dominated) of the prefilter survivors in the context.
template <int D>
int split_d(sky_ctx* c, const float* rows_dev, uint32_t n_rows, uint32_t d, sky::DimList dims,
            const uint32_t* cand_dev, uint32_t n_cand, int need_dom, uint32_t* m_out);

// Membership of the ne extended-skyline points (rows in c->erows, ids in
// c->epid, output rows in c->pos): grid codes, then the membership launch.
template <int D>
int member_d(sky_ctx* c, uint32_t ne, uint32_t max_d, uint32_t* minkey, bool ties);

// One batch group's whole pass, enqueued on the context stream.
template <int D>
int batch_group_d(sky_ctx* c, BatchGroup& g);

// One chunk of an SDSC lattice level: T cuboids of D columns (subs[t]), each
// over its candidates cand[off[t] .. off[t+1]) (row ids of the n x d rows,
// ascending per cuboid), N candidates in all, prefilter tasks of 256
// positions; leaves keep[i] (sky or ext) and sky[i] per candidate position.
struct SdscChunk {
  const float* rows = nullptr;
  uint32_t d = 0;
  const uint32_t* subs = nullptr;
  const uint32_t* cand = nullptr;
  const uint32_t* off = nullptr;
  const uint2* tasks = nullptr;
  uint32_t ntasks = 0, T = 0, N = 0, max_n = 0;  // max_n: the most candidates of a cuboid
  uint8_t* keep = nullptr;
  uint8_t* sky = nullptr;
  uint32_t* cnt = nullptr;  // kGroupCounters device words ([0] bit 1: NaN seen)
};
template <int D>
int sdsc_chunk_d(sky_ctx* c, const SdscChunk& k);

}  // namespace skyi

#define SKY_DISPATCH_D(DV, ...)                            \
  switch (DV) {                                            \
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
    default: break;                                        \
  }
