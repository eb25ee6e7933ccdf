// core.hip -- C ABI of the MI355X skycube engine (include/skycube_amd.h) and
// the host-side orchestration that does not depend on the dimensionality.
//
// One context = one device + one stream + the device buffers of its calls,
// grown on demand within an HBM budget.  The stages templated on the
// dimensionality live in stage_single.hip and stage_batch.hip.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_run_length_encode.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "batch.cuh"
#include "internal.hpp"
#include "kernels.cuh"
#include "prims.cuh"

using sky::kBlock;

namespace skyi {

int fail(sky_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int ensure(sky_ctx* c, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.cap >= bytes) return SKY_OK;
  if (c->carve) {  // the allocation-free path: a slice of the scratch array
    Carve& k = *c->carve;
    const size_t sz = (bytes + 255) & ~size_t(255);
    if (k.top + sz > k.cap)
      return fail(c, SKY_ERR_CAPACITY, "scratch of %zu bytes exhausted (%zu more needed)", k.cap,
                  k.top + sz - k.cap);
    b.p = k.base + k.top;
    b.cap = sz;
    k.top += sz;
    return SKY_OK;
  }
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

void free_buf(sky_ctx* c, DevBuf& b) {
  if (b.p) {
    (void)hipFree(b.p);
    c->in_use -= b.cap;
  }
  b.p = nullptr;
  b.cap = 0;
}

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

int copy_count(sky_ctx* c, uint32_t* host) {
  HIPCHK(hipMemcpyAsync(host, P<uint32_t>(c->counter), sizeof(uint32_t), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SKY_OK;
}

int select_flagged(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out) {
  rocprim::counting_iterator<uint32_t> it(0);
  size_t bytes = 0;
  HIPCHK(rocprim::select(nullptr, bytes, it, flags, out, P<uint32_t>(c->counter), n, c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(rocprim::select(c->cub_tmp.p, bytes, it, flags, out, P<uint32_t>(c->counter), n,
                         c->stream));
  return SKY_OK;
}

namespace {

template <typename K>
int sort_pairs_t(sky_ctx* c, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout,
                 uint32_t n, int end_bit) {
  size_t bytes = 0;
  HIPCHK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, (unsigned)end_bit,
                                   c->stream));
  TRY(ensure(c, c->cub_tmp, bytes));
  HIPCHK(rocprim::radix_sort_pairs(c->cub_tmp.p, bytes, kin, kout, vin, vout, n, 0u,
                                   (unsigned)end_bit, c->stream));
  return SKY_OK;
}

}  // namespace

int sort_pairs(sky_ctx* c, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
               uint32_t* vout, uint32_t n, int end_bit) {
  return sort_pairs_t(c, kin, kout, vin, vout, n, end_bit);
}

int sort_pairs(sky_ctx* c, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
               uint32_t* vout, uint32_t n, int end_bit) {
  return sort_pairs_t(c, kin, kout, vin, vout, n, end_bit);
}

int fill32(sky_ctx* c, void* p, uint64_t n, uint32_t v) {
  if (!n) return SKY_OK;
  hipLaunchKernelGGL(sky::k_fill_u32, dim3((unsigned)std::min<uint64_t>(grid_for(n), 4096)),
                     dim3(kBlock), 0, c->stream, static_cast<uint32_t*>(p), n, v);
  LAUNCHCHK();
  return SKY_OK;
}

int side_stream(sky_ctx* c) {
  if (!c->copy_stream)
    HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  if (!c->sorted_ev) HIPCHK(hipEventCreateWithFlags(&c->sorted_ev, hipEventDisableTiming));
  if (!c->side_ev) HIPCHK(hipEventCreateWithFlags(&c->side_ev, hipEventDisableTiming));
  return SKY_OK;
}

int fill8(sky_ctx* c, void* p, uint64_t n, uint8_t v) {
  if (!n) return SKY_OK;
  hipLaunchKernelGGL(sky::k_fill_u8, dim3((unsigned)std::min<uint64_t>(grid_for(n), 4096)),
                     dim3(kBlock), 0, c->stream, static_cast<uint8_t*>(p), n, v);
  LAUNCHCHK();
  return SKY_OK;
}

int exclusive_scan_dev(sky_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n) {
  HIPCHK(prim_scan(in, out, n, nullptr, c->batch.tmp.p, c->stream));
  return SKY_OK;
}

int seg_count_scan(sky_ctx* c, const uint8_t* flags, const uint32_t* seg, uint32_t T,
                   uint32_t* cnt, uint32_t* off, uint32_t max_len) {
  TRY(fill32(c, cnt, T + 1, 0));
  if (T) {
    hipLaunchKernelGGL(sky::k_seg_count, dim3(T, sky::seg_slices(max_len)),
                       dim3(sky::kPrimThreads), 0, c->stream, flags, seg, cnt);
    LAUNCHCHK();
  }
  return exclusive_scan_dev(c, cnt, off, T + 1);
}

int select_dev(sky_ctx* c, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t* count) {
  HIPCHK(prim_select(flags, n, out, count, c->batch.tmp.p, c->stream));
  return SKY_OK;
}

}  // namespace skyi

namespace {

using namespace skyi;

int run_split(sky_ctx* c, const float* rows_dev, uint32_t n_rows, uint32_t d, uint32_t sub,
              const uint32_t* cand_dev, uint32_t n_cand, int need_dom, uint32_t* m_out) {
  uint32_t ds = 0;
  const sky::DimList dims = dims_of(sub, &ds);
  int rc = SKY_ERR_ARG;
  SKY_DISPATCH_D(ds, rc = split_d<D>(c, rows_dev, n_rows, d, dims, cand_dev, n_cand, need_dom,
                                     m_out));
  return rc;
}

int check_ctx(sky_ctx* c) {
  if (!c) return SKY_ERR_ARG;
  c->err.clear();
  HIPCHK(hipSetDevice(c->device));
  return SKY_OK;
}

}  // namespace

int skyi::run_mdmc(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d,
                   uint32_t* n_ext) {
  c->has_mdmc = false;
  c->has_hc = false;
  const uint32_t W = sky_words_per_point(d);
  if (max_d == 0 || max_d > d) max_d = d;
  *n_ext = 0;
  c->res_rowcnt = false;
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
  TRY(fill8(c, c->flag.p, n, 0));
  hipLaunchKernelGGL(sky::k_mark_ext, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<uint8_t>(c->status), P<uint32_t>(c->s1pid), m, P<uint8_t>(c->flag));
  LAUNCHCHK();
  TRY(select_flagged(c, P<uint8_t>(c->flag), n, P<uint32_t>(c->ext_pids)));
  // the extended skyline's size, and the tied sample values the split's
  // pivot pass counted (counter[2])
  uint32_t cnt3[3] = {0u, 0u, 0u};
  HIPCHK(hipMemcpyAsync(cnt3, P<uint32_t>(c->counter), sizeof(cnt3), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const uint32_t ne = cnt3[0];
  // more than 1 in 64 sample values tied: the membership keeps its tie
  // lattices in every workgroup
  const bool ties = (uint64_t)cnt3[2] * 64 > (uint64_t)std::min<uint32_t>(m, 2048) * d;
  if (ne >= (1u << 29))
    return fail(c, SKY_ERR_CAPACITY, "extended skyline of %u points (limit 2^29 - 1)", ne);
  TRY(ensure(c, c->erows, sizeof(float) * (size_t)ne * d));
  TRY(ensure(c, c->dom, sizeof(uint32_t) * (size_t)ne * W));
  TRY(ensure(c, c->eidx, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->epid, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->pos, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->codes, sizeof(uint2) * (size_t)ne));
  TRY(ensure(c, c->masks, sizeof(uint32_t) * (size_t)ne));
  TRY(ensure(c, c->piv, sizeof(uint32_t) * 16 * sky::kCodePivots));
  uint32_t* minkey = P<uint32_t>(c->counter) + 16;
  TRY(fill32(c, minkey, 16, 0xFFFFFFFFu));
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
    int rc = SKY_OK;
    SKY_DISPATCH_D(d, {
      hipLaunchKernelGGL(sky::k_gather_ext<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                         P<float>(c->s1rows), P<uint32_t>(c->s1pid), P<uint32_t>(c->eidx), ne,
                         P<float>(c->erows), P<uint32_t>(c->epid));
    });
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_pos, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                       P<uint32_t>(c->epid), ne, P<uint32_t>(c->key_a), P<uint32_t>(c->pos));
    LAUNCHCHK();
    SKY_DISPATCH_D(d, rc = member_d<D>(c, ne, max_d, minkey, ties));
    TRY(rc);
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
  c->res_rowcnt = ne && !c->carve;
  *n_ext = ne;
  return SKY_OK;
}

int skyi::batch_buffers(sky_ctx* c, size_t N, size_t T, uint64_t tab) {
  BatchState& B = c->batch;
  TRY(ensure(c, B.rows, sizeof(float) * N * 16));
  TRY(ensure(c, B.key_a, sizeof(uint64_t) * N));
  TRY(ensure(c, B.key_b, sizeof(uint64_t) * N));
  TRY(ensure(c, B.rs_key, sizeof(uint64_t) * N));
  TRY(ensure(c, B.val_a, sizeof(uint32_t) * N));
  TRY(ensure(c, B.val_b, sizeof(uint32_t) * N));
  TRY(ensure(c, B.rs_val, sizeof(uint32_t) * N));
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
  TRY(ensure(c, B.gsum, sizeof(uint2) * 9 * gsum_cap(N, T)));  // groups, then subgroups
  TRY(ensure(c, B.iboff, sizeof(uint32_t) * tab));
  const size_t tb = prim_tmp_bytes(std::max<uint64_t>({N, T + 1, 1}));
  TRY(ensure(c, B.tmp, tb));
  B.tmp_bytes = tb;
  return SKY_OK;
}

namespace {

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

struct ToU64 {
  __host__ __device__ uint64_t operator()(uint32_t v) const { return v; }
};

// Groups the last MDMC bitmaps into HashCube entries on the device: every
// word that is not all ones becomes an entry (j << 32 | value, point id),
// emitted in point-major order and sorted stably by key, so the ids of a
// group ascend; runs of equal keys are the groups.  The extra always-zero
// word of d >= 5 is one more group (W, 0) holding every point, added when
// the groups are copied out.  The entries of each tile of 4096 words are
// counted by a pass over the bitmaps, or, where the membership kernel counted
// each row's as it stored it (res_rowcnt), taken from those counts.
//
// If pids is given with room for every entry, the ids are copied into it
// on the context's copy stream as soon as they are sorted, while the groups
// are still being formed (*early: the copy was started; the caller waits
// for the copy stream).
int build_hashcube_impl(sky_ctx* c, uint32_t* pids, uint64_t cap_entries, bool* early) {
  const uint32_t ne = c->res_n_ext, W = c->res_W;
  c->hc_groups = c->hc_entries = 0;
  if (ne == 0) {
    c->has_hc = true;
    return SKY_OK;
  }
  const uint64_t words = (uint64_t)ne * W;
  const uint64_t tiles = (words + sky::kHcTile - 1) / sky::kHcTile;
  if (tiles >= 0xFFFFFFFFull) return fail(c, SKY_ERR_CAPACITY, "HashCube export too large");
  int wbits = 0;
  while ((1u << wbits) < W) ++wbits;
  TRY(ensure(c, c->counter, 64 * sizeof(uint32_t)));
  uint32_t* total_dev = P<uint32_t>(c->counter) + 48;
  const uint32_t* toff = nullptr;
  int tshift = 0;
  if (c->res_rowcnt) {
    // exclusive scan of the row counts: a tile's offset is its first row's
    TRY(ensure(c, c->hc_runs, prim_tmp_bytes(ne + 1)));
    HIPCHK(prim_scan(P<uint32_t>(c->hc_off), P<uint32_t>(c->hc_off), ne, total_dev,
                     c->hc_runs.p, c->stream));
    toff = P<uint32_t>(c->hc_off);
    tshift = 12 - wbits;
    c->res_rowcnt = false;  // hc_off holds offsets now
  } else {
    TRY(ensure(c, c->hc_cnt, sizeof(uint32_t) * (tiles + 1)));
    TRY(ensure(c, c->hc_runs, prim_tmp_bytes(tiles + 1)));
    hipLaunchKernelGGL(sky::k_hc_count, dim3((unsigned)tiles), dim3(256), 0, c->stream,
                       P<uint32_t>(c->dom), words, P<uint32_t>(c->hc_cnt));
    LAUNCHCHK();
    HIPCHK(prim_scan(P<uint32_t>(c->hc_cnt), P<uint32_t>(c->hc_cnt), (uint32_t)tiles, total_dev,
                     c->hc_runs.p, c->stream));
    toff = P<uint32_t>(c->hc_cnt);
  }
  uint32_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, total_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const bool extra = c->res_d >= 5;
  if ((uint64_t)total + (extra ? ne : 0) > 0xFFFFFFFFull)
    return fail(c, SKY_ERR_CAPACITY, "HashCube export of more than 2^32 entries");
  TRY(ensure(c, c->hc_key_a, sizeof(uint64_t) * std::max<uint32_t>(total, 1)));
  TRY(ensure(c, c->hc_key_b, sizeof(uint64_t) * std::max<uint32_t>(total, 1)));
  TRY(ensure(c, c->hc_val_a, sizeof(uint32_t) * std::max<uint32_t>(total, 1)));
  TRY(ensure(c, c->hc_val_b, sizeof(uint32_t) * std::max<uint32_t>(total, 1)));
  hipLaunchKernelGGL(sky::k_hc_emit, dim3((unsigned)tiles), dim3(256), 0, c->stream,
                     P<uint32_t>(c->dom), words, wbits, P<uint32_t>(c->ext_pids), toff, tshift,
                     P<uint64_t>(c->hc_key_a), P<uint32_t>(c->hc_val_a));
  LAUNCHCHK();
  uint32_t ng = 0;
  if (total) {
    TRY(sort_pairs(c, P<uint64_t>(c->hc_key_a), P<uint64_t>(c->hc_key_b),
                   P<uint32_t>(c->hc_val_a), P<uint32_t>(c->hc_val_b), total, 32 + wbits));
    if (pids && cap_entries >= (uint64_t)total + (extra ? ne : 0)) {
      TRY(side_stream(c));
      HIPCHK(hipEventRecord(c->sorted_ev, c->stream));
      HIPCHK(hipStreamWaitEvent(c->copy_stream, c->sorted_ev, 0));
      *early = true;
      HIPCHK(hipMemcpyAsync(pids, c->hc_val_b.p, sizeof(uint32_t) * total, hipMemcpyDeviceToHost,
                            c->copy_stream));
      if (extra)  // the extra word's group: every point, ascending
        HIPCHK(hipMemcpyAsync(pids + total, c->ext_pids.p, sizeof(uint32_t) * ne,
                              hipMemcpyDeviceToHost, c->copy_stream));
    }
    // runs of equal keys = groups
    TRY(ensure(c, c->hc_ukey, sizeof(uint64_t) * total));
    uint32_t* run_len = P<uint32_t>(c->hc_val_a);  // reuse: the unsorted values are done
    size_t bytes = 0;
    HIPCHK(rocprim::run_length_encode(nullptr, bytes, P<uint64_t>(c->hc_key_b), total,
                                      P<uint64_t>(c->hc_ukey), run_len, P<uint32_t>(c->counter),
                                      c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(rocprim::run_length_encode(c->cub_tmp.p, bytes, P<uint64_t>(c->hc_key_b), total,
                                      P<uint64_t>(c->hc_ukey), run_len, P<uint32_t>(c->counter),
                                      c->stream));
    TRY(copy_count(c, &ng));
    TRY(ensure(c, c->hc_gw, sizeof(uint32_t) * (ng + 1)));
    TRY(ensure(c, c->hc_gv, sizeof(uint32_t) * (ng + 1)));
    TRY(ensure(c, c->hc_goff, sizeof(uint64_t) * (ng + 2)));
    rocprim::transform_iterator<const uint32_t*, ToU64, uint64_t> it(run_len, ToU64());
    bytes = 0;
    HIPCHK(rocprim::exclusive_scan(nullptr, bytes, it, P<uint64_t>(c->hc_goff), (uint64_t)0, ng,
                                   rocprim::plus<uint64_t>(), c->stream));
    TRY(ensure(c, c->cub_tmp, bytes));
    HIPCHK(rocprim::exclusive_scan(c->cub_tmp.p, bytes, it, P<uint64_t>(c->hc_goff), (uint64_t)0,
                                   ng, rocprim::plus<uint64_t>(), c->stream));
    hipLaunchKernelGGL(sky::k_hc_groups, dim3(grid_for(ng)), dim3(kBlock), 0, c->stream,
                       P<uint64_t>(c->hc_ukey), (uint64_t)ng, P<uint32_t>(c->hc_gw),
                       P<uint32_t>(c->hc_gv));
    LAUNCHCHK();
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hc_groups = ng + (extra ? 1 : 0);
  c->hc_entries = (uint64_t)total + (extra ? ne : 0);
  c->hc_sorted = total;
  c->has_hc = true;
  return SKY_OK;
}

// build_hashcube_impl; an early copy it started has ended if it fails.
int build_hashcube(sky_ctx* c, uint32_t* pids, uint64_t cap_entries, bool* early) {
  *early = false;
  const int rc = build_hashcube_impl(c, pids, cap_entries, early);
  if (rc != SKY_OK && *early) (void)hipStreamSynchronize(c->copy_stream);
  return rc;
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
    DevBuf* gb[] = {&g.d_off, &g.d_src, &g.d_tasks, &g.d_maxd, &g.ext_pids, &g.eoff, &g.dom,
                    &g.cnt};
    for (DevBuf* b : gb) free_buf(c, *b);
  }
  if (B.desc_done) {
    (void)hipEventSynchronize(B.desc_done);
    (void)hipEventDestroy(B.desc_done);
    B.desc_done = nullptr;
  }
  if (B.desc_host) (void)hipHostFree(B.desc_host);
  B.desc_host = nullptr;
  B.desc_cap = 0;
  DevBuf* sb[] = {&B.rows,  &B.key_a,  &B.key_b,  &B.val_a, &B.val_b,  &B.srows,  &B.skeys,
                  &B.spid,  &B.alive,  &B.s1idx,  &B.s1rows, &B.s1keys, &B.s1pid, &B.masks,
                  &B.status, &B.flag_pt, &B.flag_s1, &B.eidx, &B.erows, &B.epid,  &B.pos,
                  &B.codes, &B.rank,   &B.piv,    &B.minkey, &B.seg_a, &B.seg_b, &B.tmp,
                  &B.ikey,  &B.ikey2,  &B.ival,   &B.isort,  &B.ibkey, &B.ibcode, &B.iblocal,
                  &B.ihead, &B.ibmk,   &B.ibstart, &B.ibcnt, &B.ibrun, &B.igfirst, &B.iboff,
                  &B.gcnt,  &B.goff,   &B.gsum,   &B.sminkey, &B.rs_key, &B.rs_val};
  for (DevBuf* b : sb) free_buf(c, *b);
  B.tmp_bytes = 0;
}

// Builds the groups of a batch and every device buffer it needs.
// Builds the batch's groups and uploads their descriptors.  src[t] is the
// device address of instance t's rows.  The descriptors go up from pinned
// staging owned by the context, asynchronously: the host does not wait, and
// the staging is reused once the previous batch's copies have completed.
int batch_setup(sky_ctx* c, const uint64_t* src, uint32_t n_inst, const uint32_t* n,
                const uint32_t* d, const uint32_t* max_d) {
  free_batch(c);
  BatchState& B = c->batch;
  B.n_inst = n_inst;
  // descriptor bytes of every group, 16-byte aligned parts
  auto up16 = [](size_t b) { return (b + 15) & ~size_t(15); };
  uint32_t cnt[17] = {0};
  uint64_t ntask[17] = {0};
  for (uint32_t t = 0; t < n_inst; ++t)
    if (n[t] > 0) {
      cnt[d[t]]++;
      ntask[d[t]] += (n[t] + kBlock - 1) / kBlock;
    }
  size_t stage_bytes = 0;
  for (int D = 1; D <= 16; ++D)
    if (cnt[D])
      stage_bytes += up16(4ull * (cnt[D] + 1)) + up16(8ull * cnt[D]) + up16(8ull * ntask[D]) +
                     up16(4ull * cnt[D]);
  if (B.desc_done) HIPCHK(hipEventSynchronize(B.desc_done));  // staging free again
  else HIPCHK(hipEventCreateWithFlags(&B.desc_done, hipEventDisableTiming));
  if (stage_bytes > B.desc_cap) {
    if (B.desc_host) HIPCHK(hipHostFree(B.desc_host));
    B.desc_host = nullptr;
    B.desc_cap = 0;
    const size_t want = std::max<size_t>(stage_bytes + stage_bytes / 2, 64u << 10);
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&B.desc_host), want, hipHostMallocDefault));
    B.desc_cap = want;
  }
  uint8_t* sp = B.desc_host;
  auto stage = [&](size_t bytes) {
    uint8_t* p = sp;
    sp += up16(bytes);
    return p;
  };
  uint32_t maxN = 0, maxT = 0;
  for (int D = 1; D <= 16; ++D) {
    BatchGroup& g = B.pool[D];
    g.D = D;
    g.inst.clear();
    if (!cnt[D]) continue;
    for (uint32_t t = 0; t < n_inst; ++t)
      if ((int)d[t] == D && n[t] > 0) g.inst.push_back(t);
    g.T = (uint32_t)g.inst.size();
    g.off.assign(g.T + 1, 0);
    g.max_n = 0;
    uint32_t* h_off = reinterpret_cast<uint32_t*>(stage(4ull * (g.T + 1)));
    uint64_t* h_src = reinterpret_cast<uint64_t*>(stage(8ull * g.T));
    uint2* h_tasks = reinterpret_cast<uint2*>(stage(8ull * ntask[D]));
    uint32_t* h_md = reinterpret_cast<uint32_t*>(stage(4ull * g.T));
    uint32_t nt = 0;
    for (uint32_t k = 0; k < g.T; ++k) {
      const uint32_t t = g.inst[k];
      if ((uint64_t)g.off[k] + n[t] > 0xFFFFFFFFull)
        return fail(c, SKY_ERR_ARG, "batch group of d=%d exceeds 2^32 points", D);
      if (n[t] >= (1u << 29))
        return fail(c, SKY_ERR_CAPACITY, "instance %u: %u points (limit 2^29 - 1)", t, n[t]);
      g.off[k + 1] = g.off[k] + n[t];
      g.max_n = std::max(g.max_n, n[t]);
      h_src[k] = src[t];
      h_md[k] = (max_d && max_d[t] && max_d[t] < (uint32_t)D) ? max_d[t] : (uint32_t)D;
      for (uint32_t s0 = 0; s0 < n[t]; s0 += kBlock) h_tasks[nt++] = make_uint2(k, g.off[k] + s0);
    }
    std::copy(g.off.begin(), g.off.end(), h_off);
    g.N = g.off[g.T];
    g.ntasks = nt;
    TRY(ensure(c, g.d_off, sizeof(uint32_t) * (g.T + 1)));
    TRY(ensure(c, g.d_src, sizeof(uint64_t) * g.T));
    TRY(ensure(c, g.d_tasks, sizeof(uint2) * nt));
    TRY(ensure(c, g.d_maxd, sizeof(uint32_t) * g.T));
    TRY(ensure(c, g.ext_pids, sizeof(uint32_t) * g.N));
    TRY(ensure(c, g.eoff, sizeof(uint32_t) * (g.T + 1)));
    TRY(ensure(c, g.dom, sizeof(uint32_t) * (size_t)g.N * sky_words_per_point(D)));
    TRY(ensure(c, g.cnt, sizeof(uint32_t) * kGroupCounters));
    HIPCHK(hipMemcpyAsync(g.d_off.p, h_off, sizeof(uint32_t) * (g.T + 1), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(g.d_src.p, h_src, sizeof(uint64_t) * g.T, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(g.d_tasks.p, h_tasks, sizeof(uint2) * nt, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(g.d_maxd.p, h_md, sizeof(uint32_t) * g.T, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipEventRecord(B.desc_done, c->stream));
    maxN = std::max(maxN, g.N);
    maxT = std::max(maxT, g.T);
    B.groups.push_back(&g);
  }
  uint64_t tab = 1;
  for (const BatchGroup* g : B.groups)
    if (batch_self_index(g->D, g->T)) tab = std::max<uint64_t>(tab, ((uint64_t)g->T << g->D) + 1);
  TRY(batch_buffers(c, maxN, maxT, tab));
  B.sig_src.assign(src, src + n_inst);
  B.sig_tune = c->tune_ver;
  B.sig_n.assign(n, n + n_inst);
  B.sig_d.assign(d, d + n_inst);
  B.sig_m.assign(n_inst, 0);
  if (max_d)
    for (uint32_t t = 0; t < n_inst; ++t) B.sig_m[t] = max_d[t];
  B.ready = true;
  return SKY_OK;
}

int batch_enqueue(sky_ctx* c) {
  for (BatchGroup* g : c->batch.groups) {
    int rc = SKY_OK;
    SKY_DISPATCH_D(g->D, rc = batch_group_d<D>(c, *g));
    if (rc != SKY_OK) return rc;
  }
  return SKY_OK;
}

// A captured pass must hold kernel nodes only: memset nodes have given wrong
// results on replays after the first on these devices, and no step of the
// batched pass is meant to make one.
int check_graph(sky_ctx* c, hipGraph_t graph) {
  size_t count = 0;
  HIPCHK(hipGraphGetNodes(graph, nullptr, &count));
  std::vector<hipGraphNode_t> nodes(count);
  if (count) HIPCHK(hipGraphGetNodes(graph, nodes.data(), &count));
  for (size_t i = 0; i < count; ++i) {
    hipGraphNodeType t;
    HIPCHK(hipGraphNodeGetType(nodes[i], &t));
    if (t != hipGraphNodeTypeKernel)
      return fail(c, SKY_ERR_DEVICE, "captured batch holds a node of type %d (kernels only)",
                  (int)t);
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

int batch_finish(sky_ctx* c, uint64_t* total_ext);

bool batch_same(const BatchState& B, const uint64_t* src, uint32_t n_inst, const uint32_t* n,
                const uint32_t* d, const uint32_t* max_d) {
  if (!B.ready || B.n_inst != n_inst) return false;
  for (uint32_t t = 0; t < n_inst; ++t)
    if (B.sig_src[t] != src[t] || B.sig_n[t] != n[t] || B.sig_d[t] != d[t] ||
        B.sig_m[t] != (max_d ? max_d[t] : 0u))
      return false;
  return true;
}

// Frees every device buffer the context holds; results are gone afterwards.
void free_all(sky_ctx* c) {
  DevBuf* bufs[kSingleBufs];
  single_bufs(c, bufs);
  for (DevBuf* b : bufs) free_buf(c, *b);
  free_batch_buffers(c);
  DevBuf* sd[] = {&c->sd_pool, &c->sd_flag,  &c->sd_lat_off, &c->sd_lat_cnt, &c->sd_lat_sky,
                  &c->sd_subs, &c->sd_src,   &c->sd_pcnt,    &c->sd_coff,    &c->sd_tasks,
                  &c->sd_cand, &c->sd_keep,  &c->sd_skyf,    &c->sd_sel,     &c->sd_kcnt,
                  &c->sd_koff, &c->sd_scnt,  &c->sd_soff,    &c->sd_counters};
  for (DevBuf* b : sd) free_buf(c, *b);
  c->has_mdmc = c->has_hc = c->has_batch = c->has_sdsc = false;
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
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return SKY_ERR_DEVICE;
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : free_b;
  // the allocation-free MDMC's scratch address on this device (once); a
  // failure only disables sky_mdmc_device_noalloc, which then reports it
  (void)noalloc_init_device(device);
  *out = c;
  return SKY_OK;
}

void sky_ctx_destroy(sky_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->copy_stream) {
    (void)hipStreamSynchronize(c->copy_stream);
    (void)hipStreamDestroy(c->copy_stream);
  }
  if (c->sorted_ev) (void)hipEventDestroy(c->sorted_ev);
  if (c->side_ev) (void)hipEventDestroy(c->side_ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int sky_ctx_set_option(sky_ctx* c, int option, int64_t value) {
  if (!c) return SKY_ERR_ARG;
  const int opt = option & 0xFF, dd = option >> 8;
  if (dd < 0 || dd > 16) return fail(c, SKY_ERR_ARG, "option %d: dimensionality %d", option, dd);
  if (opt == SKY_OPT_MEMBER_INDEX) {
    if (value < -1 || value > 3 || dd) return fail(c, SKY_ERR_ARG, "bad member index option");
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
      {SKY_OPT_BATCH_ORDER, 0, 3, &Tuning::order},
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
      {"order", &Tuning::order, {3, 2, 0}},
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
        if (rc == SKY_OK) SKY_DISPATCH_D(g.D, rc = batch_group_d<D>(c, g));
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

void* sky_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void sky_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

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
  bool early = false;  // the ids are already on their way into pids
  if (!c->has_hc) TRY(build_hashcube(c, group_word && group_value && group_off ? pids : nullptr,
                                     cap_entries, &early));
  *n_groups = c->hc_groups;
  *n_entries = c->hc_entries;
  if (early && cap_groups < c->hc_groups)
    HIPCHK(hipStreamSynchronize(c->copy_stream));
  if (!group_word || !group_value || !group_off || !pids) return SKY_OK;
  if (cap_groups < c->hc_groups || cap_entries < c->hc_entries)
    return fail(c, SKY_ERR_CAPACITY, "HashCube needs %llu groups and %llu entries",
                (unsigned long long)c->hc_groups, (unsigned long long)c->hc_entries);
  const uint64_t ng = c->hc_groups, nt = c->hc_entries, ns = c->hc_sorted;
  const uint64_t nsg = ng - (c->res_d >= 5 && ng ? 1 : 0);  // groups from the sort
  if (nsg) {
    HIPCHK(hipMemcpyAsync(group_word, c->hc_gw.p, sizeof(uint32_t) * nsg, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipMemcpyAsync(group_value, c->hc_gv.p, sizeof(uint32_t) * nsg,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(group_off, c->hc_goff.p, sizeof(uint64_t) * nsg,
                          hipMemcpyDeviceToHost, c->stream));
    if (!early)
      HIPCHK(hipMemcpyAsync(pids, c->hc_val_b.p, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost,
                            c->stream));
  }
  if (ng > nsg && !early)  // the extra word's group: every point, ascending
    HIPCHK(hipMemcpyAsync(pids + ns, c->ext_pids.p, sizeof(uint32_t) * c->res_n_ext,
                          hipMemcpyDeviceToHost, c->stream));
  const hipError_t e1 = hipStreamSynchronize(c->stream);
  if (early) HIPCHK(hipStreamSynchronize(c->copy_stream));
  HIPCHK(e1);
  if (ng > nsg) {
    group_word[nsg] = c->res_W;
    group_value[nsg] = 0;
    group_off[nsg] = ns;
  }
  group_off[ng] = nt;
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
  std::vector<uint64_t> src(n_inst);
  uint64_t f = 0;
  for (uint32_t t = 0; t < n_inst; ++t) {
    src[t] = reinterpret_cast<uint64_t>(rows_dev + f);
    f += (uint64_t)n[t] * d[t];
  }
  BatchState& B = c->batch;
  if (!batch_same(B, src.data(), n_inst, n, d, max_d) || B.sig_tune != c->tune_ver)
    TRY(batch_setup(c, src.data(), n_inst, n, d, max_d));
  if (use_graph) {
    if (!B.exec) {
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = batch_enqueue(c);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(c->stream, &graph);
      if (rc != SKY_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess)
        return fail(c, SKY_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
      const int grc = check_graph(c, graph);
      if (grc != SKY_OK) {
        (void)hipGraphDestroy(graph);
        return grc;
      }
      B.graph = graph;
      HIPCHK(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(B.exec, c->stream));
  } else {
    TRY(batch_enqueue(c));
  }
  return batch_finish(c, total_ext);
}

int sky_batch_mdmc_dev(sky_ctx* c, const float* const* inst_rows_dev, uint32_t n_inst,
                       const uint32_t* n, const uint32_t* d, const uint32_t* max_d,
                       uint64_t* total_ext) {
  TRY(check_ctx(c));
  if (!n || !d || !total_ext || (n_inst && !inst_rows_dev))
    return fail(c, SKY_ERR_ARG, "null argument");
  std::vector<uint64_t> src(n_inst);
  for (uint32_t t = 0; t < n_inst; ++t) {
    if (d[t] < 1 || d[t] > 16) return fail(c, SKY_ERR_ARG, "instance %u: d=%u outside [1,16]", t, d[t]);
    if (n[t] && !inst_rows_dev[t]) return fail(c, SKY_ERR_ARG, "instance %u: null rows", t);
    src[t] = reinterpret_cast<uint64_t>(inst_rows_dev[t]);
  }
  c->has_mdmc = false;
  c->has_hc = false;
  c->has_batch = false;
  BatchState& B = c->batch;
  if (!batch_same(B, src.data(), n_inst, n, d, max_d) || B.sig_tune != c->tune_ver)
    TRY(batch_setup(c, src.data(), n_inst, n, d, max_d));
  TRY(batch_enqueue(c));
  return batch_finish(c, total_ext);
}

void* sky_ctx_stream(sky_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

}  // extern "C"

namespace {

// Waits for the batch and reads its counters: the NaN flag and the
// extended-skyline sizes of every group.
int batch_finish(sky_ctx* c, uint64_t* total_ext) {
  BatchState& B = c->batch;
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

}  // namespace

extern "C" {

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
  c->has_hc = false;
  TRY(upload_rows(c, rows, n, d));
  TRY(sdsc_walk(c, P<float>(c->rows), n, d, max_d));
  return sdsc_copy_out(c, sky_off, sky_ids, sky_cap, ext_off, ext_ids, ext_cap);
}

int sky_sdsc_device(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d, uint32_t max_d) {
  TRY(check_ctx(c));
  TRY(check_dims(c, d));
  if (n && !rows_dev) return fail(c, SKY_ERR_ARG, "null argument");
  if (max_d == 0 || max_d > d) max_d = d;
  c->has_mdmc = false;
  c->has_hc = false;
  return sdsc_walk(c, rows_dev, n, d, max_d);
}

int sky_sdsc_fetch(sky_ctx* c, uint64_t* sky_off, uint32_t* sky_ids, uint64_t sky_cap,
                   uint64_t* ext_off, uint32_t* ext_ids, uint64_t ext_cap) {
  TRY(check_ctx(c));
  if (!sky_off || !ext_off) return fail(c, SKY_ERR_ARG, "null argument");
  if (!c->has_sdsc) return fail(c, SKY_ERR_STATE, "no SDSC result is held");
  return sdsc_copy_out(c, sky_off, sky_ids, sky_cap, ext_off, ext_ids, ext_cap);
}

}  // extern "C"
