// stage_batch.hip -- the batched pass of one dimensionality group (compiled
// once per D = 1..16 with -DSKY_D=D).  Every step is one launch over all
// instances of the group; counts only the device knows stay on the device;
// no step allocates, synchronises with the host or issues a memset, so a
// whole pass can be captured in a HIP graph.
#include <algorithm>

#include "batch.cuh"
#include "internal.hpp"
#include "prims.cuh"

#ifndef SKY_D
#error "compile once per dimensionality with -DSKY_D=1..16"
#endif

namespace skyi {
namespace {

using sky::kBlock;

int tbits_of(uint32_t T) {
  int tb = 0;
  while ((1u << tb) < T) ++tb;
  return tb;
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
  HIPCHK(prim_sort_pairs(P<const uint32_t>(B.ikey), P<uint32_t>(B.ikey2), P<const uint32_t>(B.ival),
                         P<uint32_t>(B.isort), P<uint32_t>(B.rs_key), P<uint32_t>(B.rs_val), N,
                         std::min(32, tbits_of(T) + D), B.tmp.p, c->stream));
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
  TRY(exclusive_scan_dev(c, P<uint32_t>(B.seg_a), P<uint32_t>(B.seg_b), T + 1));
  constexpr int NT = sky::member_threads<D>();
  if (tu.order == 2 || tu.order == 3) {  // code-order groups (or subgroups) for every instance
    auto kern = tu.order == 3 ? sky::k_bmember<D, PP, NT, sky::kOrderSub>
                              : sky::k_bmember<D, PP, NT, sky::kOrderGroup>;
    hipLaunchKernelGGL(kern, dim3(tu.grid), dim3(NT), 0, c->stream,
                       P<float>(B.erows), P<uint2>(B.codes), P<uint32_t>(B.pos), eoff, T,
                       P<uint32_t>(B.seg_b), P<uint32_t>(g.d_maxd), P<uint32_t>(B.minkey),
                       P<uint32_t>(g.dom), cnt + 4, tu.cover, tu.head,
                       tu.min_bucket, P<uint32_t>(B.goff), P<uint2>(B.gsum),
                       P<uint2>(B.gsum) + gsum_cap(g.N, T), P<uint2>(B.ibcode),
                       P<uint32_t>(B.iblocal), nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
  } else {
    hipLaunchKernelGGL((sky::k_bmember<D, PP, NT>), dim3(tu.grid), dim3(NT), 0, c->stream,
                       P<float>(B.erows), P<uint2>(B.codes), P<uint32_t>(B.pos), eoff, T,
                       P<uint32_t>(B.seg_b), P<uint32_t>(g.d_maxd), P<uint32_t>(B.minkey),
                       P<uint32_t>(g.dom), cnt + 4, tu.cover, tu.head,
                       tu.min_bucket, nullptr, P<uint2>(B.gsum), nullptr, P<uint2>(B.ibcode),
                       P<uint32_t>(B.iblocal), P<uint32_t>(B.igfirst), P<uint32_t>(B.ibmk),
                       P<uint32_t>(B.ibstart), P<uint32_t>(B.ibcnt), P<uint32_t>(B.iblocal),
                       P<uint2>(B.ibcode));
  }
  LAUNCHCHK();
  return SKY_OK;
}

}  // namespace

// Enqueues one group's whole pass on the context stream.
template <int D>
int batch_group_d(sky_ctx* c, BatchGroup& g) {
  BatchState& B = c->batch;
  const uint32_t N = g.N, T = g.T;
  uint32_t* cnt = P<uint32_t>(g.cnt);  // [0] err [1] survivors [2] ext ids [3] ext score [4] work
  const uint32_t* off = P<uint32_t>(g.d_off);
  const unsigned gN = grid_for(N);
  const int tbits = tbits_of(T);
  TRY(fill32(c, cnt, kGroupCounters, 0));
  TRY(fill8(c, B.flag_pt.p, N, 0));
  hipLaunchKernelGGL(sky::k_bcopy<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<uint64_t>(g.d_src), off, T, N, P<float>(B.rows));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bscore<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.rows), off,
                     T, N, P<uint64_t>(B.key_a), P<uint32_t>(B.val_a), cnt);
  LAUNCHCHK();
  HIPCHK(prim_sort_pairs(P<const uint64_t>(B.key_a), P<uint64_t>(B.key_b), P<const uint32_t>(B.val_a),
                         P<uint32_t>(B.val_b), P<uint64_t>(B.rs_key), P<uint32_t>(B.rs_val), N,
                         32 + tbits, B.tmp.p, c->stream));
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
  TRY(seg_count_scan(c, P<uint8_t>(B.alive), off, T, P<uint32_t>(B.seg_a), s1off, g.max_n));
  hipLaunchKernelGGL(sky::k_bgather_idx<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<float>(B.srows), P<uint32_t>(B.skeys), P<uint32_t>(B.spid),
                     P<uint32_t>(B.s1idx), cnt + 1, P<float>(B.s1rows), P<uint32_t>(B.s1keys),
                     P<uint32_t>(B.s1pid));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpivots<D>, dim3(T * D), dim3(kBlock), 0, c->stream,
                     P<float>(B.s1rows), s1off, T, P<uint32_t>(B.piv));
  LAUNCHCHK();
  const dim3 gcol(T, sky::seg_slices(g.max_n));  // column minima, per instance
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
    TRY(fill32(c, P<uint32_t>(B.sminkey), (uint64_t)T * 16, 0xFFFFFFFFu));
    hipLaunchKernelGGL(sky::k_bcolmin<D>, gcol, dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, P<uint32_t>(B.sminkey));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bself_idx<D>, dim3((N + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint2>(B.codes), s1off, T,
                       cnt + 1, P<uint32_t>(B.iboff), P<uint32_t>(B.isort), P<uint32_t>(B.ibkey),
                       P<uint2>(B.ibcode), P<uint32_t>(B.sminkey), 0, P<uint8_t>(B.status));
    LAUNCHCHK();
  } else {
    hipLaunchKernelGGL(sky::k_bmasks<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, T, cnt + 1, P<uint32_t>(B.piv),
                       P<uint32_t>(B.masks));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bself<D>, dim3((N + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint32_t>(B.masks), s1off, T,
                       cnt + 1, 0, P<uint8_t>(B.status));
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(sky::k_bext_flags, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<uint8_t>(B.status), P<uint32_t>(B.s1pid), cnt + 1, N,
                     P<uint8_t>(B.flag_pt), P<uint8_t>(B.flag_s1));
  LAUNCHCHK();
  // extended skyline in id order (the result) and in score order
  TRY(select_dev(c, P<uint8_t>(B.flag_pt), N, P<uint32_t>(g.ext_pids), cnt + 2));
  uint32_t* eoff = P<uint32_t>(g.eoff);
  TRY(seg_count_scan(c, P<uint8_t>(B.flag_pt), off, T, P<uint32_t>(B.seg_a), eoff, g.max_n));
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
  TRY(fill32(c, P<uint32_t>(B.minkey), (uint64_t)T * 16, 0xFFFFFFFFu));
  hipLaunchKernelGGL(sky::k_bcolmin<D>, gcol, dim3(kBlock), 0, c->stream, P<float>(B.erows),
                     eoff, P<uint32_t>(B.minkey));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpivots<D>, dim3(T * D), dim3(kBlock), 0, c->stream,
                     P<float>(B.erows), eoff, T, P<uint32_t>(B.piv));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bcodes<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.erows),
                     eoff, T, cnt + 3, P<uint32_t>(B.piv), P<uint2>(B.codes));
  LAUNCHCHK();
  if (tu.order == 2 || tu.order == 3) {
    // code order per instance, groups of 64 with their summaries: one sort
    // of (instance, code key) -- the instance in the top bits, the lowest
    // code-key bits dropped where both do not fit 64 bits (any order is
    // correct; the code order only makes the group summaries tight)
    const int kbits = std::min(4 * D, 64 - tbits);
    hipLaunchKernelGGL(sky::k_bplane_keys<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<uint2>(B.codes), eoff, T, cnt + 3, N, kbits, P<uint64_t>(B.key_a),
                       P<uint32_t>(B.ival));
    LAUNCHCHK();
    HIPCHK(prim_sort_pairs(P<const uint64_t>(B.key_a), P<uint64_t>(B.key_b),
                           P<const uint32_t>(B.ival), P<uint32_t>(B.isort), P<uint64_t>(B.rs_key),
                           P<uint32_t>(B.rs_val), N, kbits + tbits, B.tmp.p, c->stream));
    hipLaunchKernelGGL(sky::k_bgroup_counts, dim3(grid_for(T + 1)), dim3(kBlock), 0, c->stream,
                       eoff, T, P<uint32_t>(B.gcnt));
    LAUNCHCHK();
    TRY(exclusive_scan_dev(c, P<uint32_t>(B.gcnt), P<uint32_t>(B.goff), T + 1));
    hipLaunchKernelGGL(sky::k_bgroup_sum<D>, dim3(grid_for(N / 64 + T + 1)), dim3(kBlock), 0,
                       c->stream, P<uint2>(B.codes), P<uint32_t>(B.isort), eoff, T,
                       P<uint32_t>(B.goff), P<uint2>(B.ibcode), P<uint32_t>(B.iblocal),
                       P<uint2>(B.gsum), P<uint2>(B.gsum) + gsum_cap(N, T));
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
  if (tu.tile == 0) return batch_member<D, sky::member_points<D>()>(c, g, eoff, tu);
  return batch_member<D, sky::member_points_alt<D>()>(c, g, eoff, tu);
}

// One chunk of an SDSC lattice level: the T cuboids of D columns in k, each
// a batched instance over its candidates, through the batched split with
// dominance told apart; leaves keep / sky flags per candidate position.
template <int D>
int sdsc_chunk_d(sky_ctx* c, const SdscChunk& k) {
  BatchState& B = c->batch;
  const uint32_t N = k.N, T = k.T;
  uint32_t* cnt = k.cnt;  // [0] err [1] survivors
  const uint32_t* off = k.off;
  const unsigned gN = grid_for(N);
  TRY(fill32(c, cnt, kGroupCounters, 0));
  TRY(fill8(c, k.keep, N, 0));
  TRY(fill8(c, k.sky, N, 0));
  hipLaunchKernelGGL(sky::k_sdsc_rows<D>, dim3(gN), dim3(kBlock), 0, c->stream, k.rows, k.d,
                     k.cand, k.subs, off, T, N, P<float>(B.rows));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bscore<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.rows), off,
                     T, N, P<uint64_t>(B.key_a), P<uint32_t>(B.val_a), cnt);
  LAUNCHCHK();
  HIPCHK(prim_sort_pairs(P<const uint64_t>(B.key_a), P<uint64_t>(B.key_b), P<const uint32_t>(B.val_a),
                         P<uint32_t>(B.val_b), P<uint64_t>(B.rs_key), P<uint32_t>(B.rs_val), N,
                         32 + tbits_of(T), B.tmp.p, c->stream));
  hipLaunchKernelGGL(sky::k_bgather<D>, dim3(gN), dim3(kBlock), 0, c->stream, P<float>(B.rows),
                     P<uint64_t>(B.key_b), P<uint32_t>(B.val_b), N, P<float>(B.srows),
                     P<uint32_t>(B.skeys), P<uint32_t>(B.spid));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bprune<D>, dim3(k.ntasks), dim3(kBlock), 0, c->stream,
                     P<float>(B.srows), P<uint32_t>(B.skeys), off, k.tasks, c->tune[D].pruners,
                     P<uint8_t>(B.alive));
  LAUNCHCHK();
  TRY(select_dev(c, P<uint8_t>(B.alive), N, P<uint32_t>(B.s1idx), cnt + 1));
  uint32_t* s1off = P<uint32_t>(B.seg_b);
  TRY(seg_count_scan(c, P<uint8_t>(B.alive), off, T, P<uint32_t>(B.seg_a), s1off, k.max_n));
  hipLaunchKernelGGL(sky::k_bgather_idx<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<float>(B.srows), P<uint32_t>(B.skeys), P<uint32_t>(B.spid),
                     P<uint32_t>(B.s1idx), cnt + 1, P<float>(B.s1rows), P<uint32_t>(B.s1keys),
                     P<uint32_t>(B.s1pid));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bpivots<D>, dim3(T * D), dim3(kBlock), 0, c->stream,
                     P<float>(B.s1rows), s1off, T, P<uint32_t>(B.piv));
  LAUNCHCHK();
  if (batch_self_index(D, T)) {
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
    // one wave per workgroup and per survivor (the dispatcher balances the
    // uneven survivors; a grid with fewer workgroups that stride measured
    // 15% slower), up to 2^26 - 1 workgroups, beyond which the waves stride
    hipLaunchKernelGGL((sky::k_bself_idx<D, 64>), dim3(std::min<uint32_t>(N, (1u << 26) - 1u)),
                       dim3(64),
                       0, c->stream, P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint2>(B.codes),
                       s1off, T, cnt + 1, P<uint32_t>(B.iboff), P<uint32_t>(B.isort),
                       P<uint32_t>(B.ibkey), P<uint2>(B.ibcode), nullptr, 1,
                       P<uint8_t>(B.status));
    LAUNCHCHK();
  } else {
    hipLaunchKernelGGL(sky::k_bmasks<D>, dim3(gN), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), s1off, T, cnt + 1, P<uint32_t>(B.piv),
                       P<uint32_t>(B.masks));
    LAUNCHCHK();
    hipLaunchKernelGGL(sky::k_bself<D>, dim3((N + 3) / 4), dim3(kBlock), 0, c->stream,
                       P<float>(B.s1rows), P<uint32_t>(B.s1keys), P<uint32_t>(B.masks), s1off, T,
                       cnt + 1, 1, P<uint8_t>(B.status));
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(sky::k_sdsc_flags, dim3(gN), dim3(kBlock), 0, c->stream,
                     P<uint8_t>(B.status), P<uint32_t>(B.s1pid), cnt + 1, k.keep, k.sky);
  LAUNCHCHK();
  return SKY_OK;
}

template int batch_group_d<SKY_D>(sky_ctx*, BatchGroup&);
template int sdsc_chunk_d<SKY_D>(sky_ctx*, const SdscChunk&);

}  // namespace skyi
