// stage_single.hip -- the single-instance stages of one dimensionality
// (compiled once per D = 1..16 with -DSKY_D=D): the extended-skyline split
// of a candidate set over a D-column subspace (split_d) and the membership
// stage of a D-dimensional skycube (member_d).
#include <algorithm>

#include "internal.hpp"
#include "kernels.cuh"

#ifndef SKY_D
#error "compile once per dimensionality with -DSKY_D=1..16"
#endif

namespace skyi {
namespace {

using sky::kBlock;

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
  TRY(sort_pairs(c, P<uint32_t>(c->masks), P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                 P<uint32_t>(c->bidx), m, (int)ds));
  hipLaunchKernelGGL(sky::k_bucket_bounds, dim3(grid_for(nb + 1)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bmask), m, nb, P<uint32_t>(c->boff));
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_bucket_gather, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<uint32_t>(c->bidx), P<uint32_t>(c->s1keys), P<uint2>(c->codes), m,
                     P<uint32_t>(c->bkey), P<uint2>(c->bcode));
  LAUNCHCHK();
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
  TRY(sort_pairs(c, P<uint32_t>(c->masks), P<uint32_t>(c->bmask), P<uint32_t>(c->biota),
                 P<uint32_t>(c->bidx), ne, (int)d));
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
// per-column minimum code of every group of 64, then of every subgroup of 8.
template <int D>
int member_groups(sky_ctx* c, uint32_t ne) {
  const uint32_t ng = (ne + 63) / 64;
  TRY(ensure(c, c->hc_key_a, sizeof(uint64_t) * ne));
  TRY(ensure(c, c->hc_key_b, sizeof(uint64_t) * ne));
  TRY(ensure(c, c->biota, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bidx, sizeof(uint32_t) * ne));
  TRY(ensure(c, c->bcode, sizeof(uint2) * ne));
  TRY(ensure(c, c->gsum, sizeof(uint2) * 9 * ng));  // groups, then their subgroups
  hipLaunchKernelGGL(sky::k_plane_keys<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<uint2>(c->codes), ne, P<uint64_t>(c->hc_key_a), P<uint32_t>(c->biota));
  LAUNCHCHK();
  TRY(sort_pairs(c, P<uint64_t>(c->hc_key_a), P<uint64_t>(c->hc_key_b), P<uint32_t>(c->biota),
                 P<uint32_t>(c->bidx), ne, 4 * D));
  hipLaunchKernelGGL(sky::k_group_sum<D>, dim3(grid_for(ng)), dim3(kBlock), 0, c->stream,
                     P<uint2>(c->codes), P<uint32_t>(c->bidx), ne, P<uint2>(c->bcode),
                     P<uint2>(c->gsum));
  LAUNCHCHK();
  return SKY_OK;
}

// The membership launch of the single-instance path.  Where the kernel runs
// with a narrower boundary table (d = 16), the points it defers follow in a
// second launch with the full table; there, unless the data has ties
// (ties), the first launch also leaves out the tie lattices and defers the
// points that meet a residual tie.  With the tie lattices, one point per
// workgroup (24 KB of lattices, four workgroups per CU) measured faster at
// d = 16 than two (three workgroups per CU): 7.9 vs 8.8 ms on the
// integer-valued n = 1M workload.  The points beyond the narrow table are
// listed before the launch (k_defer_list) and computed on the context's
// second stream beside it, where their few long workgroups hide in the
// main launch; the allocation-free call keeps to one stream and lets the
// kernel list them.
template <int D, int PP, int NT, int O>
int launch_member(sky_ctx* c, uint32_t ne, uint32_t max_d, const uint32_t* minkey,
                  uint32_t cover, sky::CandIndex ix, bool ties) {
  const dim3 grid((ne + PP - 1) / PP), blk(NT);
  constexpr int ZN = sky::narrow_zbits<D>();
  if constexpr (ZN < sky::zbits<D>()) {
    // one point per workgroup for the deferred points: they are few
    const dim3 grid2(std::min<uint32_t>(ne, 512u));
    const bool side = !c->carve;
    const size_t L = (size_t)ne + 1;
    // lists: [0, L) the points beyond the table (allocation-free: every
    // deferred point), [L, 2L) the points that met a residual tie
    TRY(ensure(c, c->defer, sizeof(uint32_t) * (side ? 2 * L : L)));
    uint32_t* pre = P<uint32_t>(c->defer);
    uint32_t* tie = side ? pre + L : pre;
    TRY(fill32(c, pre, 1, 0u));
    if (side) {
      TRY(fill32(c, tie, 1, 0u));
      TRY(side_stream(c));
      hipLaunchKernelGGL(sky::k_defer_list<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                         P<float>(c->erows), ne, minkey, ZN, pre);
      LAUNCHCHK();
      HIPCHK(hipEventRecord(c->side_ev, c->stream));
      HIPCHK(hipStreamWaitEvent(c->copy_stream, c->side_ev, 0));
      hipLaunchKernelGGL((sky::k_member_deferred<D, 1, NT, O>), grid2, blk, 0, c->copy_stream,
                         P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                         (int)max_d, minkey, P<uint32_t>(c->dom), cover, pre, ix);
      LAUNCHCHK();
      HIPCHK(hipEventRecord(c->side_ev, c->copy_stream));
      ix.defer_listed = true;
      ix.defer = ties ? nullptr : tie;
    } else {
      ix.defer = pre;
    }
    if (ties)
      hipLaunchKernelGGL((sky::k_member<D, 1, NT, O, ZN>), dim3(ne), blk, 0, c->stream,
                         P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                         (int)max_d, minkey, P<uint32_t>(c->dom), cover, ix);
    else
      hipLaunchKernelGGL((sky::k_member<D, PP, NT, O, ZN, false>), grid, blk, 0, c->stream,
                         P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                         (int)max_d, minkey, P<uint32_t>(c->dom), cover, ix);
    LAUNCHCHK();
    ix.defer = nullptr;
    ix.defer_listed = false;
    if (!side || !ties) {  // the points the launch deferred, after it
      hipLaunchKernelGGL((sky::k_member_deferred<D, 1, NT, O>), grid2, blk, 0, c->stream,
                         P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                         (int)max_d, minkey, P<uint32_t>(c->dom), cover, tie, ix);
      LAUNCHCHK();
    }
    if (side) HIPCHK(hipStreamWaitEvent(c->stream, c->side_ev, 0));
  } else {
    hipLaunchKernelGGL((sky::k_member<D, PP, NT, O>), grid, blk, 0, c->stream,
                       P<float>(c->erows), P<uint2>(c->codes), P<uint32_t>(c->pos), ne,
                       (int)max_d, minkey, P<uint32_t>(c->dom), cover, ix);
    LAUNCHCHK();
  }
  return SKY_OK;
}

}  // namespace

template <int D>
int split_d(sky_ctx* c, const float* rows_dev, uint32_t n_rows, uint32_t d, sky::DimList dims,
            const uint32_t* cand_dev, uint32_t n_cand, int need_dom, uint32_t* m_out) {
  constexpr uint32_t ds = D;
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
  TRY(fill32(c, c->counter.p, 4, 0u));
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

  TRY(sort_pairs(c, P<uint32_t>(c->key_a), P<uint32_t>(c->key_b), P<uint32_t>(c->pid_a),
                 P<uint32_t>(c->pid_b), nc, 32));

  const uint32_t npr = std::min<uint32_t>(nc, c->tune[ds].pruners);
  hipLaunchKernelGGL(sky::k_gather<D>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream, rows_dev, d,
                     P<uint32_t>(c->pid_b), nc, dims, P<float>(c->srows));
  LAUNCHCHK();
  TRY(mark(c, SKY_STAGE_PRUNE));
  hipLaunchKernelGGL(sky::k_prune<D>, dim3(grid_for(nc)), dim3(kBlock), 0, c->stream,
                     P<float>(c->srows), P<uint32_t>(c->key_b), nc, npr, P<uint8_t>(c->alive));
  LAUNCHCHK();
  TRY(select_flagged(c, P<uint8_t>(c->alive), nc, P<uint32_t>(c->s1idx)));
  uint32_t m = 0;
  TRY(copy_count(c, &m));
  if (m == 0) return fail(c, SKY_ERR_DEVICE, "prefilter removed every point");
  hipLaunchKernelGGL(sky::k_gather_idx<D>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<float>(c->srows), P<uint32_t>(c->key_b), P<uint32_t>(c->pid_b),
                     P<uint32_t>(c->s1idx), m, P<float>(c->s1rows), P<uint32_t>(c->s1keys),
                     P<uint32_t>(c->s1pid));
  LAUNCHCHK();
  TRY(mark(c, SKY_STAGE_SELF));
  // counter[2]: tied sample values of the survivors' columns
  hipLaunchKernelGGL(sky::k_pivots<D>, dim3(D * sky::kPivotSplit), dim3(kBlock), 0, c->stream,
                     P<float>(c->s1rows), m, P<uint32_t>(c->piv), P<uint32_t>(c->counter) + 2);
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_codes<D>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                     P<float>(c->s1rows), m, P<uint32_t>(c->piv), P<uint2>(c->codes),
                     P<uint32_t>(c->masks));
  LAUNCHCHK();
  TRY(self_index(c, ds, m));
  // column minima of the survivors (a point holding one has no strict dominator)
  uint32_t* sminkey = P<uint32_t>(c->counter) + 32;
  if (!need_dom) {
    TRY(fill32(c, sminkey, 16, 0xFFFFFFFFu));
    hipLaunchKernelGGL(sky::k_colmin<D>, dim3(std::min<unsigned>(grid_for(m), 1024u)),
                       dim3(kBlock), 0, c->stream, P<float>(c->s1rows), m, sminkey);
    LAUNCHCHK();
  }
  hipLaunchKernelGGL(sky::k_self_idx<D>, dim3(m), dim3(64), 0, c->stream,
                     P<float>(c->s1rows), P<uint32_t>(c->s1keys), P<uint2>(c->codes),
                     P<uint32_t>(c->boff), P<uint32_t>(c->bidx), P<uint32_t>(c->bkey),
                     P<uint2>(c->bcode), m, need_dom, need_dom ? nullptr : sminkey,
                     P<uint8_t>(c->status));
  LAUNCHCHK();
  *m_out = m;
  return SKY_OK;
}

template <int D>
int member_d(sky_ctx* c, uint32_t ne, uint32_t max_d, uint32_t* minkey, bool ties) {
  hipLaunchKernelGGL(sky::k_colmin<D>, dim3(std::min<unsigned>(grid_for(ne), 1024u)), dim3(kBlock),
                     0, c->stream, P<float>(c->erows), ne, minkey);
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_pivots<D>, dim3(D * sky::kPivotSplit), dim3(kBlock), 0, c->stream,
                     P<float>(c->erows), ne, P<uint32_t>(c->piv), nullptr);
  LAUNCHCHK();
  hipLaunchKernelGGL(sky::k_codes<D>, dim3(grid_for(ne)), dim3(kBlock), 0, c->stream,
                     P<float>(c->erows), ne, P<uint32_t>(c->piv), P<uint2>(c->codes),
                     P<uint32_t>(c->masks));
  LAUNCHCHK();
  uint32_t nbk = 0;
  if (c->opt_member_index == sky::kOrderBucket) TRY(member_index(c, D, ne, &nbk));
  constexpr int PP = sky::member_points<D>();
  constexpr int NT = sky::member_threads<D>();
  const Tuning& tu = c->tune[D];
  int order = c->opt_member_index;
  if (order < 0) order = sky::single_order<D>();
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
  } else if (order == sky::kOrderGroup || order == sky::kOrderSub) {
    TRY(member_groups<D>(c, ne));
    ix.gsum = P<uint2>(c->gsum);
    ix.ssum = P<uint2>(c->gsum) + (ne + 63) / 64;
    ix.gcode = P<uint2>(c->bcode);
    ix.gidx = P<uint32_t>(c->bidx);
  }
  if (!c->carve) {  // the allocation-free call builds no HashCube
    // per-row entry counts of the HashCube export, counted as the rows are
    // stored
    TRY(ensure(c, c->hc_off, sizeof(uint32_t) * ((size_t)ne + 1)));
    TRY(fill32(c, c->hc_off.p, ne, 0u));
    ix.rowcnt = P<uint32_t>(c->hc_off);
  }
  TRY(mark(c, SKY_STAGE_MEMBER));  // the membership kernel launches alone
  if (order == sky::kOrderSub)
    return launch_member<D, PP, NT, sky::kOrderSub>(c, ne, max_d, minkey, cover, ix, ties);
  if (order == sky::kOrderGroup)
    return launch_member<D, PP, NT, sky::kOrderGroup>(c, ne, max_d, minkey, cover, ix, ties);
  if (order == sky::kOrderBucket)
    return launch_member<D, PP, NT, sky::kOrderBucket>(c, ne, max_d, minkey, cover, ix, ties);
  return launch_member<D, PP, NT, sky::kOrderScan>(c, ne, max_d, minkey, cover, ix, ties);
}

template int split_d<SKY_D>(sky_ctx*, const float*, uint32_t, uint32_t, sky::DimList,
                            const uint32_t*, uint32_t, int, uint32_t*);
template int member_d<SKY_D>(sky_ctx*, uint32_t, uint32_t, uint32_t*, bool);

}  // namespace skyi
