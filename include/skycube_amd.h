/*
 * skycube_amd.h -- C ABI of the MI355X skycube engine (libskycube_amd.so).
 *
 * Plain C types only: pointers, sizes and status codes.  No exception and no
 * exit() crosses this boundary.  Every function returns SKY_OK (0) or a
 * negative SKY_ERR_* code; sky_last_error() describes the last failure of a
 * context.
 *
 * Semantics (identical to the reference's): points are FP32 rows, smaller is
 * better in every dimension, compared with IEEE ordered compares.  q dominates
 * p in subspace s (a bitmask over the d columns, bit j = column j) iff q_i <=
 * p_i for every i in s and q_i < p_i for at least one i in s; points equal on
 * all of s do not dominate each other.  q strictly dominates p iff q_i < p_i
 * for every i in s.  The skyline sky(s) holds the points no point dominates;
 * the extended skyline ext(s) the points no point strictly dominates.  NaN
 * inputs are rejected with SKY_ERR_NAN.
 *
 * Membership words: a point's bitmap has one bit per subspace s, in word s/32
 * at bit s%32, sky_words_per_point(d) words per point (2^d bits, at least one
 * word).  Bit s is 1 iff the point is NOT in sky(s).  Subspaces outside the
 * requested scope (s = 0, popcount(s) > max_d, and for d < 5 the unused bits
 * s >= 2^d) have no computed skyline and are reported as 1.
 *
 * Contexts: one per (host thread, device).  A context owns every device
 * allocation it makes, within the HBM budget given at creation; results of
 * the last call stay resident on the device until the next call.  Calls on
 * one context must not overlap.
 */
#ifndef SKYCUBE_AMD_H_
#define SKYCUBE_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SKY_OK 0
#define SKY_ERR_ARG (-1)        /* bad argument (d outside [1,16], null, ...) */
#define SKY_ERR_NAN (-2)        /* an input value is NaN */
#define SKY_ERR_OOM_BUDGET (-3) /* the call would exceed the context's HBM budget */
#define SKY_ERR_DEVICE (-4)     /* a HIP runtime call failed */
#define SKY_ERR_CAPACITY (-5)   /* an output buffer is too small */
#define SKY_ERR_STATE (-6)      /* no result of the requested kind is held */

typedef struct sky_ctx sky_ctx;

/* Creates a context on HIP device `device`; hbm_budget_bytes = 0 means the
 * device's free memory.  Replaces the per-device SDSCGpu set-up done by
 * initializeGpu (src/mdmc/MDMC.cu:1060-1084, src/sdsc/SDSC.cu:207-227). */
int sky_ctx_create(int device, size_t hbm_budget_bytes, sky_ctx **out);

/* Frees every device allocation of the context (the reference never frees
 * its GPU state: src/mdmc/MDMC.cu:1087-1108). */
void sky_ctx_destroy(sky_ctx *ctx);

/* Text of the last failure on this context ("" if none).  Valid until the
 * next call on the context. */
const char *sky_last_error(const sky_ctx *ctx);

/* Words per membership bitmap: max(1, 2^d / 32).  The reference's NodeAux
 * bitmap has one extra, always-zero word (src/skycube/NodeAuxilliary.h:31). */
uint32_t sky_words_per_point(uint32_t d);

/* Device memory currently held by the context, in bytes.  Buffers grow on
 * demand and are kept for the next call. */
size_t sky_ctx_bytes_in_use(const sky_ctx *ctx);

/* Page-locked host memory for rows and results (NULL if none is left):
 * copies between it and the device run asynchronously at full link speed,
 * where pageable memory is staged by the runtime.  The reference pins its
 * input before the upload (src/mdmc/MDMC.cu:1068-1071); the skycube driver
 * reads the rows into such a buffer.  Free with sky_host_free. */
void *sky_host_alloc(size_t bytes);
void sky_host_free(void *ptr);

/* Tuning options of a context (results never depend on them).  No
 * reference counterpart (the reference's kernels have fixed launch shapes).
 *   SKY_OPT_MEMBER_INDEX  how the membership kernel of the single-instance
 *                         path visits candidates: 0 in score order; 1
 *                         through median-mask buckets; 2 in groups of 64 of
 *                         the code order, skipped whole by their
 *                         column-minimum codes; 3 the same groups narrowed
 *                         to their live subgroups of 8 first; -1 (default):
 *                         3 up to d = 14, 2 above
 *   SKY_OPT_COVER         lowest-norm candidates covered before the
 *                         coverage pruning starts (64 .. 65536; 0, the
 *                         default: about n_ext / 32 as a power of two in
 *                         [2048, 16384], per instance in a batch)
 *   SKY_OPT_COVER_HEAD    how many of them are tested pair by pair before
 *                         the rest of the cover is code-tested against the
 *                         closure of the head (0: an eighth of the cover)
 *   SKY_OPT_PRUNERS       lowest-norm points of the prefilter (128; 64 .. 8192)
 *   SKY_OPT_MIN_BUCKET    bucket-size threshold above (16; 1 .. 2^30)
 *   SKY_OPT_MEMBER_GRID   workgroups of the batched persistent membership
 *                         kernel (2048; 64 .. 8192)
 *   SKY_OPT_TILE          points per membership workgroup in the batched
 *                         path: 0 the default, 1 half of it
 *   SKY_OPT_BATCH_ORDER   candidate order of the batched membership: 3
 *                         (default) code-order groups and subgroups, 2
 *                         code-order groups, 0 scan or buckets as
 *                         SKY_OPT_MIN_BUCKET decides
 * option | (d << 8) sets the value for dimensionality d only (d = 1..16);
 * a plain option sets it for every d.  sky_ctx_get_option reads one back
 * (for d = 1 when no d is given). */
#define SKY_OPT_MEMBER_INDEX 1
#define SKY_OPT_COVER 2
#define SKY_OPT_PRUNERS 3
#define SKY_OPT_MIN_BUCKET 4
#define SKY_OPT_MEMBER_GRID 5
#define SKY_OPT_TILE 6
#define SKY_OPT_BATCH_ORDER 7
#define SKY_OPT_COVER_HEAD 8
int sky_ctx_set_option(sky_ctx *ctx, int option, int64_t value);
int sky_ctx_get_option(sky_ctx *ctx, int option, int64_t *value);

/* Frees every device buffer the context holds (held results included) and
 * keeps the context usable.  No reference counterpart (the reference keeps
 * its GPU state until exit). */
int sky_ctx_release(sky_ctx *ctx);

/* Device-time breakdown of the last MDMC call, from HIP events recorded on
 * the context's stream while timing is enabled (off by default; it adds a
 * few event records per call).  No reference counterpart: the reference
 * times only the whole Execute() window (src/main.h:324-336). */
#define SKY_STAGE_SCORE_SORT 0 /* Manhattan-norm keys, radix sort, gather */
#define SKY_STAGE_PRUNE 1      /* prefilter against the lowest-score points */
#define SKY_STAGE_SELF 2       /* survivors against each other */
#define SKY_STAGE_COMPACT 3    /* extended skyline in id order, row gather,
                                  column minima, codes, candidate order */
#define SKY_STAGE_MEMBER 4     /* the membership kernel launches alone */
#define SKY_STAGE_COUNT 5
int sky_ctx_set_timing(sky_ctx *ctx, int enable);
int sky_ctx_stage_times(const sky_ctx *ctx, float *ms, int count);

/*
 * MDMC: cuboid membership of every point.  Replaces MDMC<DIMS>::Execute's GPU
 * path -- the full-space extended skyline (MDMC::gpu_extended_skyline ->
 * run_skyalign, src/mdmc/MDMC.cu:504-522), the static-grid build
 * (MDMC::build_tree_gpu, src/mdmc/MDMC.cu:664-750) and the per-point bitmap
 * kernel loop (MDMC::tombstoning_hearting_gpu, src/mdmc/MDMC.cu:752-857).
 *
 * rows: n x d row-major FP32 on the host (borrowed, not kept).  max_d caps
 * the popcount of the subspaces computed (the -m flag; max_d >= d means all).
 * Only extended-skyline points can be in any cuboid skyline; they are
 * returned in ext_pids (capacity n) in ascending order, their count in
 * *n_ext.  dom_words (capacity dom_cap_words) receives n_ext bitmaps in the
 * same order; if it is NULL or too small, the bitmaps stay on the device and
 * sky_mdmc_fetch copies them later (the call then returns SKY_OK when
 * dom_words is NULL and SKY_ERR_CAPACITY when it is too small).
 */
int sky_mdmc(sky_ctx *ctx, const float *rows, uint32_t n, uint32_t d,
             uint32_t max_d, uint32_t *ext_pids, uint32_t *n_ext,
             uint32_t *dom_words, uint64_t dom_cap_words);

/* Same as sky_mdmc with rows already in device memory (a device pointer) and
 * every result left on the device: nothing is copied to the host except the
 * count.  The stream is the context's own. */
int sky_mdmc_device(sky_ctx *ctx, const float *rows_dev, uint32_t n, uint32_t d,
                    uint32_t max_d, uint32_t *n_ext);

/* sky_mdmc_device for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No reference counterpart.  It calls no hipMalloc, hipFree or allocating
 * library routine: the caller passes the rows and the outputs in device
 * memory -- ext_pids_dev (capacity n ids) and dom_words_dev (capacity
 * dom_cap_words >= n * sky_words_per_point(d)) receive the same ids and
 * bitmaps sky_mdmc_device would leave on the device, *n_ext their count --
 * and every other buffer is carved from a 64 MiB scratch array that the
 * library holds on each device until the process ends.  Sizes with
 * n * (138 + 12 d) + 2^(d+2) + 2 MiB above 64 MiB are rejected with
 * SKY_ERR_CAPACITY (at d = 16 about 196,000 points, at d = 8 about 278,000).
 * The context supplies only its stream; its own buffers and held results are
 * left alone, and no result is held afterwards (sky_mdmc_fetch / _hashcube
 * report SKY_ERR_STATE).  All calls on one device share its scratch array:
 * calls that overlap on one device (from several contexts or threads) race. */
int sky_mdmc_device_noalloc(sky_ctx *ctx, const float *rows_dev, uint32_t n, uint32_t d,
                            uint32_t max_d, uint32_t *ext_pids_dev,
                            uint32_t *dom_words_dev, uint64_t dom_cap_words,
                            uint32_t *n_ext);

/* Copies the bitmaps (and ext pids, if ext_pids is not NULL) of the last
 * sky_mdmc / sky_mdmc_device call to host memory. */
int sky_mdmc_fetch(sky_ctx *ctx, uint32_t *ext_pids, uint32_t *dom_words,
                   uint64_t dom_cap_words);

/* Device pointers of the last MDMC result (ext pids, n_ext words each; the
 * bitmaps, n_ext * sky_words_per_point(d) words).  Owned by the context. */
int sky_mdmc_device_results(sky_ctx *ctx, const uint32_t **ext_pids_dev,
                            const uint32_t **dom_words_dev);

/*
 * HashCube of the last MDMC result: the reference's result structure
 * (HashCube::populateFromSingleNodeAux, src/skycube/Hashcube.h:107-115, built
 * by MDMC::build_hashcube, src/mdmc/MDMC.cu:648-661) -- for every word index
 * j, a map from word value v to the ids of the points whose bitmap word j is
 * v, leaving out values with all 32 bits set.  Like the reference's NodeAux
 * bitmap, a point has sky_words_per_point(d) + 1 words when d >= 5, the last
 * one always 0.  Grouped on the device (sort by (j, v)); returned as groups
 * ascending by (j, v): group_word[g], group_value[g], and the ids
 * pids[group_off[g] .. group_off[g+1]) ascending.  Call with NULL arrays (or
 * too small capacities) to learn *n_groups and *n_entries first; the grouping
 * is kept on the device until the next call.  Arrays given to the first call
 * with room enough receive the ids while the groups are still being formed
 * (the copy overlaps the grouping); with too little room the call fails with
 * SKY_ERR_CAPACITY, and the sizes are set.
 */
int sky_mdmc_hashcube(sky_ctx *ctx, uint64_t *n_groups, uint64_t *n_entries,
                      uint32_t *group_word, uint32_t *group_value,
                      uint64_t *group_off, uint32_t *pids, uint64_t cap_groups,
                      uint64_t cap_entries);

/*
 * Batched MDMC: n_inst independent skycubes in one pass.  Instance t has
 * n[t] rows of d[t] columns (1..16) and subspace cap max_d[t] (0 = d[t]).
 * rows holds the instances' row-major matrices concatenated in instance
 * order, on the host (rows_on_device = 0) or in device memory (1).
 *
 * Instances of equal d form a group; every stage runs on all instances of a
 * group in one launch, and the membership stage is one persistent launch
 * per group fed by a device work queue, so the launch count does not grow
 * with n_inst and no host synchronisation happens between stages.  With
 * use_graph != 0 the whole pass is captured in a HIP graph on first use and
 * replayed while the batch (pointer, counts, dimensions, caps) is the same.
 * *total_ext receives the number of extended-skyline points of all
 * instances.  No reference counterpart: the reference computes one skycube
 * per process (src/main.h:152-356).
 */
int sky_batch_mdmc(sky_ctx *ctx, const float *rows, int rows_on_device,
                   uint32_t n_inst, const uint32_t *n, const uint32_t *d,
                   const uint32_t *max_d, int use_graph, uint64_t *total_ext);

/* sky_batch_mdmc over instances in separate device allocations:
 * inst_rows_dev[t] holds instance t's n[t] x d[t] FP32 rows in device
 * memory.  The pass is enqueued on the context's stream (sky_ctx_stream)
 * behind the work already there, so a caller that uploads rows on another
 * stream makes that stream's work visible first (hipStreamWaitEvent on the
 * context's stream).  No graph; descriptors go up asynchronously from
 * pinned staging, and the only host wait is the one that reads the result
 * counts.  Results as for sky_batch_mdmc.  No reference counterpart. */
int sky_batch_mdmc_dev(sky_ctx *ctx, const float *const *inst_rows_dev,
                       uint32_t n_inst, const uint32_t *n, const uint32_t *d,
                       const uint32_t *max_d, uint64_t *total_ext);

/* The HIP stream (a hipStream_t) that every call on the context enqueues
 * its work on, for callers that order their own work with it.  (Calls may
 * also run work on a second stream of the context -- the d = 16
 * membership's deferred points, the HashCube's early copy of the ids --
 * whose work is joined to this stream, or waited for, before the call
 * returns.) */
void *sky_ctx_stream(sky_ctx *ctx);

/* Results of the last batch, in instance order: ext_off and dom_off
 * (n_inst + 1 entries each) delimit instance t's extended-skyline row ids
 * (local, ascending) in ext_pids and its bitmaps (sky_words_per_point(d[t])
 * words per point) in dom_words.  With NULL ext_pids / dom_words only the
 * offsets are written. */
int sky_batch_fetch(sky_ctx *ctx, uint64_t *ext_off, uint32_t *ext_pids,
                    uint64_t *dom_off, uint32_t *dom_words, uint64_t cap_ext,
                    uint64_t cap_words);

/* Autotuner of the batched path: runs the batch (as sky_batch_mdmc would)
 * once per candidate value of each setting in turn -- candidate order,
 * tile (SKY_OPT_TILE), cover depth, prefilter size, bucket threshold,
 * persistent grid -- timing
 * every dimensionality group with HIP events, and keeps for each group's d
 * the fastest value (coordinate descent, about 21 passes; tune on a sample
 * of a large batch).  The chosen settings stay on the context; report (if
 * not NULL) receives them as JSON text.  The last batch's results are not
 * kept.  No reference counterpart. */
int sky_batch_autotune(sky_ctx *ctx, const float *rows, int rows_on_device,
                       uint32_t n_inst, const uint32_t *n, const uint32_t *d,
                       const uint32_t *max_d, char *report, size_t report_cap);

/* Results of the last batch into per-instance buffers: instance t's ids to
 * ext_dst[t] and its bitmaps to dom_dst[t] (sized from sky_batch_fetch's
 * offsets; a NULL pointer skips that array of the instance).  Copies go
 * straight from device memory, at full link speed into pinned memory. */
int sky_batch_fetch_into(sky_ctx *ctx, uint32_t *const *ext_dst,
                         uint32_t *const *dom_dst);

/*
 * One-cuboid SkyAlign: sky(s) and ext(s) \ sky(s) over a candidate set.
 * Replaces run_skyalign(full_dims, n, dimension_set, sky, ext, working_set,
 * conf) (src/sdsc/skyAlign.cu:9-53, src/sdsc/skyAlign.cuh:321-322).
 *
 * rows: n_rows x d FP32 on the host.  subspace: non-zero mask over the d
 * columns (the reference's dimension_set).  cand: candidate row ids (the
 * reference's working_set), or NULL for all rows.  Only candidates are
 * compared with each other.  sky_out / ext_out (capacity n_cand, or n_rows
 * when cand is NULL) receive the row ids of sky(s) and of ext(s) \ sky(s) in
 * ascending order.  Strictly dominated candidates are in neither.
 */
int sky_skyalign(sky_ctx *ctx, const float *rows, uint32_t n_rows, uint32_t d,
                 uint32_t subspace, const uint32_t *cand, uint32_t n_cand,
                 uint32_t *sky_out, uint32_t *n_sky, uint32_t *ext_out,
                 uint32_t *n_ext);

/*
 * SDSC: the full lattice of cuboid skylines, top-down.  Replaces
 * SDSC<DIMS>::Execute's GPU path (src/sdsc/SDSC.cu:248-335): every cuboid s
 * with popcount(s) <= max_d (and the full space) is a SkyAlign over the
 * smallest parent's sky+ext (src/sdsc/SDSC.cu:81-109,365-376).
 *
 * Results, per cuboid s in 1 .. 2^d - 1, as CSR arrays: sky_off and ext_off
 * (2^d + 1 entries each; entry 0 and skipped cuboids are empty ranges),
 * sky_ids / ext_ids (capacities sky_cap / ext_cap) with ascending row ids per
 * cuboid.  If an id array is too small (or NULL) the call returns
 * SKY_ERR_CAPACITY with the offsets holding the sizes needed; the lattice
 * stays on the device either way, for sky_sdsc_fetch.
 *
 * The walk runs on the device one lattice level at a time (the reference's
 * level barrier): every cuboid of a level is one instance of a batched
 * SkyAlign pass over its parent's list, and the lists stay on the device.
 */
int sky_sdsc(sky_ctx *ctx, const float *rows, uint32_t n, uint32_t d,
             uint32_t max_d, uint64_t *sky_off, uint32_t *sky_ids,
             uint64_t sky_cap, uint64_t *ext_off, uint32_t *ext_ids,
             uint64_t ext_cap);

/* sky_sdsc on rows already in device memory; the lattice stays on the
 * device for sky_sdsc_fetch. */
int sky_sdsc_device(sky_ctx *ctx, const float *rows_dev, uint32_t n, uint32_t d,
                    uint32_t max_d);

/* Copies the lattice of the last sky_sdsc / sky_sdsc_device call into CSR
 * arrays, as sky_sdsc returns them (SDSC<DIMS>::getLattice,
 * src/sdsc/SDSC.h:18-54). */
int sky_sdsc_fetch(sky_ctx *ctx, uint64_t *sky_off, uint32_t *sky_ids,
                   uint64_t sky_cap, uint64_t *ext_off, uint32_t *ext_ids,
                   uint64_t ext_cap);

/*
 * Skycube server: continuous batching of independent skycube requests on one
 * device.  No reference counterpart -- the reference computes one skycube per
 * process (src/main.h:152-356); this serves many callers from one process.
 *
 * sky_server_submit queues a request (the rows are copied; NaN input and a
 * request whose device footprint exceeds one worker's share of the budget are
 * rejected at once with SKY_ERR_NAN / SKY_ERR_OOM_BUDGET) and enqueues the
 * upload of its rows at once, so that they wait resident in device memory
 * (a stream-ordered pool holding up to 1/16 of the budget; past that, the
 * worker uploads them when it takes the request).  `workers` threads, each
 * with its own context and stream, repeatedly take every queued request,
 * first come first served, up to `max_batch` requests and an equal share of
 * the rest of the budget of estimated device memory, and compute them as
 * one batched MDMC pass over the resident rows (sky_batch_mdmc_dev), so that
 * requests arriving while the device is busy join the next batch and one
 * worker's host work overlaps another's kernels.  Every request's result is
 * identical to sky_mdmc on the same rows alone.  hbm_budget_bytes = 0 means
 * 80% of the device's free memory.
 */
typedef struct sky_server sky_server;

int sky_server_create(int device, size_t hbm_budget_bytes, uint32_t workers,
                      uint32_t max_batch, sky_server **out);

/* Waits for every queued request, then stops the workers and frees all. */
void sky_server_destroy(sky_server *srv);

/* Estimated device bytes of one request of n points of d dimensions (an
 * upper bound of what a batched pass allocates for it). */
uint64_t sky_server_footprint(uint32_t n, uint32_t d);

/* Queues n x d rows (FP32, row-major, host memory) with subspaces of at most
 * max_d dimensions (0 = d); *id receives the request id. */
int sky_server_submit(sky_server *srv, const float *rows, uint32_t n,
                      uint32_t d, uint32_t max_d, uint64_t *id);

/* Blocks until request id is done (timeout_ms < 0: no limit); returns its
 * status, or SKY_ERR_STATE on time-out or an unknown id. */
int sky_server_wait(sky_server *srv, uint64_t id, int timeout_ms);

/* Result of a finished request, as sky_mdmc returns it: *n_ext extended-
 * skyline row ids (ascending) and their bitmaps (n_ext *
 * sky_words_per_point(d) words, cap_words available).  With NULL ext_pids or
 * dom_words only the sizes, latency (submit to result ready, ms) and batch
 * size are written and the result is kept; otherwise it is copied out and
 * released. */
int sky_server_result(sky_server *srv, uint64_t id, uint32_t *ext_pids,
                      uint32_t *n_ext, uint32_t *dom_words, uint64_t cap_words,
                      double *latency_ms, uint32_t *batch_size);

/* Batches run so far, the most batches that were in flight at once, requests
 * waiting, and the wall time (ms) during which at least one batch was in
 * flight (upload, kernels and download). */
int sky_server_stats(sky_server *srv, uint64_t *batches,
                     uint64_t *max_concurrent, uint64_t *queued,
                     double *busy_ms);

#ifdef __cplusplus
}
#endif

#endif /* SKYCUBE_AMD_H_ */
