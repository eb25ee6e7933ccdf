// server.cpp -- continuous-batching skycube server over the C ABI.
//
// Requests (one skycube each: rows, d, max_d) are queued by sky_server_submit
// from any thread.  At submission the rows are copied into pinned host
// memory and their upload to device memory is enqueued at once, on the
// server's upload stream, into a stream-ordered memory pool: a request's rows
// are resident in HBM while it waits, and the copy overlaps whatever the
// device is running.  Worker threads, each with its own device context and
// stream, repeatedly take every queued request that fits the HBM budget (the
// estimated device footprint of the requests in flight, first come first
// served) and run them as one batched MDMC pass over the resident rows
// (sky_batch_mdmc_dev): the worker's stream waits on the requests' upload
// events, the batch descriptors go up asynchronously, and the only host
// waits are the ones that read the result sizes and the results.  A request
// that arrives while the device is busy joins the next batch, and one
// worker's host work overlaps the others' kernels.  Each worker owns an
// equal share of the budget (less the part reserved for resident rows) and
// keeps its device buffers across batches.  Results come down straight from
// the batch's device buffers into pinned host memory owned by the request
// (an arena of pinned slabs) until sky_server_result copies them out.  The
// reference has no counterpart (it runs one skycube per process).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/skycube_amd.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Pinned host memory for results: slabs of at least kSlab bytes, each a bump
// allocator that starts over once every region in it has been released.
// Past kMaxPinned bytes of slabs, results fall back to pageable memory.
class PinnedArena {
 public:
  static constexpr size_t kSlab = 256ull << 20;
  static constexpr size_t kMaxPinned = 64ull << 30;
  ~PinnedArena() {
    for (Slab& sl : slabs_) (void)hipHostFree(sl.p);
  }
  // Region of `bytes` (256-byte aligned); *slab = its slab, or -1 if none.
  void* alloc(size_t bytes, int* slab) {
    bytes = (bytes + 255) & ~size_t(255);
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t i = 0; i < slabs_.size(); ++i) {
      Slab& sl = slabs_[i];
      if (sl.cap - sl.top >= bytes) {
        void* p = sl.p + sl.top;
        sl.top += bytes;
        sl.live++;
        *slab = (int)i;
        return p;
      }
    }
    const size_t cap = std::max(kSlab, bytes);
    void* p = nullptr;
    if (total_ + cap > kMaxPinned || hipHostMalloc(&p, cap, hipHostMallocDefault) != hipSuccess) {
      *slab = -1;
      return nullptr;
    }
    total_ += cap;
    slabs_.push_back({static_cast<uint8_t*>(p), cap, bytes, 1});
    *slab = (int)slabs_.size() - 1;
    return p;
  }
  void release(int slab) {
    if (slab < 0) return;
    std::lock_guard<std::mutex> lk(mu_);
    Slab& sl = slabs_[(size_t)slab];
    if (--sl.live == 0) sl.top = 0;
  }

 private:
  struct Slab {
    uint8_t* p;
    size_t cap, top;
    uint32_t live;
  };
  std::mutex mu_;
  std::vector<Slab> slabs_;
  size_t total_ = 0;
};

struct Request {
  uint64_t id = 0;
  // rows on the host: a pinned region (rows_slab >= 0) or rows_v
  float* rows = nullptr;
  int rows_slab = -1;
  std::vector<float> rows_v;
  // rows in device memory once the upload is enqueued, and its event
  float* dev = nullptr;
  hipEvent_t up = nullptr;
  uint64_t resident = 0;  // bytes counted against the resident-row reserve
  uint32_t n = 0, d = 0, max_d = 0;
  uint64_t footprint = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 on the device, 2 done
  int rc = SKY_OK;
  std::string err;
  uint32_t batch_size = 0;
  // result: n_ext ids and n_words bitmap words, in a pinned region (slab
  // >= 0) or in the fallback vectors
  uint32_t n_ext = 0;
  uint64_t n_words = 0;
  uint32_t *ext_p = nullptr, *dom_p = nullptr;
  int slab = -1;
  std::vector<uint32_t> ext_v, dom_v;
};

}  // namespace

struct sky_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
  uint64_t budget = 0, share = 0, in_flight = 0, next_id = 1;
  uint32_t max_batch = 0;
  bool stop = false;
  std::vector<sky_ctx*> ctxs;
  std::vector<std::thread> workers;
  PinnedArena arena;
  int device = 0;
  hipStream_t up_stream = nullptr;  // uploads of submitted rows
  hipMemPool_t pool = nullptr;      // device memory of resident rows
  uint64_t resident = 0, resident_cap = 0;
  uint64_t batches = 0, max_concurrent = 0, concurrent = 0;
  double busy_s = 0, busy_since = 0;  // time with at least one batch running
};

extern "C" uint64_t sky_server_footprint(uint32_t n, uint32_t d) {
  // Device bytes of one instance in a batched pass, an upper bound: shared
  // scratch per point (four 16-float row copies, 64-bit sort keys in two
  // buffers, ids, flags, codes: 332 B; the mask index: 52 B), sort
  // temporaries (16 B), the row upload (4d B), and ids and bitmaps as if
  // every point were in the extended skyline; per instance, pivots, minima,
  // offsets (1.1 KB) and the bucket table (4 B per median mask).  Buffers
  // are allocated with 1/8 headroom.
  const uint64_t per_point = 332 + 52 + 16 + 4ull * d + 4 + 4ull * sky_words_per_point(d);
  return ((uint64_t)n * per_point + 1100 + (4ull << d)) * 9 / 8;
}

namespace {

// Fixed device bytes of a batched pass (allocation minimums, counters,
// offsets of the sixteen dimensionality groups, sort temporaries).
constexpr uint64_t kBatchFixed = 4ull << 20;

// Device bytes of a request's rows.
size_t row_bytes(const Request* r) { return sizeof(float) * (size_t)r->n * r->d; }

// Gives back a request's rows: the device copy (stream-ordered on `st`, after
// the work that read it), the upload event, the host copy.
void drop_rows(sky_server* s, Request* r, hipStream_t st) {
  if (r->dev) (void)hipFreeAsync(r->dev, st);
  r->dev = nullptr;
  if (r->up) (void)hipEventDestroy(r->up);
  r->up = nullptr;
  s->arena.release(r->rows_slab);
  r->rows_slab = -1;
  r->rows = nullptr;
  r->rows_v.clear();
  r->rows_v.shrink_to_fit();
}

void run_batch(sky_server* s, sky_ctx* ctx, std::vector<Request*>& take, uint64_t need) {
  const size_t B = take.size();
  hipStream_t st = static_cast<hipStream_t>(sky_ctx_stream(ctx));
  std::vector<uint32_t> n(B), d(B), m(B);
  std::vector<const float*> src(B);
  // buffers kept from earlier batches count against the share: drop them
  // when they and this batch together might not fit
  int rc = SKY_OK;
  if (sky_ctx_bytes_in_use(ctx) + need > s->share) rc = sky_ctx_release(ctx);
  for (size_t k = 0; k < B && rc == SKY_OK; ++k) {
    Request* r = take[k];
    n[k] = r->n;
    d[k] = r->d;
    m[k] = r->max_d;
    const size_t bytes = row_bytes(r);
    if (r->dev) {  // resident: order the batch after its upload
      if (hipStreamWaitEvent(st, r->up, 0) != hipSuccess) rc = SKY_ERR_DEVICE;
    } else if (bytes) {  // not resident (reserve full): upload on this stream
      void* p = nullptr;
      if (hipMallocFromPoolAsync(&p, bytes, s->pool, st) != hipSuccess) {
        rc = SKY_ERR_OOM_BUDGET;
      } else {
        r->dev = static_cast<float*>(p);
        if (hipMemcpyAsync(r->dev, r->rows, bytes, hipMemcpyHostToDevice, st) != hipSuccess)
          rc = SKY_ERR_DEVICE;
      }
    }
    src[k] = r->dev;
  }
  uint64_t tot = 0;
  if (rc == SKY_OK)
    rc = sky_batch_mdmc_dev(ctx, src.data(), (uint32_t)B, n.data(), d.data(), m.data(), &tot);
  std::vector<uint64_t> eo(B + 1), dof(B + 1);
  if (rc == SKY_OK) rc = sky_batch_fetch(ctx, eo.data(), nullptr, dof.data(), nullptr, 0, 0);
  if (rc == SKY_OK) {
    std::vector<uint32_t*> ed(B), dd(B);
    for (size_t k = 0; k < B; ++k) {
      Request* r = take[k];
      r->n_ext = (uint32_t)(eo[k + 1] - eo[k]);
      r->n_words = dof[k + 1] - dof[k];
      const size_t ext_bytes = ((size_t)r->n_ext * 4 + 255) & ~size_t(255);
      const size_t bytes = ext_bytes + (size_t)r->n_words * 4;
      uint8_t* p = bytes ? static_cast<uint8_t*>(s->arena.alloc(bytes, &r->slab)) : nullptr;
      if (bytes && !p) {  // no pinned memory left: pageable
        r->ext_v.resize(r->n_ext);
        r->dom_v.resize(r->n_words);
        r->ext_p = r->ext_v.data();
        r->dom_p = r->dom_v.data();
      } else if (p) {
        r->ext_p = reinterpret_cast<uint32_t*>(p);
        r->dom_p = reinterpret_cast<uint32_t*>(p + ext_bytes);
      }
      ed[k] = r->n_ext ? r->ext_p : nullptr;
      dd[k] = r->n_words ? r->dom_p : nullptr;
    }
    rc = sky_batch_fetch_into(ctx, ed.data(), dd.data());
  }
  const std::string err =
      rc == SKY_OK ? "" : (rc == SKY_ERR_OOM_BUDGET ? "no device memory for the rows" : sky_last_error(ctx));
  const double t = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  for (Request* r : take) {
    drop_rows(s, r, st);
    s->resident -= r->resident;
    r->resident = 0;
    r->rc = rc;
    r->err = err;
    r->t_done = t;
    r->state = 2;
    r->batch_size = (uint32_t)B;
    if (rc != SKY_OK) {
      s->arena.release(r->slab);
      r->slab = -1;
    }
  }
  s->in_flight -= need;
  s->batches++;
  if (--s->concurrent == 0) s->busy_s += t - s->busy_since;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

void worker(sky_server* s, size_t w) {
  sky_ctx* ctx = s->ctxs[w];
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
    if (s->stop && s->queue.empty()) return;
    // admit, first come first served, while the footprint fits this
    // worker's share (every request fits a share alone, checked at submit),
    // and at most an equal split of the queue over the workers, so that a
    // burst keeps every worker's batch in the pipeline
    std::vector<Request*> take;
    uint64_t need = kBatchFixed;
    const size_t nw = s->ctxs.size();
    const size_t cap =
        std::min<size_t>(s->max_batch, std::max<size_t>(1, (s->queue.size() + nw - 1) / nw));
    while (!s->queue.empty() && take.size() < cap) {
      Request* r = s->queue.front();
      if (!take.empty() && need + r->footprint > s->share) break;
      need += r->footprint;
      r->state = 1;
      r->t_start = now_s();
      take.push_back(r);
      s->queue.pop_front();
    }
    s->in_flight += need;
    if (s->concurrent++ == 0) s->busy_since = now_s();
    s->max_concurrent = std::max(s->max_concurrent, s->concurrent);
    lk.unlock();
    run_batch(s, ctx, take, need);
    lk.lock();
  }
}

// Frees the server's contexts, resident-row pool and upload stream.
void release_device(sky_server* s) {
  for (sky_ctx* c : s->ctxs) sky_ctx_destroy(c);
  s->ctxs.clear();
  if (s->up_stream) {
    (void)hipStreamSynchronize(s->up_stream);
    (void)hipStreamDestroy(s->up_stream);
  }
  s->up_stream = nullptr;
  if (s->pool) (void)hipMemPoolDestroy(s->pool);
  s->pool = nullptr;
}

}  // namespace

extern "C" {

int sky_server_create(int device, size_t hbm_budget_bytes, uint32_t workers, uint32_t max_batch,
                      sky_server** out) {
  if (!out || workers == 0 || workers > 8) return SKY_ERR_ARG;
  *out = nullptr;
  auto s = std::make_unique<sky_server>();
  s->max_batch = max_batch ? max_batch : 4096;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count ||
      hipSetDevice(device) != hipSuccess)
    return SKY_ERR_DEVICE;
  if (hbm_budget_bytes == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return SKY_ERR_DEVICE;
    hbm_budget_bytes = free_b / 10 * 8;
  }
  s->budget = hbm_budget_bytes;
  s->device = device;
  // 1/16 of the budget holds the rows of waiting requests; the rest is
  // split evenly over the workers
  s->resident_cap = hbm_budget_bytes / 16;
  s->share = (hbm_budget_bytes - s->resident_cap) / workers;
  if (s->share <= kBatchFixed) return SKY_ERR_OOM_BUDGET;
  int rc = hipStreamCreateWithFlags(&s->up_stream, hipStreamNonBlocking) == hipSuccess
               ? SKY_OK : SKY_ERR_DEVICE;
  if (rc == SKY_OK) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    uint64_t keep = ~0ull;  // freed blocks stay in the pool for the next requests
    int no = 0;             // reuse never makes one stream wait for another
    if (hipMemPoolCreate(&s->pool, &props) != hipSuccess ||
        hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess ||
        hipMemPoolSetAttribute(s->pool, hipMemPoolReuseAllowInternalDependencies, &no) !=
            hipSuccess)
      rc = SKY_ERR_DEVICE;
  }
  for (uint32_t w = 0; w < workers && rc == SKY_OK; ++w) {
    sky_ctx* c = nullptr;
    rc = sky_ctx_create(device, s->share, &c);
    if (rc == SKY_OK) s->ctxs.push_back(c);
  }
  if (rc != SKY_OK) {
    release_device(s.get());
    return rc;
  }
  for (uint32_t w = 0; w < workers; ++w) s->workers.emplace_back(worker, s.get(), (size_t)w);
  *out = s.release();
  return SKY_OK;
}

void sky_server_destroy(sky_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (std::thread& t : s->workers) t.join();
  release_device(s);
  delete s;
}

int sky_server_submit(sky_server* s, const float* rows, uint32_t n, uint32_t d, uint32_t max_d,
                      uint64_t* id) {
  if (!s || !id || (n && !rows) || d < 1 || d > 16 || max_d > 16) return SKY_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->n = n;
  r->d = d;
  r->max_d = max_d;
  r->footprint = sky_server_footprint(n, d);
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (r->footprint + kBatchFixed > s->share) return SKY_ERR_OOM_BUDGET;
  }
  const size_t count = (size_t)n * d, bytes = sizeof(float) * count;
  for (size_t i = 0; i < count; ++i)
    if (std::isnan(rows[i])) return SKY_ERR_NAN;
  if (bytes) {
    r->rows = static_cast<float*>(s->arena.alloc(bytes, &r->rows_slab));
    if (!r->rows) {  // no pinned memory left: pageable
      r->rows_v.assign(rows, rows + count);
      r->rows = r->rows_v.data();
    } else {
      std::memcpy(r->rows, rows, bytes);
    }
    // reserve room among the resident rows, then enqueue the upload
    bool room = false;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->resident + bytes <= s->resident_cap) {
        s->resident += bytes;
        r->resident = bytes;
        room = true;
      }
    }
    if (room) {
      int prev = 0;
      const bool dev_ok = hipGetDevice(&prev) == hipSuccess && hipSetDevice(s->device) == hipSuccess;
      void* p = nullptr;
      bool ok = dev_ok && hipMallocFromPoolAsync(&p, bytes, s->pool, s->up_stream) == hipSuccess;
      if (ok) {
        r->dev = static_cast<float*>(p);
        ok = hipMemcpyAsync(r->dev, r->rows, bytes, hipMemcpyHostToDevice, s->up_stream) ==
                 hipSuccess &&
             hipEventCreateWithFlags(&r->up, hipEventDisableTiming) == hipSuccess &&
             hipEventRecord(r->up, s->up_stream) == hipSuccess;
      }
      if (!ok) {  // the worker uploads the rows instead
        if (r->up) (void)hipEventDestroy(r->up);
        r->up = nullptr;
        if (r->dev) {
          (void)hipStreamSynchronize(s->up_stream);
          (void)hipFreeAsync(r->dev, s->up_stream);
        }
        r->dev = nullptr;
        std::lock_guard<std::mutex> lk(s->mu);
        s->resident -= bytes;
        r->resident = 0;
      }
      if (dev_ok) (void)hipSetDevice(prev);
    }
  }
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  *id = r->id;
  s->queue.push_back(r.get());
  s->all.emplace(r->id, std::move(r));
  s->cv_work.notify_one();
  return SKY_OK;
}

int sky_server_wait(sky_server* s, uint64_t id, int timeout_ms) {
  if (!s) return SKY_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->all.find(id) == s->all.end()) return SKY_ERR_STATE;
  // the request is looked up again after every wake-up: another thread may
  // have taken its result (and freed it) meanwhile
  auto done = [&] {
    auto it = s->all.find(id);
    return it == s->all.end() || it->second->state == 2;
  };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
    return SKY_ERR_STATE;
  }
  auto it = s->all.find(id);
  return it == s->all.end() ? SKY_ERR_STATE : it->second->rc;
}

int sky_server_result(sky_server* s, uint64_t id, uint32_t* ext_pids, uint32_t* n_ext,
                      uint32_t* dom_words, uint64_t cap_words, double* latency_ms,
                      uint32_t* batch_size) {
  if (!s || !n_ext) return SKY_ERR_ARG;
  std::unique_ptr<Request> own;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return SKY_ERR_STATE;
    Request* r = it->second.get();
    if (r->state != 2) return SKY_ERR_STATE;
    if (r->rc != SKY_OK) {
      const int rc = r->rc;
      s->all.erase(it);
      return rc;
    }
    *n_ext = r->n_ext;
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    if (batch_size) *batch_size = r->batch_size;
    if (!ext_pids || !dom_words) return SKY_OK;  // sizes only; the result stays
    if (cap_words < r->n_words) return SKY_ERR_CAPACITY;
    own = std::move(it->second);  // the copy runs outside the lock
    s->all.erase(it);
  }
  if (own->n_ext) std::memcpy(ext_pids, own->ext_p, sizeof(uint32_t) * own->n_ext);
  if (own->n_words) std::memcpy(dom_words, own->dom_p, sizeof(uint32_t) * own->n_words);
  s->arena.release(own->slab);
  return SKY_OK;
}

int sky_server_stats(sky_server* s, uint64_t* batches, uint64_t* max_concurrent,
                     uint64_t* queued, double* busy_ms) {
  if (!s) return SKY_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (busy_ms) *busy_ms = (s->busy_s + (s->concurrent ? now_s() - s->busy_since : 0.0)) * 1e3;
  if (batches) *batches = s->batches;
  if (max_concurrent) *max_concurrent = s->max_concurrent;
  if (queued) *queued = s->queue.size();
  return SKY_OK;
}

}  // extern "C"
