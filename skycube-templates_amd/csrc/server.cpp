// server.cpp -- continuous-batching skycube server over the C ABI.
//
// Requests (one skycube each: rows, d, max_d) are queued by sky_server_submit
// from any thread.  Worker threads, each with its own device context and
// stream, repeatedly take every queued request that fits the HBM budget (the
// estimated device footprint of the requests in flight, first come first
// served) and run them as one batched MDMC pass (sky_batch_mdmc).  A request
// that arrives while the device is busy joins the next batch, and one
// worker's host work (row upload, set-up, result download) overlaps the
// other's kernels.  Each worker owns hbm_budget / workers bytes of device
// memory; a batch is admitted only while the estimated footprint of its
// requests fits that share.  Results are copied to host memory owned by the
// request until sky_server_result hands them over.  The reference has no
// counterpart (it runs one skycube per process).
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

struct Request {
  uint64_t id = 0;
  std::vector<float> rows;
  uint32_t n = 0, d = 0, max_d = 0;
  uint64_t footprint = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 on the device, 2 done
  int rc = SKY_OK;
  std::string err;
  uint32_t batch_size = 0;
  std::vector<uint32_t> ext, dom;
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
  uint64_t batches = 0, max_concurrent = 0, concurrent = 0;
  double busy_s = 0, busy_since = 0;  // time with at least one batch running
};

extern "C" uint64_t sky_server_footprint(uint32_t n, uint32_t d) {
  // Device bytes of one instance in a batched pass, an upper bound: shared
  // scratch per point (four 16-float row copies, 64-bit sort keys in two
  // buffers, ids, flags, codes: 332 B), sort temporaries (16 B), the row
  // upload (4d B), and ids and bitmaps as if every point were in the
  // extended skyline; per instance, pivots, minima and offsets (1.1 KB).
  // Buffers are allocated with 1/8 headroom.
  const uint64_t per_point = 332 + 16 + 4ull * d + 4 + 4ull * sky_words_per_point(d);
  return ((uint64_t)n * per_point + 1100) * 9 / 8;
}

namespace {

// Fixed device bytes of a batched pass (allocation minimums, counters,
// offsets of the sixteen dimensionality groups, sort temporaries).
constexpr uint64_t kBatchFixed = 4ull << 20;

void run_batch(sky_server* s, sky_ctx* ctx, std::vector<Request*>& take, uint64_t need) {
  std::vector<uint32_t> n(take.size()), d(take.size()), m(take.size());
  size_t total = 0;
  for (size_t k = 0; k < take.size(); ++k) {
    n[k] = take[k]->n;
    d[k] = take[k]->d;
    m[k] = take[k]->max_d;
    total += take[k]->rows.size();
  }
  // buffers kept from earlier batches count against the share: drop them
  // when they and this batch together might not fit
  int rc0 = SKY_OK;
  if (sky_ctx_bytes_in_use(ctx) + need > s->share) rc0 = sky_ctx_release(ctx);
  std::vector<float> rows(std::max<size_t>(total, 1));
  size_t o = 0;
  for (Request* r : take) {
    std::copy(r->rows.begin(), r->rows.end(), rows.begin() + (long)o);
    o += r->rows.size();
  }
  uint64_t tot = 0;
  int rc = rc0;
  if (rc == SKY_OK)
    rc = sky_batch_mdmc(ctx, rows.data(), 0, (uint32_t)take.size(), n.data(), d.data(), m.data(),
                        0, &tot);
  std::vector<uint64_t> eo(take.size() + 1), dof(take.size() + 1);
  std::vector<uint32_t> ids, words;
  if (rc == SKY_OK) rc = sky_batch_fetch(ctx, eo.data(), nullptr, dof.data(), nullptr, 0, 0);
  if (rc == SKY_OK) {
    ids.resize(std::max<uint64_t>(eo.back(), 1));
    words.resize(std::max<uint64_t>(dof.back(), 1));
    rc = sky_batch_fetch(ctx, eo.data(), ids.data(), dof.data(), words.data(), ids.size(),
                         words.size());
  }
  const std::string err = rc == SKY_OK ? "" : sky_last_error(ctx);
  const double t = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  for (size_t k = 0; k < take.size(); ++k) {
    Request* r = take[k];
    r->rc = rc;
    r->err = err;
    if (rc == SKY_OK) {
      r->ext.assign(ids.begin() + (long)eo[k], ids.begin() + (long)eo[k + 1]);
      r->dom.assign(words.begin() + (long)dof[k], words.begin() + (long)dof[k + 1]);
    }
    r->rows.clear();
    r->rows.shrink_to_fit();
    r->t_done = t;
    r->state = 2;
    r->batch_size = (uint32_t)take.size();
  }
  s->in_flight -= need;
  s->batches++;
  if (--s->concurrent == 0) s->busy_s += t - s->busy_since;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

void worker(sky_server* s, size_t w) {
  sky_ctx* ctx = s->ctxs[w];
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
    if (s->stop && s->queue.empty()) return;
    // admit, first come first served, while the footprint fits this
    // worker's share; every request fits a share alone (checked at submit)
    std::vector<Request*> take;
    uint64_t need = kBatchFixed;
    while (!s->queue.empty() && take.size() < s->max_batch) {
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
  s->share = hbm_budget_bytes / workers;
  if (s->share <= kBatchFixed) return SKY_ERR_OOM_BUDGET;
  for (uint32_t w = 0; w < workers; ++w) {
    sky_ctx* c = nullptr;
    const int rc = sky_ctx_create(device, s->share, &c);
    if (rc != SKY_OK) {
      for (sky_ctx* x : s->ctxs) sky_ctx_destroy(x);
      return rc;
    }
    s->ctxs.push_back(c);
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
  for (sky_ctx* c : s->ctxs) sky_ctx_destroy(c);
  delete s;
}

int sky_server_submit(sky_server* s, const float* rows, uint32_t n, uint32_t d, uint32_t max_d,
                      uint64_t* id) {
  if (!s || !id || (n && !rows) || d < 1 || d > 16 || max_d > 16) return SKY_ERR_ARG;
  auto r = std::make_unique<Request>();
  r->n = n;
  r->d = d;
  r->max_d = max_d;
  r->rows.assign(rows, rows + (size_t)n * d);
  for (float v : r->rows)
    if (std::isnan(v)) return SKY_ERR_NAN;
  r->footprint = sky_server_footprint(n, d);
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->footprint + kBatchFixed > s->share) return SKY_ERR_OOM_BUDGET;
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
  auto it = s->all.find(id);
  if (it == s->all.end()) return SKY_ERR_STATE;
  Request* r = it->second.get();
  auto done = [&] { return r->state == 2; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, done);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) {
    return SKY_ERR_STATE;
  }
  return r->rc;
}

int sky_server_result(sky_server* s, uint64_t id, uint32_t* ext_pids, uint32_t* n_ext,
                      uint32_t* dom_words, uint64_t cap_words, double* latency_ms,
                      uint32_t* batch_size) {
  if (!s || !n_ext) return SKY_ERR_ARG;
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
  *n_ext = (uint32_t)r->ext.size();
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  if (batch_size) *batch_size = r->batch_size;
  if (!ext_pids || !dom_words) return SKY_OK;  // sizes only; the result stays
  if (cap_words < r->dom.size()) return SKY_ERR_CAPACITY;
  std::copy(r->ext.begin(), r->ext.end(), ext_pids);
  std::copy(r->dom.begin(), r->dom.end(), dom_words);
  s->all.erase(it);
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
