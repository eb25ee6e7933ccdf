// noalloc.hip -- sky_mdmc_device_noalloc: the MDMC of sky_mdmc_device for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// Every buffer of the call other than the caller's inputs and outputs is
// carved from `scratch`, one file-scope device array per GPU that the
// library holds until the process ends.  Its device address is looked up
// once per device (sky_ctx_create -> noalloc_init_device) and only that
// address is handed to kernels and copies.  Calls on one GPU share the
// array, so calls that overlap on one GPU race; every call clears what
// must start at zero (counters, flags) with fill kernels.
#include <mutex>

#include "internal.hpp"

namespace {

static __device__ __align__(256) uint8_t scratch[skyi::kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
bool g_done[kMaxDevices] = {};

}  // namespace

namespace skyi {

int noalloc_init_device(int device) {
  if (device < 0 || device >= kMaxDevices) return SKY_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[device]) return g_ptr[device] ? SKY_OK : SKY_ERR_DEVICE;
  g_done[device] = true;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return SKY_ERR_DEVICE;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SKY_ERR_DEVICE;
  g_ptr[device] = p;
  return SKY_OK;
}

// rocPRIM temporaries of one call: the largest is the 64-bit key sort of the
// code order (12 bytes per point and its histograms and look-back state)
uint64_t noalloc_sort_tmp(uint64_t n) { return 16ull * n + (1ull << 20); }

uint64_t noalloc_bytes(uint64_t n, uint32_t d) {
  // every buffer of the call with n_ext = n: 122 + 12 d bytes per point (the
  // score-order keys, ids, rows and codes of the split, its mask index, the
  // extended skyline's rows, ids and code order with its group and subgroup
  // summaries, the deferral list), the
  // rocPRIM temporaries, the median-mask bucket table of 2^d + 1 entries, and
  // 1 MiB for the 256-byte rounding of about 50 carves and the small tables
  return n * (122ull + 12ull * d) + noalloc_sort_tmp(n) + (4ull << d) + (1ull << 20);
}

}  // namespace skyi

extern "C" int sky_mdmc_device_noalloc(sky_ctx* c, const float* rows_dev, uint32_t n, uint32_t d,
                                       uint32_t max_d, uint32_t* ext_pids_dev,
                                       uint32_t* dom_words_dev, uint64_t dom_cap_words,
                                       uint32_t* n_ext) {
  using namespace skyi;
  if (!c) return SKY_ERR_ARG;
  c->err.clear();
  HIPCHK(hipSetDevice(c->device));
  if (d < 1 || d > 16) return fail(c, SKY_ERR_ARG, "d=%u outside [1,16]", d);
  if (!n_ext || (n && (!rows_dev || !ext_pids_dev || !dom_words_dev)))
    return fail(c, SKY_ERR_ARG, "null argument");
  const uint64_t need = noalloc_bytes(n, d);
  if (need > kNoallocScratchBytes)
    return fail(c, SKY_ERR_CAPACITY,
                "n=%u, d=%u needs %llu bytes of scratch; the allocation-free MDMC holds %llu",
                n, d, (unsigned long long)need, (unsigned long long)kNoallocScratchBytes);
  if (c->device < 0 || c->device >= kMaxDevices || !g_ptr[c->device])
    return fail(c, SKY_ERR_DEVICE, "scratch of device %d was not set up", c->device);
  const uint64_t W = sky_words_per_point(d);
  if (dom_cap_words < (uint64_t)n * W)
    return fail(c, SKY_ERR_CAPACITY, "dom_words_dev holds %llu words, %llu may be needed",
                (unsigned long long)dom_cap_words, (unsigned long long)((uint64_t)n * W));
  // the context's own buffers step aside for the call; the results go
  // straight to the caller's buffers
  DevBuf* bufs[kSingleBufs];
  single_bufs(c, bufs);
  DevBuf saved[kSingleBufs];
  for (int i = 0; i < kSingleBufs; ++i) {
    saved[i] = *bufs[i];
    *bufs[i] = DevBuf{};
  }
  c->ext_pids = DevBuf{ext_pids_dev, sizeof(uint32_t) * (size_t)n};
  c->dom = DevBuf{dom_words_dev, sizeof(uint32_t) * (size_t)dom_cap_words};
  Carve carve{static_cast<uint8_t*>(g_ptr[c->device]), kNoallocScratchBytes, 0};
  const bool timing = c->timing;
  c->timing = false;  // no event creation on this path
  c->carve = &carve;
  // the sorts' temporaries first, at their largest, so that every sort of
  // the call reuses one slice
  int rc = ensure(c, c->cub_tmp, noalloc_sort_tmp(n));
  if (rc == SKY_OK) rc = run_mdmc(c, rows_dev, n, d, max_d, n_ext);
  c->carve = nullptr;
  c->timing = timing;
  for (int i = 0; i < kSingleBufs; ++i) *bufs[i] = saved[i];
  // the last result held by the context is gone (its buffers were not used)
  c->has_mdmc = false;
  c->has_hc = false;
  return rc;
}
