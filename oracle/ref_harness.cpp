// ref_harness.cpp -- test infrastructure: a C ABI over the reference's own
// CPU skyline code, for pinning the oracle and timing the reference's CPU
// path.  Never part of the product.
//
// Built by oracle/Makefile (target `ref`) into oracle/_ref/libskyref.so from
// the reference sources where they lie: src/hybrid/mommyq.cpp (MommyQ, the
// per-cuboid skyline of the reference's STSC, with PQFilter and the
// DominateLeft_array tests of src/utilities/dt_avx2.h it includes) and
// src/utilities/lattice_utilities.cpp (generate_lattice, get_tds_list).
//
// The reference's STSC driver, HybridCube<DIMS>::Execute
// (src/stsc/HybridCube.cpp:30-136), includes src/skycube/Skycube.h, which
// pulls in Boost.Thread through src/skycube/NodeAuxilliary.h:20; Boost is
// not installed, so that file is not compiled.  ref_stsc below walks the
// lattice the way it does -- full space first with every thread, then each
// level in descending popcount order, levels above max_d skipped, every
// cuboid of a level in parallel (one thread each), its candidates the sky
// and ext lists of its parent with the fewest, or of the full space when
// parents[0] has an empty skyline (HybridCube.cpp:76-100, findMinParent
// :166-176), MommyQ above 50 candidates and the pairwise DominateLeft loop
// at or below (:102-131) -- calling the reference's MommyQ and
// DominateLeft_array for all skyline work.
#include <omp.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hybrid/mommyq.h"
#include "utilities/lattice_utilities.h"

std::vector<unsigned int> get_tds_list(int m);
lattice_node* generate_lattice(int m);

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One cuboid: sky / ext pids of the candidates `cand` over subspace `sub`
// (MommyQ above 50 candidates or when `always_mommyq`, as for the full
// space, else the pairwise loop).
template <int DIMS>
void cuboid(const float* rows, uint32_t sub, const uint32_t* cand, uint32_t nc, uint32_t threads,
            uint32_t accum, uint32_t pq, std::vector<unsigned>* sky, std::vector<unsigned>* ext,
            bool always_mommyq = false) {
  std::vector<int> dims;
  for (int j = 0; j < DIMS; ++j)
    if (sub >> j & 1u) dims.push_back(j);
  std::vector<HybridTuple<DIMS>> t(nc);
  for (uint32_t k = 0; k < nc; ++k) {
    t[k].pid = (int)cand[k];
    t[k].elems = const_cast<float*>(rows) + (size_t)cand[k] * DIMS;
    t[k].score = 0.0f;
    t[k].partition = 0;
  }
  if (nc > 50 || (always_mommyq && nc > 0)) {
    MommyQ<DIMS> mom(threads, nc, accum, pq, sky, ext, &dims);
    mom.Init(t.data());
    mom.Execute();
    return;
  }
  std::vector<int> res(nc, 0);
  for (uint32_t i = 0; i < nc; ++i)
    for (uint32_t j = 0; j < nc; ++j) {
      if (j == i) continue;
      const int r = DominateLeft_array<DIMS>(t[j].elems, t[i].elems, sub);
      if (r == -2) {
        res[i] = r;
        break;
      }
      if (r == -1) res[i] = r;
    }
  for (uint32_t i = 0; i < nc; ++i) {
    if (res[i] == 0) sky->push_back((unsigned)t[i].pid);
    if (res[i] == -1) ext->push_back((unsigned)t[i].pid);
  }
}

template <int DIMS>
int stsc(const float* rows, uint32_t n, uint32_t max_d, uint32_t threads, uint32_t accum,
         uint32_t pq, std::vector<std::vector<unsigned>>& sky,
         std::vector<std::vector<unsigned>>& ext, double* level_s) {
  const int m = 1 << DIMS;
  lattice_node* lat = generate_lattice(m);
  const std::vector<unsigned> order = get_tds_list(m);
  sky.assign(m, {});
  ext.assign(m, {});
  const uint32_t full = order.at(0);
  std::vector<uint32_t> all(n);
  for (uint32_t i = 0; i < n; ++i) all[i] = i;
  double t0 = now_s();
  cuboid<DIMS>(rows, full, all.data(), n, threads, accum, pq, &sky[full], &ext[full], true);
  level_s[DIMS] = now_s() - t0;
  // the remaining cuboids by level, descending popcount
  std::vector<std::vector<uint32_t>> levels(DIMS + 1);
  for (size_t k = 1; k < order.size(); ++k) levels[__builtin_popcount(order[k])].push_back(order[k]);
  for (int lv = DIMS - 1; lv >= 1; --lv) {
    if ((uint32_t)lv > max_d) continue;
    const std::vector<uint32_t>& L = levels[lv];
    t0 = now_s();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (size_t k = 0; k < L.size(); ++k) {
      const uint32_t s = L[k];
      uint32_t src = full;
      if (!sky[lat[s].parents[0]].empty()) {
        size_t best = sky[lat[s].parents[0]].size() + ext[lat[s].parents[0]].size();
        src = lat[s].parents[0];
        for (size_t i = 1; i < lat[s].parents.size(); ++i) {
          const uint32_t p = lat[s].parents[i];
          if (sky[p].size() + ext[p].size() < best) {
            best = sky[p].size() + ext[p].size();
            src = p;
          }
        }
      }
      std::vector<uint32_t> cand(sky[src].begin(), sky[src].end());
      cand.insert(cand.end(), ext[src].begin(), ext[src].end());
      cuboid<DIMS>(rows, s, cand.data(), (uint32_t)cand.size(), 1, accum, pq, &sky[s], &ext[s]);
    }
    level_s[lv] = now_s() - t0;
  }
  delete[] lat;
  return 0;
}

template <int DIMS>
double timed_level(const float* rows, const uint32_t* subs, const uint64_t* cand_off,
                   const uint32_t* cand_ids, uint32_t k, uint32_t threads, uint32_t accum,
                   uint32_t pq, double* each_s) {
  const double t0 = now_s();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (uint32_t i = 0; i < k; ++i) {
    std::vector<unsigned> sky, ext;
    const double c0 = now_s();
    cuboid<DIMS>(rows, subs[i], cand_ids + cand_off[i], (uint32_t)(cand_off[i + 1] - cand_off[i]),
                 1, accum, pq, &sky, &ext);
    if (each_s) each_s[i] = now_s() - c0;
  }
  return now_s() - t0;
}

#define REF_DISPATCH(DV, ...)                            \
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
    default: return -1;                                  \
  }

}  // namespace

extern "C" {

// The lattice of the last ref_stsc call, for ref_stsc_fetch.
static std::vector<std::vector<unsigned>> g_sky, g_ext;

// The skycube by the reference's STSC walk.  Per cuboid s (1 .. 2^d - 1),
// CSR offsets of the sky and ext pid lists (sky_off / ext_off: 2^d + 1
// entries); level_s[k] = seconds of level k (level d is the full space).
// The lists are kept for ref_stsc_fetch.
int ref_stsc(const float* rows, uint32_t n, uint32_t d, uint32_t max_d, uint32_t threads,
             uint32_t accum, uint32_t pq, uint64_t* sky_off, uint64_t* ext_off, double* level_s) {
  if (d < 1 || d > 16 || !sky_off || !ext_off || !level_s) return -1;
  if (max_d == 0 || max_d > d) max_d = d;
  for (uint32_t k = 0; k <= d; ++k) level_s[k] = 0.0;
  REF_DISPATCH(d, stsc<D>(rows, n, max_d, threads, accum, pq, g_sky, g_ext, level_s));
  const uint32_t S = 1u << d;
  uint64_t os = 0, oe = 0;
  for (uint32_t s = 0; s < S; ++s) {
    sky_off[s] = os;
    ext_off[s] = oe;
    os += g_sky[s].size();
    oe += g_ext[s].size();
  }
  sky_off[S] = os;
  ext_off[S] = oe;
  return 0;
}

// The lists of the last ref_stsc call, in the order MommyQ left them, at
// the offsets it returned.
int ref_stsc_fetch(uint32_t* sky_ids, uint32_t* ext_ids) {
  uint64_t os = 0, oe = 0;
  for (size_t s = 0; s < g_sky.size(); ++s) {
    if (!g_sky[s].empty()) std::memcpy(sky_ids + os, g_sky[s].data(), sizeof(uint32_t) * g_sky[s].size());
    if (!g_ext[s].empty()) std::memcpy(ext_ids + oe, g_ext[s].data(), sizeof(uint32_t) * g_ext[s].size());
    os += g_sky[s].size();
    oe += g_ext[s].size();
  }
  return 0;
}

// Wall seconds of k cuboids of the reference's walk run as one parallel
// level would run them (schedule(dynamic, 1) over `threads`, one thread per
// cuboid): cuboid i is subspace subs[i] over candidates
// cand_ids[cand_off[i] .. cand_off[i+1]); results are discarded.  each_s
// (nullable, k entries) receives each cuboid's own seconds.
int ref_cuboids_timed(const float* rows, uint32_t d, const uint32_t* subs, const uint64_t* cand_off,
                      const uint32_t* cand_ids, uint32_t k, uint32_t threads, uint32_t accum,
                      uint32_t pq, double* wall_s, double* each_s) {
  if (d < 1 || d > 16 || !wall_s) return -1;
  REF_DISPATCH(d, *wall_s = timed_level<D>(rows, subs, cand_off, cand_ids, k, threads, accum, pq,
                                           each_s));
  return 0;
}

// One cuboid with `threads` threads (the full-space pass uses them all):
// its sky / ext pids (capacity n_cand each) and its seconds.
int ref_cuboid(const float* rows, uint32_t d, uint32_t sub, const uint32_t* cand, uint32_t n_cand,
               uint32_t threads, uint32_t accum, uint32_t pq, uint32_t* sky_out, uint32_t* n_sky,
               uint32_t* ext_out, uint32_t* n_ext, double* seconds) {
  if (d < 1 || d > 16 || !n_sky || !n_ext || !seconds) return -1;
  std::vector<unsigned> sky, ext;
  REF_DISPATCH(d, {
    const double t0 = now_s();
    cuboid<D>(rows, sub, cand, n_cand, threads, accum, pq, &sky, &ext, sub == (1u << D) - 1u);
    *seconds = now_s() - t0;
  });
  *n_sky = (uint32_t)sky.size();
  *n_ext = (uint32_t)ext.size();
  if (sky_out && !sky.empty()) std::memcpy(sky_out, sky.data(), sizeof(uint32_t) * sky.size());
  if (ext_out && !ext.empty()) std::memcpy(ext_out, ext.data(), sizeof(uint32_t) * ext.size());
  return 0;
}

}  // extern "C"
