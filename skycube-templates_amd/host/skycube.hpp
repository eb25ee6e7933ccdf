// skycube.hpp -- C++ host side of the MI355X skycube engine.
//
// Mirrors the reference's algorithm objects and result types over the C ABI
// of include/skycube_amd.h:
//   MDMC   (src/mdmc/MDMC.h:34-67): Init(rows), Execute(), getHashCubeStructure()
//   SDSC   (src/sdsc/SDSC.h:18-54): Init(rows), Execute(), getLattice()
//   Config (src/utilities/config.h:13-27), lattice_node
//   (src/utilities/lattice_utilities.h:21-31), getsubspace
//   (src/utilities/utilities.h:253-267), CSV input (utilities.h:140-227).
// All device work happens in libskycube_amd.so; errors raise SkyError.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/skycube_amd.h"

namespace skycube {

struct SkyError : std::runtime_error {
  int code;
  SkyError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Flag set of the reference driver (src/utilities/config.h:13-27).
struct Config {
  std::string input_fname, algo, papi;
  uint32_t threads = 8, alpha_size = 1024, pq_size = 128, partitioning_depth = 3;
  int max_d = -1;
  bool check = false, verbose = false, cpu = true, cross_distribution = false;
  std::vector<int> devices;
};

// One cuboid of the lattice (src/utilities/lattice_utilities.h:21-31).
struct lattice_node {
  unsigned int index = 0;
  std::vector<int> dimensions;
  std::vector<unsigned int> skyline;
  std::vector<unsigned int> extended_sky;
};

using HashCubeMaps = std::map<unsigned int, std::vector<int>>;

// Host buffer in page-locked memory (sky_host_alloc), or in ordinary memory
// when none is left: copies between it and the device then run at full
// link speed.
template <typename T>
class HostBuffer {
 public:
  HostBuffer() = default;
  ~HostBuffer() { release(); }
  HostBuffer(const HostBuffer&) = delete;
  HostBuffer& operator=(const HostBuffer&) = delete;
  // n elements; the contents are not kept when the buffer grows
  void resize(size_t n) {
    if (n > cap_) {
      release();
      const size_t bytes = sizeof(T) * (n ? n : 1);
      void* p = sky_host_alloc(bytes);
      pinned_ = p != nullptr;
      if (!p) p = std::malloc(bytes);
      if (!p) throw std::bad_alloc();
      p_ = static_cast<T*>(p);
      cap_ = n;
    }
    n_ = n;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  T* begin() { return p_; }
  T* end() { return p_ + n_; }
  T& operator[](size_t i) { return p_[i]; }
  bool pinned() const { return pinned_; }

 private:
  void release() {
    if (p_) {
      if (pinned_) sky_host_free(p_);
      else std::free(p_);
    }
    p_ = nullptr;
    cap_ = n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
  bool pinned_ = false;
};

// Splits one CSV line on ',' (a trailing comma adds no field) and converts
// every field with strtof, correctly rounded like istringstream >> float.
inline std::vector<float> split_line(const std::string& line) {
  std::vector<float> out;
  std::stringstream ss(line);
  std::string tok;
  while (std::getline(ss, tok, ',')) {
    if (tok.empty() || tok == "\r") continue;
    char* end = nullptr;
    const float v = std::strtof(tok.c_str(), &end);
    if (end == tok.c_str()) throw SkyError(SKY_ERR_ARG, "PARSE FAILED: '" + tok + "'");
    out.push_back(v);
  }
  return out;
}

// Number of fields on the first line (read_dims, utilities.h:213-227).
inline int read_dims(const std::string& fname) {
  std::ifstream f(fname);
  if (!f.is_open()) throw SkyError(SKY_ERR_ARG, "..No such file: " + fname);
  std::string line;
  std::getline(f, line);
  return (int)split_line(line).size();
}

// Row-major n x d FP32 matrix of a CSV file (read_data, utilities.h:140-178).
inline std::vector<float> read_data(const std::string& fname, uint32_t* n, uint32_t* d) {
  std::ifstream f(fname);
  if (!f.is_open()) throw SkyError(SKY_ERR_ARG, "..No such file: " + fname);
  std::vector<float> rows;
  std::string line;
  uint32_t dims = 0, count = 0;
  while (std::getline(f, line)) {
    const std::vector<float> r = split_line(line);
    if (r.empty()) continue;
    if (count == 0) dims = (uint32_t)r.size();
    if (r.size() != dims) throw SkyError(SKY_ERR_ARG, "ragged CSV row " + std::to_string(count));
    rows.insert(rows.end(), r.begin(), r.end());
    ++count;
  }
  *n = count;
  *d = dims;
  return rows;
}

// Owns one device context.
class Device {
 public:
  explicit Device(int device, size_t budget = 0) {
    const int rc = sky_ctx_create(device, budget, &ctx_);
    if (rc != SKY_OK) throw SkyError(rc, "sky_ctx_create failed on device " + std::to_string(device));
  }
  ~Device() { sky_ctx_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  sky_ctx* get() const { return ctx_; }
  void check(int rc) const {
    if (rc != SKY_OK) throw SkyError(rc, sky_last_error(ctx_));
  }

 private:
  sky_ctx* ctx_ = nullptr;
};

inline int first_device(const Config& cfg) { return cfg.devices.empty() ? 0 : cfg.devices[0]; }

// MDMC: per-point cuboid membership, HashCube result (MDMC.h:34-67).
class MDMC {
 public:
  MDMC(const Config& cfg, uint32_t n, uint32_t d) : cfg_(cfg), n_(n), d_(d), dev_(first_device(cfg)) {}
  // Copies the rows into pinned memory (the reference copies them into a
  // flat buffer, MDMC.cu:204-225, and pins it, MDMC.cu:1068-1071).
  void Init(float** data) {
    rows_.resize((size_t)n_ * d_);
    for (uint32_t i = 0; i < n_; ++i)
      for (uint32_t j = 0; j < d_; ++j) rows_[(size_t)i * d_ + j] = data[i][j];
  }
  void Init(const std::vector<float>& rows) {
    rows_.resize(rows.size());
    std::copy(rows.begin(), rows.end(), rows_.begin());
  }
  // Device skycube, then the HashCube grouped on the device and loaded
  // into the reference's map layout.
  void Execute() {
    const uint32_t max_d = cfg_.max_d > 0 ? (uint32_t)cfg_.max_d : d_;
    ext_.assign(n_ ? n_ : 1, 0);
    uint32_t ne = 0;
    dev_.check(sky_mdmc(dev_.get(), rows_.data(), n_, d_, max_d, ext_.data(), &ne, nullptr, 0));
    ext_.resize(ne);
    uint64_t ng = 0, nt = 0;
    dev_.check(sky_mdmc_hashcube(dev_.get(), &ng, &nt, nullptr, nullptr, nullptr, nullptr, 0, 0));
    std::vector<uint32_t> gw(ng ? ng : 1), gv(ng ? ng : 1);
    std::vector<uint64_t> go(ng + 1);
    ids_.resize(nt ? nt : 1);
    HostBuffer<uint32_t>& ids = ids_;
    dev_.check(sky_mdmc_hashcube(dev_.get(), &ng, &nt, gw.data(), gv.data(), go.data(), ids.data(),
                                 gw.size(), ids.size()));
    const uint32_t words = sky_words_per_point(d_) + (d_ >= 5 ? 1u : 0u);
    maps_.assign(words, HashCubeMaps());
    // the groups come sorted by (word, value): each insertion goes at the
    // end of its map
    for (uint64_t g = 0; g < ng; ++g) {
      HashCubeMaps& mp = maps_[gw[g]];
      mp.emplace_hint(mp.end(), gv[g], std::vector<int>(ids.begin() + go[g], ids.begin() + go[g + 1]));
    }
  }
  // (2^d / 32) + 1 maps, word value -> point ids (Hashcube.h:171-177).
  HashCubeMaps* getHashCubeStructure() { return maps_.data(); }
  const std::vector<uint32_t>& extendedSkyline() const { return ext_; }
  size_t words() const { return maps_.size(); }

 private:
  Config cfg_;
  uint32_t n_, d_;
  Device dev_;
  HostBuffer<float> rows_;
  HostBuffer<uint32_t> ids_;  // the HashCube's ids as they come from the device
  std::vector<uint32_t> ext_;
  std::vector<HashCubeMaps> maps_;
};

// SDSC: one SkyAlign per cuboid, top-down (SDSC.h:18-54).
class SDSC {
 public:
  SDSC(const Config& cfg, uint32_t n, uint32_t d) : cfg_(cfg), n_(n), d_(d), dev_(first_device(cfg)) {}
  void Init(float** data) {
    rows_.resize((size_t)n_ * d_);
    for (uint32_t i = 0; i < n_; ++i)
      for (uint32_t j = 0; j < d_; ++j) rows_[(size_t)i * d_ + j] = data[i][j];
  }
  void Init(const std::vector<float>& rows) {
    rows_.resize(rows.size());
    std::copy(rows.begin(), rows.end(), rows_.begin());
  }
  void Execute() {
    const uint32_t max_d = cfg_.max_d > 0 ? (uint32_t)cfg_.max_d : d_;
    const uint32_t S = 1u << d_;
    std::vector<uint64_t> so(S + 1), eo(S + 1);
    int rc = sky_sdsc(dev_.get(), rows_.data(), n_, d_, max_d, so.data(), nullptr, 0, eo.data(),
                      nullptr, 0);
    if (rc != SKY_ERR_CAPACITY) dev_.check(rc);
    std::vector<uint32_t> sky(so[S] ? so[S] : 1), ext(eo[S] ? eo[S] : 1);
    dev_.check(sky_sdsc(dev_.get(), rows_.data(), n_, d_, max_d, so.data(), sky.data(), sky.size(),
                        eo.data(), ext.data(), ext.size()));
    lattice_.assign(S, lattice_node());
    for (uint32_t s = 1; s < S; ++s) {
      lattice_node& L = lattice_[s];
      L.index = s;
      for (uint32_t j = 0; j < d_; ++j)
        if (s >> j & 1u) L.dimensions.push_back((int)j);
      L.skyline.assign(sky.begin() + (long)so[s], sky.begin() + (long)so[s + 1]);
      L.extended_sky.assign(ext.begin() + (long)eo[s], ext.begin() + (long)eo[s + 1]);
    }
  }
  lattice_node* getLattice() { return lattice_.data(); }

 private:
  Config cfg_;
  uint32_t n_, d_;
  Device dev_;
  HostBuffer<float> rows_;
  std::vector<lattice_node> lattice_;
};

// Skyline of cuboid s from a HashCube (getsubspace, utilities.h:253-267).
inline std::vector<int> getsubspace(unsigned int subspace, HashCubeMaps* hashcube) {
  std::vector<int> ret;
  const unsigned int index = subspace / 32;
  const unsigned int mask = 1u << (subspace % 32);
  for (auto it = hashcube[index].begin(); it != hashcube[index].end(); ++it)
    if (!(it->first & mask)) ret.insert(ret.end(), it->second.begin(), it->second.end());
  return ret;
}

}  // namespace skycube
