// main.cpp -- the `skycube` command: the reference driver's flags and output
// (src/main.cu:66-177, src/main.h:152-356) over the MI355X engine.
//
//   skycube -f <file.csv> -s mdmc|sdsc [-m max_d] [-g dev[,dev..]] [-c] ...
//
// Prints the milliseconds of the Execute() window followed by a tab, like the
// reference (src/main.h:324-346).  -c checks the result: the reference
// compares against its STSC template; here MDMC and SDSC, two independent
// device algorithms, are compared cuboid by cuboid (sizes and members), and
// each mismatch is printed.  -w <file> (not a reference flag) writes the
// result after the timed window: the HashCube maps of mdmc or the lattice's
// sky / ext lists of sdsc, as little-endian uint32 records (write_result).
#include <getopt.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <iostream>

#include "skycube.hpp"

using namespace skycube;

static void printUsage() {
  printf("\nskycube - the MI355X skycube engine (MDMC and SDSC templates)\n\n");
  printf("USAGE: ./skycube -f <filename> -s <algorithm>\n");
  printf(" -f: input filename\n");
  printf(" -t: run with num_threads (accepted; the device does the work)\n");
  printf(" -s: skycube algorithm to run\n");
  printf("     Supported algorithms: [mdmc sdsc]\n");
  printf(" -a: alpha block size (accepted)\n");
  printf(" -d: depth of mdmc search tree (accepted)\n");
  printf(" -q: priority queue size (accepted)\n");
  printf(" -c: check result (MDMC against SDSC, cuboid by cuboid)\n");
  printf(" -p: do not use cpus in sdsc and mdmc (always the case here)\n");
  printf(" -g: gpu device numbers to use in a comma separated list (the first is used)\n");
  printf(" -m: maximum amount of dimensions in each subspace of result\n");
  printf(" -i: custom papi counters (not supported: use rocprofv3)\n");
  printf(" -o: print work distribution for cross device computation\n");
  printf(" -w: write the result to a file (HashCube maps or lattice lists)\n\n");
  printf("Example: ./skycube -f ../workloads/house.csv -s mdmc\n\n");
}

// Result file of -w: "SKYC", d, kind (0 HashCube, 1 lattice), then for a
// HashCube the number of maps and per map its group count and per group
// (value, id count, ids); for a lattice, per cuboid s = 1 .. 2^d - 1, the
// sky count and ids, then the ext count and ids.
static void put(std::FILE* f, uint32_t v) { std::fwrite(&v, sizeof(v), 1, f); }

static bool write_hashcube(const std::string& path, uint32_t d, HashCubeMaps* maps, size_t words) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fwrite("SKYC", 1, 4, f);
  put(f, d);
  put(f, 0);
  put(f, (uint32_t)words);
  for (size_t j = 0; j < words; ++j) {
    put(f, (uint32_t)maps[j].size());
    for (const auto& kv : maps[j]) {
      put(f, kv.first);
      put(f, (uint32_t)kv.second.size());
      for (int id : kv.second) put(f, (uint32_t)id);
    }
  }
  return std::fclose(f) == 0;
}

static bool write_lattice(const std::string& path, uint32_t d, const lattice_node* lat) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fwrite("SKYC", 1, 4, f);
  put(f, d);
  put(f, 1);
  for (uint32_t s = 1; s < (1u << d); ++s) {
    put(f, (uint32_t)lat[s].skyline.size());
    for (unsigned id : lat[s].skyline) put(f, id);
    put(f, (uint32_t)lat[s].extended_sky.size());
    for (unsigned id : lat[s].extended_sky) put(f, id);
  }
  return std::fclose(f) == 0;
}

static void int_split(const char* s, std::vector<int>* out) {
  std::string str(s), tok;
  std::stringstream ss(str);
  while (std::getline(ss, tok, ',')) out->push_back(atoi(tok.c_str()));
}

static uint64_t now_ms() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::milliseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int main(int argc, char** argv) {
  Config cfg;
  std::string out_path;
  int c = 0;
  opterr = 0;
  while ((c = getopt(argc, argv, "f:t:s:a:d:q:g:m:i:pcow:")) != -1) {
    switch (c) {
      case 'f': cfg.input_fname = optarg; break;
      case 't': cfg.threads = atoi(optarg); break;
      case 's': cfg.algo = optarg; break;
      case 'a': cfg.alpha_size = atoi(optarg); break;
      case 'd': cfg.partitioning_depth = atoi(optarg); break;
      case 'q': cfg.pq_size = atoi(optarg); break;
      case 'g': int_split(optarg, &cfg.devices); break;
      case 'm': cfg.max_d = atoi(optarg); break;
      case 'p': cfg.cpu = false; break;
      case 'c': cfg.check = true; break;
      case 'o': cfg.cross_distribution = true; break;
      case 'i': cfg.papi = optarg; break;
      case 'w': out_path = optarg; break;
      default:
        if (isprint(optopt)) fprintf(stderr, "Unknown option `-%c'.\n", optopt);
        printUsage();
        return 1;
    }
  }
  if (cfg.input_fname.empty() || cfg.algo.empty()) {
    printUsage();
    return 1;
  }
  if (cfg.algo != "mdmc" && cfg.algo != "sdsc") {
    fprintf(stderr, "algorithm '%s' is not provided by this engine (mdmc, sdsc)\n", cfg.algo.c_str());
    return 1;
  }
  if (cfg.devices.size() > 1)
    fprintf(stderr, "note: one device is used (device %d)\n", cfg.devices[0]);
  if (!cfg.papi.empty()) fprintf(stderr, "note: -i is not supported; use rocprofv3\n");
  try {
    const int d = read_dims(cfg.input_fname);
    if (d < 2 || d > 16) {
      printf("not supported yet, add it in the main method and template instantiations!\n");
      return 0;
    }
    if (cfg.max_d == -1) cfg.max_d = d;
    uint32_t n = 0, dd = 0;
    const std::vector<float> rows = read_data(cfg.input_fname, &n, &dd);
    const uint32_t S = 1u << dd;
    uint64_t q_sta = 0, q_end = 0;
    if (cfg.algo == "mdmc") {
      MDMC sky(cfg, n, dd);
      sky.Init(rows);
      q_sta = now_ms();
      sky.Execute();
      q_end = now_ms();
      if (cfg.cross_distribution) printf("GPU%d\n1.000000\n", first_device(cfg));
      if (!out_path.empty() &&
          !write_hashcube(out_path, dd, sky.getHashCubeStructure(), sky.words()))
        throw SkyError(SKY_ERR_ARG, "cannot write " + out_path);
      if (cfg.check) {
        SDSC ref(cfg, n, dd);
        ref.Init(rows);
        ref.Execute();
        printf("checking\n");
        for (uint32_t s = 1; s < S; ++s) {
          if (__builtin_popcount(s) > cfg.max_d) continue;
          std::vector<int> a = getsubspace(s, sky.getHashCubeStructure());
          std::vector<unsigned int> b = ref.getLattice()[s].skyline;
          std::sort(a.begin(), a.end());
          const bool same = a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin(),
                                                               [](int x, unsigned y) { return (unsigned)x == y; });
          if (!same) printf("In arrays, truth: %lu, in: %lu\n", b.size(), a.size());
        }
        printf("done checking\n");
      }
    } else {
      SDSC sky(cfg, n, dd);
      sky.Init(rows);
      q_sta = now_ms();
      sky.Execute();
      q_end = now_ms();
      if (cfg.cross_distribution) printf("GPU%d\n1.000000\n", first_device(cfg));
      if (!out_path.empty() && !write_lattice(out_path, dd, sky.getLattice()))
        throw SkyError(SKY_ERR_ARG, "cannot write " + out_path);
      if (cfg.check) {
        MDMC ref(cfg, n, dd);
        ref.Init(rows);
        ref.Execute();
        printf("checking\n");
        for (uint32_t s = 1; s < S; ++s) {
          if (__builtin_popcount(s) > cfg.max_d) continue;
          std::vector<int> b = getsubspace(s, ref.getHashCubeStructure());
          std::vector<unsigned int> a = sky.getLattice()[s].skyline;
          std::sort(b.begin(), b.end());
          const bool same = a.size() == b.size() && std::equal(a.begin(), a.end(), b.begin(),
                                                               [](unsigned x, int y) { return x == (unsigned)y; });
          if (!same) printf("In arrays, truth: %lu, in: %lu\n", b.size(), a.size());
        }
        printf("done checking\n");
      }
    }
    std::cout << (q_end - q_sta) << "\t";
    std::cout.flush();
  } catch (const SkyError& e) {
    fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 2;
  }
  return 0;
}
