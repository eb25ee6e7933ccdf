// csv_rows.cpp -- test helper: parses a CSV file with the host driver's
// reader (skycube-templates_amd/host/skycube.hpp: read_dims, read_data) and
// writes "n d" on the first line of stdout, then the n x d FP32 rows as raw
// little-endian bytes.  No device call is made.
#include <cstdio>

#include "../../skycube-templates_amd/host/skycube.hpp"

int main(int argc, char** argv) {
  if (argc != 2) return 1;
  try {
    const int dims = skycube::read_dims(argv[1]);
    uint32_t n = 0, d = 0;
    const std::vector<float> rows = skycube::read_data(argv[1], &n, &d);
    if (n && (int)d != dims) return 3;
    std::printf("%u %u\n", n, d);
    std::fflush(stdout);
    std::fwrite(rows.data(), sizeof(float), rows.size(), stdout);
  } catch (const skycube::SkyError& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
  return 0;
}
