// Host-code fuzz harness, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitizers.py: the STL parser and normaliser (untrusted bytes) and the mesh statistics and
// frontal-area raster (user meshes) of aerosim_amd/csrc/stl.cpp and geometry.cpp, fed seeded random, truncated
// and mutated binary and ASCII STL files, degenerate and transformed meshes and the built-in mesh generators.
// Prints "host fuzz OK <cases>" and exits 0 when no sanitizer fired (a finding aborts with its report).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"

namespace aero {
// The library defines this in its C-ABI file, which needs the HIP runtime; the harness keeps the last message.
static std::string g_err;
void set_host_error(const std::string& m) { g_err = m; }
}  // namespace aero

static std::mt19937_64 rng(20261204);
static double uni(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static uint32_t u32(uint32_t n) { return (uint32_t)(rng() % (n ? n : 1)); }

static void put_f32(std::vector<uint8_t>& b, float v) {
  uint8_t t[4];
  std::memcpy(t, &v, 4);
  b.insert(b.end(), t, t + 4);
}

static std::vector<uint8_t> binary_stl(uint32_t ntri, uint32_t declared, bool solid_header) {
  std::vector<uint8_t> b(80, ' ');
  if (solid_header) std::memcpy(b.data(), "solid x", 7);
  uint8_t c[4];
  std::memcpy(c, &declared, 4);
  b.insert(b.end(), c, c + 4);
  for (uint32_t i = 0; i < ntri; ++i) {
    for (int k = 0; k < 12; ++k) {
      const int r = (int)u32(20);
      float v = (float)uni(-2.0, 2.0);
      if (r == 0) v = NAN;
      if (r == 1) v = INFINITY;
      if (r == 2) v = 1e38f;
      if (r == 3) v = 0.0f;
      put_f32(b, v);
    }
    b.push_back((uint8_t)u32(256));
    b.push_back((uint8_t)u32(256));
  }
  return b;
}

static std::vector<uint8_t> ascii_stl(int ntri) {
  std::string s = "solid fuzz\n";
  const char* nums[] = {"1", "-2.5", "3.", ".5", "1e3", "-1E-2", "nan", "inf", "1e400", "0x10", "+7", "--1", "1.2.3", ""};
  for (int i = 0; i < ntri; ++i) {
    s += " facet normal 0 0 1\n  outer loop\n";
    const int nv = 3 + ((u32(10) == 0) ? (int)u32(3) - 1 : 0);  // sometimes 2 or 4 vertices
    for (int v = 0; v < nv; ++v) {
      s += "   vertex";
      for (int k = 0; k < 3; ++k) {
        s += u32(3) ? " " : "\t";
        s += u32(4) ? std::to_string(uni(-5, 5)) : nums[u32(14)];
      }
      s += "\n";
    }
    s += "  endloop\n endfacet\n";
  }
  s += "endsolid fuzz\n";
  return std::vector<uint8_t>(s.begin(), s.end());
}

static void mutate(std::vector<uint8_t>& b) {
  const int n = 1 + (int)u32(8);
  for (int i = 0; i < n && !b.empty(); ++i) {
    switch (u32(4)) {
      case 0: b[u32((uint32_t)b.size())] = (uint8_t)u32(256); break;
      case 1: b.resize(u32((uint32_t)b.size() + 1)); break;
      case 2: b.insert(b.begin() + u32((uint32_t)b.size()), (uint8_t)u32(256)); break;
      default: b.erase(b.begin() + u32((uint32_t)b.size())); break;
    }
  }
}

static long cases = 0;

static void stats_of(const std::vector<float>& t, int32_t ntri) {
  double out[6];
  double world[16] = {0};
  for (int i = 0; i < 16; ++i) world[i] = (i % 5 == 0) ? 1.0 : 0.0;
  aero_mesh_stats(t.data(), ntri, nullptr, out);
  for (int i = 0; i < 12; ++i) world[u32(16)] = uni(-3, 3);  // arbitrary, possibly singular or projective
  aero_mesh_stats(t.data(), ntri, world, out);
  ++cases;
}

static void parse_and_use(const std::vector<uint8_t>& b) {
  int64_t n = 0;
  const int rc = aero_stl_parse(b.data(), (int64_t)b.size(), nullptr, 0, &n);
  ++cases;
  if (rc != AERO_OK || n <= 0 || n > 2000000) return;
  std::vector<float> t((size_t)n * 9);
  int64_t m = 0;
  if (aero_stl_parse(b.data(), (int64_t)b.size(), t.data(), n, &m) != AERO_OK) return;
  if (m > 1) {  // a too-small capacity must be refused, not overrun
    std::vector<float> small((size_t)(m - 1) * 9);
    int64_t k = 0;
    if (aero_stl_parse(b.data(), (int64_t)b.size(), small.data(), m - 1, &k) == AERO_OK) {
      std::fprintf(stderr, "parse accepted a capacity below the count\n");
      std::abort();
    }
  }
  int32_t mm = 0;
  aero_stl_normalize(t.data(), m, &mm);
  stats_of(t, (int32_t)m);
}

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? std::atoi(argv[1]) : 300;
  const int gen_rounds = argc > 2 ? std::atoi(argv[2]) : 8;
  for (int r = 0; r < rounds; ++r) {
    // binary: exact, truncated, padded, lying counts, "solid" headers
    const uint32_t nt = u32(40);
    uint32_t decl = nt;
    switch (u32(5)) {
      case 0: decl = nt + 1 + u32(5); break;
      case 1: decl = 0xFFFFFFFFu; break;
      case 2: decl = nt ? nt - 1 : 0; break;
      default: break;
    }
    auto b = binary_stl(nt, decl, u32(3) == 0);
    parse_and_use(b);
    if (u32(2)) b.resize(u32((uint32_t)b.size() + 1));
    mutate(b);
    parse_and_use(b);
    auto a = ascii_stl((int)u32(12));
    parse_and_use(a);
    mutate(a);
    parse_and_use(a);
    std::vector<uint8_t> junk(u32(300));
    for (auto& c : junk) c = (uint8_t)u32(256);
    parse_and_use(junk);
    parse_and_use(std::vector<uint8_t>());
  }
  // generated meshes, including degenerate sizes and segment counts, through the statistics
  for (int r = 0; r < gen_rounds; ++r) {
    int32_t n = 0;
    const double c[3] = {uni(-1, 1), uni(-1, 1), uni(-1, 1)};
    const int lvl = (int)u32(4);
    if (aero_mesh_icosphere(lvl, uni(0, 3), c, nullptr, 0, &n) == AERO_OK && n > 0) {
      std::vector<float> t((size_t)n * 9);
      aero_mesh_icosphere(lvl, uni(0, 3), c, t.data(), n, &n);
      stats_of(t, n);
    }
    const int seg = (int)u32(80);
    if (aero_mesh_rocket(c, uni(0, 50), uni(0, 10), uni(0, 20), uni(0, 0.1), rng(), seg, nullptr, 0, &n) == AERO_OK &&
        n > 0) {
      std::vector<float> t((size_t)n * 9);
      aero_mesh_rocket(c, 40, 6, 15, 0.02, 7, seg, t.data(), n, &n);
      stats_of(t, n);
    }
    const int np = (int)u32(60);
    if (aero_mesh_naca(uni(0, 0.3), uni(0, 50), c, uni(-30, 30), -5, 5, np, nullptr, 0, &n) == AERO_OK && n > 0) {
      std::vector<float> t((size_t)n * 9);
      aero_mesh_naca(0.12, 20, c, 5, -5, 5, np, t.data(), n, &n);
      stats_of(t, n);
    }
    const double lo[3] = {uni(-2, 0), uni(-2, 0), uni(-2, 0)}, hi[3] = {uni(-1, 2), uni(-1, 2), uni(-1, 2)};
    if (aero_mesh_box(lo, hi, nullptr, 0, &n) == AERO_OK && n > 0) {
      std::vector<float> t((size_t)n * 9);
      aero_mesh_box(lo, hi, t.data(), n, &n);
      stats_of(t, n);
    }
  }
  std::printf("host fuzz OK %ld\n", cases);
  return 0;
}
