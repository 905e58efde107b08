// Host geometry helpers shared by the C ABI and the GPU voxelizer.
#pragma once
#include <stdint.h>

#include <cmath>
#include <string>
#include <vector>

namespace aero {

// ECMAScript Math.min / Math.max / Math.round on doubles: NaN if either operand is NaN, -0 below +0,
// and halves rounded toward +infinity.
inline double js_min(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return std::signbit(a) ? a : b;
  return a < b ? a : b;
}
inline double js_max(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return std::signbit(a) ? b : a;
  return a > b ? a : b;
}
inline double js_round(double x) {
  const double r = std::floor(x);
  return (x - r >= 0.5) ? r + 1.0 : r;
}

// Message reported by aero_last_error(NULL) after a failed context-free call.
void set_host_error(const std::string& msg);

struct HostBVH {
  int32_t ntri = 0, nnodes = 0, depth = 0;
  std::vector<double> tri;   // per triangle: A, E1 = B - A, E2 = C - A
  std::vector<double> node;  // per node: min xyz, max xyz
  std::vector<int32_t> lf;   // leaf: first index into idx; inner: left child (right = left + 1)
  std::vector<int32_t> cnt;  // leaf triangle count, 0 for inner nodes
  std::vector<int32_t> idx;  // triangle order
  double bmin[3], bmax[3];
};

void world_positions(const float* tri, int32_t ntri, const double* world, std::vector<double>& out);
double frontal_area(const double* w, int32_t ntri, int gridRes);
void build_bvh(const double* w, int32_t ntri, HostBVH& b);

}  // namespace aero
