// Host geometry helpers shared by the C ABI and the GPU voxelizer.
#pragma once
#include <stdint.h>

#include <vector>

namespace aero {

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
