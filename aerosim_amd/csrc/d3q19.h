// D3Q19 lattice tables shared by the host and device code of the solver.
// Velocity order, opposite table and weights follow src/solver/lbmConstants.ts:10-63 of the reference;
// the order is part of the population layout that callers read back, so it is kept identical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero {

constexpr int Q = 19;
__host__ __device__ constexpr int cx(int d) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  return t[d];
}
__host__ __device__ constexpr int cy(int d) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
  return t[d];
}
__host__ __device__ constexpr int cz(int d) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
  return t[d];
}
__host__ __device__ constexpr int opp(int d) { return d == 0 ? 0 : (d & 1 ? d + 1 : d - 1); }
// Double weights 1/3, 1/18, 1/36 and their float rounding (LBM_WEIGHTS_F32).
__host__ __device__ constexpr double w64(int d) { return d == 0 ? 1.0 / 3.0 : (d <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }
__host__ __device__ constexpr float w32(int d) { return (float)w64(d); }

constexpr uint8_t CELL_FLUID = 0, CELL_SOLID = 1, CELL_BOUNDARY = 2;

// Domain faces and the 5 directions that point into the domain from each face. The halo that stores
// face-cell populations for the reference's clamped pulls keeps exactly these 5 per face.
enum Face { FX0 = 0, FX1 = 1, FY0 = 2, FY1 = 3, FZ0 = 4, FZ1 = 5 };
__host__ __device__ constexpr int face_dir(int face, int s) {
  constexpr int t[6][5] = {{1, 7, 9, 11, 13},  {2, 8, 10, 12, 14}, {3, 7, 10, 15, 17},
                           {4, 8, 9, 16, 18},  {5, 11, 14, 15, 18}, {6, 12, 13, 16, 17}};
  return t[face][s];
}
// slot of direction d among the inward directions of face (or -1).
__host__ __device__ constexpr int face_slot(int face, int d) {
  return face_dir(face, 0) == d ? 0
       : face_dir(face, 1) == d ? 1
       : face_dir(face, 2) == d ? 2
       : face_dir(face, 3) == d ? 3
       : face_dir(face, 4) == d ? 4
                                : -1;
}

// Counter hash of the seeded initial velocity noise (splitmix64 finaliser); uniform in [-1, 1).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double noise_unit(uint64_t seed, uint64_t idx, int comp) {
  const uint64_t h = mix64(seed ^ mix64(idx * 4ull + (uint64_t)comp));
  return ((double)(h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}

}  // namespace aero
