// GPU voxelizer (replaces voxelizeMeshToLattice, voxelizer.ts:86-281, and the MeshBVH queries it uses,
// bvh.ts:265-589). FP64 throughout; this file is compiled with -ffp-contract=off so every product and
// sum rounds exactly as the reference's double arithmetic does, which keeps the solid flags bit-exact.
//
//   k_scan   one thread per (iy, iz) column of the mesh bounding box: sorted +x scanline hits; an even
//            count fills cell centres between entry/exit pairs, an odd count falls back to the 3-ray
//            majority vote per cell (voxelizer.ts:126-173).
//   k_mark   fluid cells with an in-domain solid D3Q19 neighbour become boundary cells (voxelizer.ts:176-212).
//   compaction of the boundary cells in cell-index order (the reference's iz, iy, ix loop order) with
//            hipcub::DeviceSelect, then k_bmap.
//   k_q      per boundary cell and solid link, the segment intersection distance q (voxelizer.ts:217-256).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "d3q19.h"
#include "geometry.h"

namespace aero {
namespace {

constexpr int HIT_CAP = 64;
constexpr int STACK = 64;

struct DevBVH {
  const double* __restrict__ tri;
  const double* __restrict__ node;
  const int32_t* __restrict__ lf;
  const int32_t* __restrict__ cnt;
  const int32_t* __restrict__ idx;
  int32_t ntri;
  double bmin[3], bmax[3];
};

struct VoxGeo {
  int32_t nx, ny, nz;
  double dx, ox, oy, oz;
  int32_t minIx, maxIx, minIy, maxIy, minIz, maxIz;
};

// std::min / std::max semantics (the oracle's), including the choice between equal operands.
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// bvh.ts:375-457. Returns the hit count (sorted in hits[0..min(n, HIT_CAP))); sets *overflow when more
// than HIT_CAP hits were found.
__device__ int scan_hits(const DevBVH& B, double y, double z, double* hits, int* overflow) {
  if (B.ntri == 0) return 0;
  const double py = y + 1.37e-8, pz = z + 2.93e-8;
  if (py < B.bmin[1] || py > B.bmax[1] || pz < B.bmin[2] || pz > B.bmax[2]) return 0;
  int stack[STACK];
  int sp = 0, n = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int nd = stack[--sp];
    const double* nb = B.node + (size_t)nd * 6;
    if (py < nb[1] || py > nb[4] || pz < nb[2] || pz > nb[5]) continue;
    const int c = B.cnt[nd];
    if (c > 0) {
      const int first = B.lf[nd];
      for (int i = first; i < first + c; ++i) {
        const double* T = B.tri + (size_t)B.idx[i] * 9;
        const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
        const double e2x = T[6], e2y = T[7], e2z = T[8];
        const double det = -e1y * e2z + e1z * e2y;
        if (fabs(det) < 1e-13) continue;
        const double invDet = 1.0 / det;
        const double ty = py - ay, tz = pz - az;
        const double u = (-ty * e2z + tz * e2y) * invDet;
        if (u < 0.0 || u > 1.0) continue;
        const double v = (ty * e1z - tz * e1y) * invDet;
        if (v < 0.0 || u + v > 1.0) continue;
        const double xh = ax + u * e1x + v * e2x;
        if (n < HIT_CAP) {
          int j = n;  // insertion keeps hits sorted ascending
          while (j > 0 && hits[j - 1] > xh) {
            hits[j] = hits[j - 1];
            --j;
          }
          hits[j] = xh;
        } else {
          *overflow = 1;
        }
        ++n;
      }
    } else {
      if (sp + 2 > STACK) { *overflow = 1; return n; }
      const int left = B.lf[nd];
      stack[sp++] = left + 1;
      stack[sp++] = left;
    }
  }
  return n;
}

// bvh.ts:499-589
__device__ int count_ray(const DevBVH& B, double ox, double oy, double oz, double dx, double dy, double dz, int* overflow) {
  const double invDx = 1.0 / dx, invDy = 1.0 / dy, invDz = 1.0 / dz;
  int hitCount = 0;
  int stack[STACK];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int nd = stack[--sp];
    const double* nb = B.node + (size_t)nd * 6;
    const double tx1 = (nb[0] - ox) * invDx, tx2 = (nb[3] - ox) * invDx;
    double tmin = dmin(tx1, tx2), tmax = dmax(tx1, tx2);
    const double ty1 = (nb[1] - oy) * invDy, ty2 = (nb[4] - oy) * invDy;
    tmin = dmax(tmin, dmin(ty1, ty2));
    tmax = dmin(tmax, dmax(ty1, ty2));
    const double tz1 = (nb[2] - oz) * invDz, tz2 = (nb[5] - oz) * invDz;
    tmin = dmax(tmin, dmin(tz1, tz2));
    tmax = dmin(tmax, dmax(tz1, tz2));
    if (tmax < dmax(0.0, tmin)) continue;
    const int c = B.cnt[nd];
    if (c > 0) {
      const int first = B.lf[nd];
      for (int i = first; i < first + c; ++i) {
        const double* T = B.tri + (size_t)B.idx[i] * 9;
        const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
        const double e2x = T[6], e2y = T[7], e2z = T[8];
        const double px = dy * e2z - dz * e2y;
        const double py = dz * e2x - dx * e2z;
        const double pz = dx * e2y - dy * e2x;
        const double det = e1x * px + e1y * py + e1z * pz;
        if (fabs(det) < 1e-13) continue;
        const double invDet = 1.0 / det;
        const double tx = ox - ax, ty = oy - ay, tz = oz - az;
        const double u = (tx * px + ty * py + tz * pz) * invDet;
        if (u < 0.0 || u > 1.0) continue;
        const double qx = ty * e1z - tz * e1y;
        const double qy = tz * e1x - tx * e1z;
        const double qz = tx * e1y - ty * e1x;
        const double v = (dx * qx + dy * qy + dz * qz) * invDet;
        if (v < 0.0 || u + v > 1.0) continue;
        const double t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
        if (t > 1e-7) ++hitCount;
      }
    } else {
      if (sp + 2 > STACK) { *overflow = 1; return hitCount; }
      const int left = B.lf[nd];
      stack[sp++] = left + 1;
      stack[sp++] = left;
    }
  }
  return hitCount;
}

// bvh.ts:464-497
__device__ bool inside(const DevBVH& B, double x, double y, double z, int* overflow) {
  if (x < B.bmin[0] || x > B.bmax[0] || y < B.bmin[1] || y > B.bmax[1] || z < B.bmin[2] || z > B.bmax[2]) return false;
  const double dirs[3][3] = {{1.0, 0.0137, 0.0293}, {-0.0191, 1.0, 0.0317}, {0.0239, -0.0173, 1.0}};
  int votes = 0;
  for (int d = 0; d < 3; ++d) {
    const int h = count_ray(B, x, y, z, dirs[d][0], dirs[d][1], dirs[d][2], overflow);
    if ((h & 1) == 1) ++votes;
    if (votes >= 2) return true;
    if (d == 1 && votes == 0) return false;
  }
  return votes >= 2;
}

// bvh.ts:265-367
__device__ double segment(const DevBVH& B, double ox, double oy, double oz, double dx, double dy, double dz, double maxT,
                          int* overflow) {
  if (B.ntri == 0) return -1;
  const double invDx = fabs(dx) > 1e-15 ? 1.0 / dx : (dx >= 0 ? 1e15 : -1e15);
  const double invDy = fabs(dy) > 1e-15 ? 1.0 / dy : (dy >= 0 ? 1e15 : -1e15);
  const double invDz = fabs(dz) > 1e-15 ? 1.0 / dz : (dz >= 0 ? 1e15 : -1e15);
  double closestT = maxT + 1e-7;
  bool found = false;
  int stack[STACK];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int nd = stack[--sp];
    const double* nb = B.node + (size_t)nd * 6;
    const double tx1 = (nb[0] - ox) * invDx, tx2 = (nb[3] - ox) * invDx;
    double tmin = dmin(tx1, tx2), tmax = dmax(tx1, tx2);
    const double ty1 = (nb[1] - oy) * invDy, ty2 = (nb[4] - oy) * invDy;
    tmin = dmax(tmin, dmin(ty1, ty2));
    tmax = dmin(tmax, dmax(ty1, ty2));
    const double tz1 = (nb[2] - oz) * invDz, tz2 = (nb[5] - oz) * invDz;
    tmin = dmax(tmin, dmin(tz1, tz2));
    tmax = dmin(tmax, dmax(tz1, tz2));
    if (tmax < dmax(0.0, tmin) || tmin > closestT) continue;
    const int c = B.cnt[nd];
    if (c > 0) {
      const int first = B.lf[nd];
      for (int i = first; i < first + c; ++i) {
        const double* T = B.tri + (size_t)B.idx[i] * 9;
        const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
        const double e2x = T[6], e2y = T[7], e2z = T[8];
        const double px = dy * e2z - dz * e2y;
        const double py = dz * e2x - dx * e2z;
        const double pz = dx * e2y - dy * e2x;
        const double det = e1x * px + e1y * py + e1z * pz;
        if (fabs(det) < 1e-13) continue;
        const double invDet = 1.0 / det;
        const double tx = ox - ax, ty = oy - ay, tz = oz - az;
        const double u = (tx * px + ty * py + tz * pz) * invDet;
        if (u < -1e-7 || u > 1.0 + 1e-7) continue;
        const double qx = ty * e1z - tz * e1y;
        const double qy = tz * e1x - tx * e1z;
        const double qz = tx * e1y - ty * e1x;
        const double v = (dx * qx + dy * qy + dz * qz) * invDet;
        if (v < -1e-7 || u + v > 1.0 + 1e-7) continue;
        const double t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
        if (t > 1e-7 && t <= closestT) {
          closestT = t;
          found = true;
        }
      }
    } else {
      if (sp + 2 > STACK) { *overflow = 1; return -1; }
      const int left = B.lf[nd];
      stack[sp++] = left + 1;
      stack[sp++] = left;
    }
  }
  return found ? dmin(maxT, closestT) : -1;
}

__global__ void k_scan(DevBVH B, VoxGeo v, uint8_t* __restrict__ flags, unsigned long long* __restrict__ solid,
                       int* __restrict__ overflow) {
  const int ncy = v.maxIy - v.minIy + 1, ncz = v.maxIz - v.minIz + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncy * ncz) return;
  const int iy = v.minIy + t % ncy, iz = v.minIz + t / ncy;
  const double z = v.oz + (iz + 0.5) * v.dx;
  const double y = v.oy + (iy + 0.5) * v.dx;
  double hits[HIT_CAP];
  int of = 0;
  const int hc = scan_hits(B, y, z, hits, &of);
  if (of) { atomicOr(overflow, 1); return; }
  if (hc == 0) return;
  uint8_t* row = flags + (size_t)v.nx * (iy + (size_t)v.ny * iz);
  unsigned long long cnt = 0;
  if ((hc & 1) == 0) {
    for (int k = 0; k < hc; k += 2) {
      const double xe = hits[k], xx = hits[k + 1];
      if (xx - xe < 1e-7) continue;
      const int s = max(v.minIx, (int)ceil((xe - v.ox) / v.dx - 0.5));
      const int e = min(v.maxIx, (int)floor((xx - v.ox) / v.dx - 0.5));
      for (int ix = s; ix <= e; ++ix)
        if (row[ix] == CELL_FLUID) { row[ix] = CELL_SOLID; ++cnt; }
    }
  } else {
    for (int ix = v.minIx; ix <= v.maxIx; ++ix) {
      const double x = v.ox + (ix + 0.5) * v.dx;
      if (x >= hits[0] && x <= hits[hc - 1] && inside(B, x, y, z, &of))
        if (row[ix] == CELL_FLUID) { row[ix] = CELL_SOLID; ++cnt; }
    }
    if (of) atomicOr(overflow, 1);
  }
  if (cnt) atomicAdd(solid, cnt);
}

// Boundary marking over the search box (voxelizer.ts:176-212). Only SOLID is tested, so marking in place
// is order-independent.
__global__ void k_mark(VoxGeo v, int sx0, int sy0, int sz0, int bx, int by, int bz, uint8_t* __restrict__ flags,
                       uint8_t* __restrict__ sel) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)bx * by * bz) return;
  const int ix = sx0 + (int)(t % bx);
  const int iy = sy0 + (int)((t / bx) % by);
  const int iz = sz0 + (int)(t / ((long long)bx * by));
  const size_t idx = ix + (size_t)v.nx * (iy + (size_t)v.ny * iz);
  if (flags[idx] == CELL_SOLID) return;
  bool has = false;
  for (int d = 1; d <= 18 && !has; ++d) {
    const int a = ix + cx(d), b = iy + cy(d), c = iz + cz(d);
    if (a < 0 || a >= v.nx || b < 0 || b >= v.ny || c < 0 || c >= v.nz) continue;
    if (flags[a + (size_t)v.nx * (b + (size_t)v.ny * c)] == CELL_SOLID) has = true;
  }
  if (has) sel[idx] = 1;
}

__global__ void k_apply_boundary(const int32_t* __restrict__ blist, const int* __restrict__ nb, uint8_t* __restrict__ flags,
                                 int32_t* __restrict__ bmap) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= *nb) return;
  flags[blist[b]] = CELL_BOUNDARY;
  bmap[blist[b]] = b;
}

__global__ void k_q(DevBVH B, VoxGeo v, const uint8_t* __restrict__ flags, const int32_t* __restrict__ blist, int nb,
                    float* __restrict__ q, float* __restrict__ minq, int* __restrict__ overflow) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int cell = blist[b];
  const int ix = cell % v.nx;
  const int rem = (cell - ix) / v.nx;
  const int iy = rem % v.ny;
  const int iz = (rem - iy) / v.ny;
  const double fx = v.ox + (ix + 0.5) * v.dx, fy = v.oy + (iy + 0.5) * v.dx, fz = v.oz + (iz + 0.5) * v.dx;
  double mq = 1.0;
  int of = 0;
  for (int d = 1; d <= 18; ++d) {
    float out = -1.0f;
    const int a = ix + cx(d), bb = iy + cy(d), c = iz + cz(d);
    if (a >= 0 && a < v.nx && bb >= 0 && bb < v.ny && c >= 0 && c < v.nz &&
        flags[a + (size_t)v.nx * (bb + (size_t)v.ny * c)] == CELL_SOLID) {
      const double t = segment(B, fx, fy, fz, cx(d) * v.dx, cy(d) * v.dx, cz(d) * v.dx, 1.0, &of);
      double qv = 0.5;
      if (t > 0 && t <= 1.0) qv = dmax(1e-4, dmin(1.0, t));
      out = (float)qv;
      if (qv < mq) mq = qv;
    }
    q[(size_t)b * 18 + (d - 1)] = out;
  }
  minq[b] = (float)mq;
  if (of) atomicOr(overflow, 1);
}

struct Flagger {
  __host__ __device__ bool operator()(const uint8_t& s) const { return s != 0; }
};

}  // namespace
}  // namespace aero

using namespace aero;

#define VX_HIP(call)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) {                                                                      \
      c->err = std::string("aero_set_mesh: ") + #call + ": " + hipGetErrorString(e_);           \
      return AERO_E_HIP;                                                                         \
    }                                                                                            \
  } while (0)

extern "C" int aero_set_mesh(aero_ctx* c, const float* tri_pos, int32_t ntri, const double* world,
                             const aero_tunnel* tunnel) {
  if (!c) return AERO_E_ARG;
  if (!tri_pos || ntri <= 0 || !tunnel || !(tunnel->height > 0)) {
    c->err = "aero_set_mesh: need a non-empty triangle array and a tunnel height";
    return AERO_E_ARG;
  }
  const auto t0 = std::chrono::steady_clock::now();
  VX_HIP(hipSetDevice(c->dev));
  {
    const int rc = c->quiesce();  // the side stream may still read the geometry buffers replaced below
    if (rc != AERO_OK) return rc;
  }
  hipStream_t s = c->stream;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t N = c->N;
  std::vector<double> w;
  world_positions(tri_pos, ntri, world, w);
  HostBVH hb;
  build_bvh(w.data(), ntri, hb);
  if (hb.depth + 2 > STACK / 2) {
    c->err = "aero_set_mesh: BVH deeper than the traversal stack";
    return AERO_E_ARG;
  }
  // lattice geometry (voxelizer.ts:101-121)
  VoxGeo v;
  v.nx = nx; v.ny = ny; v.nz = nz;
  const double h = tunnel->height / ny;
  v.dx = h;
  v.ox = -nx * h * 0.5; v.oy = -ny * h * 0.5; v.oz = -nz * h * 0.5;
  v.minIx = std::max(0, (int)std::floor((hb.bmin[0] - v.ox) / h) - 1);
  v.maxIx = std::min(nx - 1, (int)std::ceil((hb.bmax[0] - v.ox) / h) + 1);
  v.minIy = std::max(0, (int)std::floor((hb.bmin[1] - v.oy) / h) - 1);
  v.maxIy = std::min(ny - 1, (int)std::ceil((hb.bmax[1] - v.oy) / h) + 1);
  v.minIz = std::max(0, (int)std::floor((hb.bmin[2] - v.oz) / h) - 1);
  v.maxIz = std::min(nz - 1, (int)std::ceil((hb.bmax[2] - v.oz) / h) + 1);

  VoxScratch& X = c->vox;
  double* d_tri = (double*)X.get(0, hb.tri.size() * 8);
  double* d_node = (double*)X.get(1, (size_t)hb.nnodes * 6 * 8);
  int32_t* d_lf = (int32_t*)X.get(2, (size_t)hb.nnodes * 4);
  int32_t* d_cnt = (int32_t*)X.get(3, (size_t)hb.nnodes * 4);
  int32_t* d_idx = (int32_t*)X.get(4, (size_t)ntri * 4);
  unsigned long long* d_solid = (unsigned long long*)X.get(5, 16);
  int* d_misc = (int*)X.get(6, 16);  // [0] overflow, [1] boundary count
  uint8_t* d_sel = (uint8_t*)X.get(7, N);
  if (!d_tri || !d_node || !d_lf || !d_cnt || !d_idx || !d_solid || !d_misc || !d_sel) {
    c->err = "aero_set_mesh: scratch allocation failed";
    return AERO_E_NOMEM;
  }
  VX_HIP(hipMemcpyAsync(d_tri, hb.tri.data(), hb.tri.size() * 8, hipMemcpyHostToDevice, s));
  VX_HIP(hipMemcpyAsync(d_node, hb.node.data(), (size_t)hb.nnodes * 6 * 8, hipMemcpyHostToDevice, s));
  VX_HIP(hipMemcpyAsync(d_lf, hb.lf.data(), (size_t)hb.nnodes * 4, hipMemcpyHostToDevice, s));
  VX_HIP(hipMemcpyAsync(d_cnt, hb.cnt.data(), (size_t)hb.nnodes * 4, hipMemcpyHostToDevice, s));
  VX_HIP(hipMemcpyAsync(d_idx, hb.idx.data(), (size_t)ntri * 4, hipMemcpyHostToDevice, s));
  VX_HIP(hipMemsetAsync(d_solid, 0, 16, s));
  VX_HIP(hipMemsetAsync(d_misc, 0, 16, s));
  VX_HIP(hipMemsetAsync(c->flags, 0, N, s));
  VX_HIP(hipMemsetAsync(d_sel, 0, N, s));
  VX_HIP(hipMemsetAsync(c->bmap, 0xff, N * 4, s));
  DevBVH B;
  B.tri = d_tri; B.node = d_node; B.lf = d_lf; B.cnt = d_cnt; B.idx = d_idx; B.ntri = ntri;
  for (int k = 0; k < 3; ++k) { B.bmin[k] = hb.bmin[k]; B.bmax[k] = hb.bmax[k]; }

  const int ncol = (v.maxIy - v.minIy + 1) * (v.maxIz - v.minIz + 1);
  if (ncol > 0 && v.maxIx >= v.minIx) {
    hipLaunchKernelGGL(k_scan, dim3((ncol + 63) / 64), dim3(64), 0, s, B, v, c->flags, d_solid, d_misc);
    VX_HIP(hipGetLastError());
  }
  const int sx0 = std::max(0, v.minIx - 1), sx1 = std::min(nx - 1, v.maxIx + 1);
  const int sy0 = std::max(0, v.minIy - 1), sy1 = std::min(ny - 1, v.maxIy + 1);
  const int sz0 = std::max(0, v.minIz - 1), sz1 = std::min(nz - 1, v.maxIz + 1);
  const int bx = sx1 - sx0 + 1, by = sy1 - sy0 + 1, bz = sz1 - sz0 + 1;
  const long long nbox = (bx > 0 && by > 0 && bz > 0) ? (long long)bx * by * bz : 0;
  if (nbox > 0) {
    hipLaunchKernelGGL(k_mark, dim3((unsigned)((nbox + 255) / 256)), dim3(256), 0, s, v, sx0, sy0, sz0, bx, by, bz,
                       c->flags, d_sel);
    VX_HIP(hipGetLastError());
  }
  // order-preserving compaction of the marked cells
  const size_t cap = (size_t)std::max<long long>(1, nbox);
  int32_t* d_blist = (int32_t*)X.get(8, cap * 4);
  if (!d_blist) { c->err = "aero_set_mesh: scratch allocation failed"; return AERO_E_NOMEM; }
  hipcub::CountingInputIterator<int32_t> cit(0);
  size_t tmp_bytes = 0;
  VX_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, cit, d_sel, d_blist, d_misc + 1, (int)N, s));
  void* d_tmp = X.get(9, std::max<size_t>(tmp_bytes, 16));
  if (!d_tmp) { c->err = "aero_set_mesh: scratch allocation failed"; return AERO_E_NOMEM; }
  VX_HIP(hipcub::DeviceSelect::Flagged(d_tmp, tmp_bytes, cit, d_sel, d_blist, d_misc + 1, (int)N, s));
  int hm[2];
  unsigned long long hsolid = 0;
  VX_HIP(hipMemcpyAsync(hm, d_misc, 8, hipMemcpyDeviceToHost, s));
  VX_HIP(hipMemcpyAsync(&hsolid, d_solid, 8, hipMemcpyDeviceToHost, s));
  VX_HIP(hipStreamSynchronize(s));
  if (hm[0]) {
    c->err = "aero_set_mesh: scanline hit buffer or traversal stack overflow (mesh too complex per column)";
    return AERO_E_ARG;
  }
  const int nb = hm[1];
  if (nb > 0) {
    hipLaunchKernelGGL(k_apply_boundary, dim3((nb + 255) / 256), dim3(256), 0, s, d_blist, d_misc + 1, c->flags, c->bmap);
    VX_HIP(hipGetLastError());
  }
  // per-cell buffers sized to nb
  if (c->blist) hipFree(c->blist);
  if (c->bq) hipFree(c->bq);
  if (c->minq) hipFree(c->minq);
  c->blist = nullptr; c->bq = nullptr; c->minq = nullptr;
  c->B = (uint32_t)nb;
  if (nb > 0) {
    VX_HIP(hipMalloc(&c->blist, (size_t)nb * 4));
    VX_HIP(hipMalloc(&c->bq, (size_t)nb * 18 * 4));
    VX_HIP(hipMalloc(&c->minq, (size_t)nb * 4));
    VX_HIP(hipMemcpyAsync(c->blist, d_blist, (size_t)nb * 4, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_q, dim3((nb + 63) / 64), dim3(64), 0, s, B, v, c->flags, c->blist, nb, c->bq, c->minq, d_misc);
    VX_HIP(hipGetLastError());
  }
  c->h_flags.resize(N);
  VX_HIP(hipMemcpyAsync(c->h_flags.data(), c->flags, N, hipMemcpyDeviceToHost, s));
  VX_HIP(hipMemcpyAsync(hm, d_misc, 4, hipMemcpyDeviceToHost, s));
  VX_HIP(hipStreamSynchronize(s));
  if (hm[0]) {
    c->err = "aero_set_mesh: traversal stack overflow in the link-distance pass";
    return AERO_E_ARG;
  }
  c->counts[0] = (int32_t)hsolid;
  c->counts[1] = nb;
  c->counts[2] = (int32_t)(N - hsolid - nb);
  const int rc = c->upload_geometry_common();
  c->vox_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
