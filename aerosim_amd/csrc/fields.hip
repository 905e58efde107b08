// Derived-field readouts on the GPU, computed from the solver's resident macro fields:
//  - the RGBA8 cross-section slice of the viewer (velocity magnitude, vorticity magnitude or Cp through the
//    CFD colour map), by the CPU solver's rules (lbmCpuReference.ts:23-46, 500-596; FP64 like its JavaScript)
//    or by the WGSL post-process shader's rules (postProcess.wgsl:26-146; FP32);
//  - the Q-criterion, vorticity magnitude and vorticity vector of every cell with their maxima
//    (vortexIso.ts:22-119);
//  - point probes: trilinear velocity, nearest-fluid density on solid hits and the physical readout
//    (pointProbe.ts:23-92 and 284-355, streamlines.ts:20-81).
// This file is compiled with -ffp-contract=off: every operation rounds where the reference's does, so the
// FP64 results equal the reference's double arithmetic bit for bit.
#include <hip/hip_runtime.h>

#include "fields.h"

namespace aero {
namespace {

constexpr uint8_t SOLID = 1;

// ECMAScript Math.min / Math.max (NaN if either is NaN), Math.round (halves toward +infinity) and the
// Uint8Array store (NaN and infinities to 0, otherwise truncation modulo 256).
__device__ inline double jmin(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b); }
__device__ inline double jmax(double a, double b) { return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b); }
__device__ inline double jround(double x) {
  const double r = floor(x);
  return (x - r >= 0.5) ? r + 1.0 : r;
}
__device__ inline uint32_t to_u8(double v) {
  if (!(v == v) || v == __builtin_inf() || v == -__builtin_inf()) return 0;
  return (uint32_t)((long long)trunc(v) & 255);
}

__device__ inline float rho_at(const FieldView& v, size_t i) {
  const float r = v.rho[i];
  if (!v.corr || (v.flags && v.flags[i] == SOLID)) return r;
  return (float)((double)r * *v.corr);
}

__device__ inline size_t cidx(const FieldView& v, uint32_t x, uint32_t y, uint32_t z) {
  return (size_t)x + (size_t)v.nx * ((size_t)y + (size_t)v.ny * z);
}

// sampleCFDColormap (lbmCpuReference.ts:23-46)
__device__ inline void colormap64(double tv, double* rgb) {
  const double t = jmax(0.0, jmin(1.0, tv));
  if (t < 0.25) {
    const double u = t * 4.0;
    rgb[0] = 0.08 + (0.06 - 0.08) * u;
    rgb[1] = 0.22 + (0.72 - 0.22) * u;
    rgb[2] = 0.75 + (0.88 - 0.75) * u;
  } else if (t < 0.5) {
    const double u = (t - 0.25) * 4.0;
    rgb[0] = 0.06 + (0.12 - 0.06) * u;
    rgb[1] = 0.72 + (0.82 - 0.72) * u;
    rgb[2] = 0.88 + (0.38 - 0.88) * u;
  } else if (t < 0.75) {
    const double u = (t - 0.5) * 4.0;
    rgb[0] = 0.12 + (0.98 - 0.12) * u;
    rgb[1] = 0.82 + (0.62 - 0.82) * u;
    rgb[2] = 0.38 + (0.08 - 0.38) * u;
  } else {
    const double u = (t - 0.75) * 4.0;
    rgb[0] = 0.98 + (0.94 - 0.98) * u;
    rgb[1] = 0.62 + (0.16 - 0.62) * u;
    rgb[2] = 0.08 + (0.18 - 0.08) * u;
  }
}

// colormapCFD (postProcess.wgsl:27-42): mix(a, b, u) = a * (1 - u) + b * u in f32
__device__ inline float mixf(float a, float b, float u) { return a * (1.0f - u) + b * u; }
__device__ inline void colormap32(float tv, float* rgb) {
  const float t = fminf(fmaxf(tv, 0.0f), 1.0f);
  const float s[5][3] = {{0.08f, 0.22f, 0.75f}, {0.06f, 0.72f, 0.88f}, {0.12f, 0.82f, 0.38f},
                         {0.98f, 0.62f, 0.08f}, {0.94f, 0.16f, 0.18f}};
  int seg;
  float u;
  if (t < 0.25f) { seg = 0; u = t * 4.0f; }
  else if (t < 0.5f) { seg = 1; u = (t - 0.25f) * 4.0f; }
  else if (t < 0.75f) { seg = 2; u = (t - 0.5f) * 4.0f; }
  else { seg = 3; u = (t - 0.75f) * 4.0f; }
  for (int k = 0; k < 3; ++k) rgb[k] = mixf(s[seg][k], s[seg + 1][k], u);
}
__device__ inline uint32_t pack8(float x) { return (uint32_t)fminf(fmaxf(x * 255.0f, 0.0f), 255.0f); }

__global__ void k_slice_rgba(FieldView v, int axis, uint32_t fixed, int field, int rules, double u_ref, float u_ref32,
                             uint32_t w, uint32_t h, uint32_t* rgba) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w * h) return;
  const uint32_t u = p % w, vv = p / w;
  uint32_t ix = u, iy = vv, iz = fixed;
  if (axis == 1) { ix = u; iy = fixed; iz = vv; }
  else if (axis == 2) { ix = fixed; iy = vv; iz = u; }
  const size_t idx = cidx(v, ix, iy, iz);
  if (v.flags && v.flags[idx] == SOLID) {
    rgba[p] = 0;
    return;
  }
  const uint32_t ixP = min(v.nx - 1, ix + 1), ixM = ix > 0 ? ix - 1 : 0;
  const uint32_t iyP = min(v.ny - 1, iy + 1), iyM = iy > 0 ? iy - 1 : 0;
  const uint32_t izP = min(v.nz - 1, iz + 1), izM = iz > 0 ? iz - 1 : 0;
  const size_t XP = cidx(v, ixP, iy, iz), XM = cidx(v, ixM, iy, iz), YP = cidx(v, ix, iyP, iz),
               YM = cidx(v, ix, iyM, iz), ZP = cidx(v, ix, iy, izP), ZM = cidx(v, ix, iy, izM);
  if (rules == SLICE_RULES_CPU) {
    double scalar = 0.5;
    uint32_t alpha = 215;
    if (field == 0) {
      const double x = v.ux[idx], y = v.uy[idx], z = v.uz[idx];
      const double speed = sqrt(x * x + y * y + z * z);
      scalar = jmin(1.0, jmax(0.0, speed / (u_ref * 1.55)));
    } else if (field == 1) {
      const double wx = 0.5 * ((double)v.uz[YP] - (double)v.uz[YM]) - 0.5 * ((double)v.uy[ZP] - (double)v.uy[ZM]);
      const double wy = 0.5 * ((double)v.ux[ZP] - (double)v.ux[ZM]) - 0.5 * ((double)v.uz[XP] - (double)v.uz[XM]);
      const double wz = 0.5 * ((double)v.uy[XP] - (double)v.uy[XM]) - 0.5 * ((double)v.ux[YP] - (double)v.ux[YM]);
      const double mag = sqrt(wx * wx + wy * wy + wz * wz);
      const double vn = jmin(1.0, mag / (u_ref * 0.42));
      scalar = 0.35 + 0.65 * vn;
      alpha = to_u8(jround(70 + 175 * vn));
    } else {
      const double cp = (((double)rho_at(v, idx) - 1.0) / 3.0) / (0.5 * u_ref * u_ref);
      scalar = jmin(1.0, jmax(0.0, (cp + 1.1) / 2.1));
    }
    double rgb[3];
    colormap64(scalar, rgb);
    rgba[p] = to_u8(jround(rgb[0] * 255)) | (to_u8(jround(rgb[1] * 255)) << 8) |
              (to_u8(jround(rgb[2] * 255)) << 16) | (alpha << 24);
  } else {
    float scalar = 0.5f, alpha = 0.82f;
    if (field == 0) {
      const float x = v.ux[idx], y = v.uy[idx], z = v.uz[idx];
      const float ratio = sqrtf(x * x + y * y + z * z) / u_ref32;
      scalar = fminf(fmaxf(ratio / 1.6f, 0.0f), 1.0f);
    } else if (field == 1) {
      const float dUzDy = 0.5f * (v.uz[YP] - v.uz[YM]);
      const float dUyDz = 0.5f * (v.uy[ZP] - v.uy[ZM]);
      const float dUxDz = 0.5f * (v.ux[ZP] - v.ux[ZM]);
      const float dUzDx = 0.5f * (v.uz[XP] - v.uz[XM]);
      const float dUyDx = 0.5f * (v.uy[XP] - v.uy[XM]);
      const float dUxDy = 0.5f * (v.ux[YP] - v.ux[YM]);
      const float cx = dUzDy - dUyDz, cy = dUxDz - dUzDx, cz = dUyDx - dUxDy;
      const float vn = sqrtf(cx * cx + cy * cy + cz * cz) / (u_ref32 * 0.45f);
      scalar = fminf(fmaxf(0.35f + 0.65f * vn, 0.0f), 1.0f);
      alpha = fminf(fmaxf(0.25f + 0.70f * vn, 0.25f), 0.92f);
    } else {
      const float cp = ((rho_at(v, idx) - 1.0f) / 3.0f) / (0.5f * u_ref32 * u_ref32);
      scalar = fminf(fmaxf((cp + 1.1f) / 2.1f, 0.0f), 1.0f);
    }
    float rgb[3];
    colormap32(scalar, rgb);
    rgba[p] = pack8(rgb[0]) | (pack8(rgb[1]) << 8) | (pack8(rgb[2]) << 16) | (pack8(alpha) << 24);
  }
}

constexpr int QB = 256;  // threads per block of the Q-criterion kernels

__device__ inline double axis_scale(uint32_t prev, uint32_t next, double inv_dx, double inv_2dx) {
  // 1 / ((next - prev) * dx) with next - prev in {1, 2}, or 1 / (2 dx) for a one-cell axis
  return next > prev ? (next - prev == 2 ? inv_2dx : inv_dx) : inv_2dx;
}

__device__ inline double wave_max(double x) {
  for (int o = 32; o > 0; o >>= 1) {
    const double y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return x;
}

// Tile of the Q-criterion kernel: 256 x-columns (one per thread) by QTY rows, marching over QTZ planes. Each
// thread keeps the velocity of planes z-1, z, z+1 for its QTY rows and one halo row on either side in
// registers, so a plane row is read from memory once per tile (the halo rows of neighbouring tiles mostly come
// from L2); x neighbours come from the adjacent lanes.
constexpr int QTY = 1;
constexpr int QTZ = 16;

struct Planes {
  float u[3][QTY + 2];  // component, row (index 0 is row y0 - 1, clamped)
};

__device__ inline void load_plane(const FieldView& v, Planes& p, uint32_t y0, uint32_t z, uint32_t xl) {
  const float* comp[3] = {v.ux, v.uy, v.uz};
#pragma unroll
  for (int i = 0; i < QTY + 2; ++i) {
    int row = (int)y0 + i - 1;
    row = row < 0 ? 0 : (row > (int)v.ny - 1 ? (int)v.ny - 1 : row);
    const size_t a = cidx(v, xl, (uint32_t)row, z);
#pragma unroll
    for (int c = 0; c < 3; ++c) p.u[c][i] = comp[c][a];
  }
}

__global__ void __launch_bounds__(QB) k_qcrit(FieldView v, double inv_dx, double inv_2dx, float* q, float* om,
                                              float* ov, double* part) {
  double mq = 1e-6, mo = 1e-6;
  const uint32_t ntx = (v.nx + QB - 1) / QB, nty = (v.ny + QTY - 1) / QTY;
  // blocks on one XCD (ids equal modulo 8) take consecutive tiles, so neighbouring tiles share its L2
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  uint32_t t = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
  const uint32_t tx = t % ntx;
  t /= ntx;
  const uint32_t y0 = (t % nty) * QTY, z0 = (t / nty) * QTZ;
  const uint32_t z1 = min(v.nz, z0 + QTZ);
  const uint32_t x = tx * QB + threadIdx.x;
  const bool live_x = x < v.nx;
  const uint32_t xl = live_x ? x : v.nx - 1;  // columns past the end read column nx - 1: the clamped neighbour
  const int lane = threadIdx.x & 63;
  const uint32_t xP = x > 0 ? min(x, v.nx) - 1 : 0, xN = min(v.nx - 1, x + 1);
  const double sx = axis_scale(xP, xN, inv_dx, inv_2dx);
  const float* comp[3] = {v.ux, v.uy, v.uz};
  // planes z-1, z, z+1 (issuing the next plane's loads before computing gave no gain: the kernel is bound by
  // its FP64 instruction stream, not by load latency)
  Planes pm, p0, pp;
  load_plane(v, pm, y0, z0 > 0 ? z0 - 1 : 0, xl);
  load_plane(v, p0, y0, z0, xl);
  for (uint32_t z = z0; z < z1; ++z) {
    load_plane(v, pp, y0, min(v.nz - 1, z + 1), xl);
    const uint32_t zP = z > 0 ? z - 1 : 0, zN = min(v.nz - 1, z + 1);
    const double sz = axis_scale(zP, zN, inv_dx, inv_2dx);
#pragma unroll
    for (int i = 0; i < QTY; ++i) {
      const uint32_t y = y0 + i;
      // x neighbours of row y from the adjacent lanes; the wave's edge lanes read them
      float xm[3], xp[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xm[c] = __shfl_up(p0.u[c][i + 1], 1);
        xp[c] = __shfl_down(p0.u[c][i + 1], 1);
      }
      if (y >= v.ny || !live_x) continue;
      if (lane == 0)
        for (int c = 0; c < 3; ++c) xm[c] = comp[c][cidx(v, xP, y, z)];
      if (lane == 63)
        for (int c = 0; c < 3; ++c) xp[c] = comp[c][cidx(v, xN, y, z)];
      const size_t idx = cidx(v, x, y, z);
      if (v.flags && v.flags[idx] == SOLID) {
        q[idx] = 0.f;
        om[idx] = 0.f;
        ov[3 * idx] = 0.f;
        ov[3 * idx + 1] = 0.f;
        ov[3 * idx + 2] = 0.f;
        continue;
      }
      const uint32_t yP = y > 0 ? y - 1 : 0, yN = min(v.ny - 1, y + 1);
      const double sy = axis_scale(yP, yN, inv_dx, inv_2dx);
      // d/dx from (x+1) - (x-1), d/dy from rows i+2 and i, d/dz from planes z+1 and z-1
      const double dux_dx = ((double)xp[0] - (double)xm[0]) * sx;
      const double dux_dy = ((double)p0.u[0][i + 2] - (double)p0.u[0][i]) * sy;
      const double dux_dz = ((double)pp.u[0][i + 1] - (double)pm.u[0][i + 1]) * sz;
      const double duy_dx = ((double)xp[1] - (double)xm[1]) * sx;
      const double duy_dy = ((double)p0.u[1][i + 2] - (double)p0.u[1][i]) * sy;
      const double duy_dz = ((double)pp.u[1][i + 1] - (double)pm.u[1][i + 1]) * sz;
      const double duz_dx = ((double)xp[2] - (double)xm[2]) * sx;
      const double duz_dy = ((double)p0.u[2][i + 2] - (double)p0.u[2][i]) * sy;
      const double duz_dz = ((double)pp.u[2][i + 1] - (double)pm.u[2][i + 1]) * sz;
      const double S11 = dux_dx, S22 = duy_dy, S33 = duz_dz;
      const double S12 = 0.5 * (dux_dy + duy_dx);
      const double S13 = 0.5 * (dux_dz + duz_dx);
      const double S23 = 0.5 * (duy_dz + duz_dy);
      const double normS = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
      const double O12 = 0.5 * (dux_dy - duy_dx);
      const double O13 = 0.5 * (dux_dz - duz_dx);
      const double O23 = 0.5 * (duy_dz - duz_dy);
      const double normO = 2.0 * (O12 * O12 + O13 * O13 + O23 * O23);
      const double qv = 0.5 * (normO - normS);
      const double ox = duz_dy - duy_dz, oy = dux_dz - duz_dx, oz = duy_dx - dux_dy;
      const double mag = sqrt(ox * ox + oy * oy + oz * oz);
      q[idx] = (float)qv;
      om[idx] = (float)mag;
      ov[3 * idx] = (float)ox;
      ov[3 * idx + 1] = (float)oy;
      ov[3 * idx + 2] = (float)oz;
      if (qv > mq) mq = qv;   // NaN never raises a maximum, as in the reference's comparison
      if (mag > mo) mo = mag;
    }
    pm = p0;
    p0 = pp;
  }
  // maxima of values >= 1e-6 are independent of the order of comparison
  __shared__ double sq[QB / 64], so[QB / 64];
  mq = wave_max(mq);
  mo = wave_max(mo);
  if ((threadIdx.x & 63) == 0) {
    sq[threadIdx.x >> 6] = mq;
    so[threadIdx.x >> 6] = mo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < QB / 64; ++i) {
      mq = sq[i] > mq ? sq[i] : mq;
      mo = so[i] > mo ? so[i] : mo;
    }
    part[2 * blockIdx.x] = mq;
    part[2 * blockIdx.x + 1] = mo;
  }
}

__global__ void __launch_bounds__(QB) k_qcrit_final(const double* part, uint32_t n, double* out) {
  double mq = 1e-6, mo = 1e-6;
  for (uint32_t i = threadIdx.x; i < n; i += QB) {
    mq = part[2 * i] > mq ? part[2 * i] : mq;
    mo = part[2 * i + 1] > mo ? part[2 * i + 1] : mo;
  }
  __shared__ double sq[QB / 64], so[QB / 64];
  mq = wave_max(mq);
  mo = wave_max(mo);
  if ((threadIdx.x & 63) == 0) {
    sq[threadIdx.x >> 6] = mq;
    so[threadIdx.x >> 6] = mo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < QB / 64; ++i) {
      mq = sq[i] > mq ? sq[i] : mq;
      mo = so[i] > mo ? so[i] : mo;
    }
    out[0] = mq;
    out[1] = mo;
  }
}

// ---- fluid mass statistics (computeFluidMassStats, lbmCpuReference.ts:461-495) ---------------------------------
// The reference walks the cells in index order, skips solid cells, stops at the first cell whose rho or velocity is
// NaN, and sums rho, counts the cells and keeps the largest speed over the cells before it. Each block takes a chunk
// of MCHUNK consecutive cells (MCPT per thread, strided by the block size so that loads coalesce), finds the chunk's
// first NaN cell, and reduces over its non-solid cells before that one; the final kernel keeps the chunks up to the
// one holding the first NaN of the grid. Sums are FP64 in a fixed order; maxima and counts are exact.
constexpr int MB = 256, MCPT = 8, MCHUNK = MB * MCPT;
constexpr double NO_NAN = 4294967295.0;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(v, o, 64);
    v = t < v ? t : v;
  }
  return v;
}

// part[4 b .. 4 b + 3] = rho sum, cell count, largest speed, index of the first NaN cell (NO_NAN if none), all over
// the non-solid cells of chunk b that come before its first NaN cell.
__global__ void __launch_bounds__(MB) k_mass_partial(FieldView v, uint32_t N, double* part) {
  const uint32_t base = blockIdx.x * (uint32_t)MCHUNK;
  double r[MCPT], u[MCPT], w[MCPT], z[MCPT];
  bool live[MCPT];
  uint32_t first = 0xFFFFFFFFu;
#pragma unroll
  for (int c = 0; c < MCPT; ++c) {
    const uint32_t i = base + (uint32_t)c * MB + threadIdx.x;
    live[c] = i < N && !(v.flags && v.flags[i] == SOLID);
    r[c] = u[c] = w[c] = z[c] = 0.0;
    if (live[c]) {
      r[c] = rho_at(v, i);
      u[c] = v.ux[i];
      w[c] = v.uy[i];
      z[c] = v.uz[i];
      if ((r[c] != r[c] || u[c] != u[c] || w[c] != w[c] || z[c] != z[c]) && i < first) first = i;
    }
  }
  __shared__ uint32_t smin[MB / 64];
  __shared__ double ssum[MB / 64], scnt[MB / 64], smax[MB / 64];
  first = wave_min_u32(first);
  if ((threadIdx.x & 63) == 0) smin[threadIdx.x >> 6] = first;
  __syncthreads();
  for (int k = 0; k < MB / 64; ++k) first = smin[k] < first ? smin[k] : first;
  double sum = 0.0, cnt = 0.0, mx = 0.0;
#pragma unroll
  for (int c = 0; c < MCPT; ++c) {
    const uint32_t i = base + (uint32_t)c * MB + threadIdx.x;
    if (live[c] && i < first) {
      sum += r[c];
      cnt += 1.0;
      const double spd = sqrt(u[c] * u[c] + w[c] * w[c] + z[c] * z[c]);
      mx = spd > mx ? spd : mx;
    }
  }
  sum = wave_sum(sum);
  cnt = wave_sum(cnt);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) {
    ssum[threadIdx.x >> 6] = sum;
    scnt[threadIdx.x >> 6] = cnt;
    smax[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < MB / 64; ++k) {
      sum += ssum[k];
      cnt += scnt[k];
      mx = smax[k] > mx ? smax[k] : mx;
    }
    part[4 * (size_t)blockIdx.x] = sum;
    part[4 * (size_t)blockIdx.x + 1] = cnt;
    part[4 * (size_t)blockIdx.x + 2] = mx;
    part[4 * (size_t)blockIdx.x + 3] = first == 0xFFFFFFFFu ? NO_NAN : (double)first;
  }
}

// out: total mass, mean rho (1 when no cell counted), largest speed, NaN flag, counted cells.
__global__ void __launch_bounds__(1024) k_mass_final(const double* part, uint32_t nb, double* out) {
  __shared__ double sa[16], sb[16], sc[16];
  // first NaN of the grid
  double f = NO_NAN;
  for (uint32_t b = threadIdx.x; b < nb; b += 1024) f = part[4 * (size_t)b + 3] < f ? part[4 * (size_t)b + 3] : f;
  f = -wave_max(-f);
  if ((threadIdx.x & 63) == 0) sa[threadIdx.x >> 6] = f;
  __syncthreads();
  for (int k = 0; k < 16; ++k) f = sa[k] < f ? sa[k] : f;
  __syncthreads();
  const uint32_t last = f == NO_NAN ? nb - 1 : (uint32_t)(f / MCHUNK);
  double sum = 0.0, cnt = 0.0, mx = 0.0;
  for (uint32_t b = threadIdx.x; b <= last; b += 1024) {
    sum += part[4 * (size_t)b];
    cnt += part[4 * (size_t)b + 1];
    mx = part[4 * (size_t)b + 2] > mx ? part[4 * (size_t)b + 2] : mx;
  }
  sum = wave_sum(sum);
  cnt = wave_sum(cnt);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) {
    sa[threadIdx.x >> 6] = sum;
    sb[threadIdx.x >> 6] = cnt;
    sc[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 16; ++k) {
      sum += sa[k];
      cnt += sb[k];
      mx = sc[k] > mx ? sc[k] : mx;
    }
    out[0] = sum;
    out[1] = cnt > 0.0 ? sum / cnt : 1.0;
    out[2] = mx;
    out[3] = f == NO_NAN ? 0.0 : 1.0;
    out[4] = cnt;
  }
}

// JavaScript `x || d` on a float read: d for 0, -0 and NaN
__device__ inline double or_default(float x, double d) { return (x == 0.0f || x != x) ? d : (double)x; }

// sampleVelocityTrilinear (streamlines.ts:20-81)
__device__ bool trilinear(const FieldView& v, double fx, double fy, double fz, double* o) {
  if (fx < 0 || fx >= (double)v.nx - 1 || fy < 0 || fy >= (double)v.ny - 1 || fz < 0 || fz >= (double)v.nz - 1) {
    o[0] = o[1] = o[2] = 0.0;
    return false;
  }
  const double fix = floor(fx), fiy = floor(fy), fiz = floor(fz);
  const uint32_t ix = (uint32_t)fix, iy = (uint32_t)fiy, iz = (uint32_t)fiz;
  const double tx = fx - fix, ty = fy - fiy, tz = fz - fiz;
  const size_t i000 = cidx(v, ix, iy, iz), i100 = cidx(v, ix + 1, iy, iz), i010 = cidx(v, ix, iy + 1, iz),
               i110 = cidx(v, ix + 1, iy + 1, iz), i001 = cidx(v, ix, iy, iz + 1), i101 = cidx(v, ix + 1, iy, iz + 1),
               i011 = cidx(v, ix, iy + 1, iz + 1), i111 = cidx(v, ix + 1, iy + 1, iz + 1);
  const float* comp[3] = {v.ux, v.uy, v.uz};
  for (int k = 0; k < 3; ++k) {
    const float* a = comp[k];
    const double c00 = (double)a[i000] * (1 - tx) + (double)a[i100] * tx;
    const double c10 = (double)a[i010] * (1 - tx) + (double)a[i110] * tx;
    const double c01 = (double)a[i001] * (1 - tx) + (double)a[i101] * tx;
    const double c11 = (double)a[i011] * (1 - tx) + (double)a[i111] * tx;
    const double c0 = c00 * (1 - ty) + c10 * ty;
    const double c1 = c01 * (1 - ty) + c11 * ty;
    o[k] = c0 * (1 - tz) + c1 * tz;
  }
  return true;
}

__global__ void k_probe(FieldView v, ProbeParams pp, const double* xyz, const uint8_t* surface, uint32_t n,
                        double* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double fx = (xyz[3 * i] - pp.origin[0]) / pp.dx - 0.5;
  const double fy = (xyz[3 * i + 1] - pp.origin[1]) / pp.dx - 0.5;
  const double fz = (xyz[3 * i + 2] - pp.origin[2]) / pp.dx - 0.5;
  const int ix = (int)jmax(0.0, jmin((double)v.nx - 1, jround(fx)));
  const int iy = (int)jmax(0.0, jmin((double)v.ny - 1, jround(fy)));
  const int iz = (int)jmax(0.0, jmin((double)v.nz - 1, jround(fz)));
  const size_t cell = cidx(v, ix, iy, iz);
  const bool cell_solid = v.flags && v.flags[cell] == SOLID;
  const bool solid = (surface && surface[i]) || cell_solid;
  double vel[3];
  const bool sampled = trilinear(v, jmax(0.0, jmin((double)v.nx - 1.001, fx)), jmax(0.0, jmin((double)v.ny - 1.001, fy)),
                                 jmax(0.0, jmin((double)v.nz - 1.001, fz)), vel);
  if ((!sampled || (vel[0] == 0 && vel[1] == 0 && vel[2] == 0)) && !solid) {
    vel[0] = or_default(v.ux[cell], 0.0);
    vel[1] = or_default(v.uy[cell], 0.0);
    vel[2] = or_default(v.uz[cell], 0.0);
  }
  double rho = or_default(rho_at(v, cell), 1.0);
  if (solid && cell_solid) {
    const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int k = 0; k < 6; ++k) {
      const int x = ix + nb[k][0], y = iy + nb[k][1], z = iz + nb[k][2];
      if (x >= 0 && x < (int)v.nx && y >= 0 && y < (int)v.ny && z >= 0 && z < (int)v.nz) {
        const size_t j = cidx(v, x, y, z);
        if (v.flags[j] == 0) {
          rho = or_default(rho_at(v, j), rho);
          break;
        }
      }
    }
  }
  // computeProbePhysics (pointProbe.ts:23-92)
  const double qdyn = 0.5 * pp.fluid_density * pp.wind_speed_mps * pp.wind_speed_mps;
  const double qlat = 0.5 * 1.0 * pp.u_inlet_lattice * pp.u_inlet_lattice;
  const double cp = ((1.0 / 3.0) * (rho - 1.0)) / jmax(1e-6, qlat);
  const double gauge = cp * qdyn;
  double* o = out + (size_t)13 * i;
  o[0] = ix;
  o[1] = iy;
  o[2] = iz;
  if (solid) {
    o[3] = o[4] = o[5] = o[6] = o[7] = 0.0;
  } else {
    const double scale = pp.wind_speed_mps / jmax(1e-4, pp.u_inlet_lattice);
    const double ux = vel[0] * scale, uy = vel[1] * scale, uz = vel[2] * scale;
    const double speed = sqrt(ux * ux + uy * uy + uz * uz);
    o[3] = speed;
    o[4] = speed * 2.23693629;
    o[5] = ux;
    o[6] = uy;
    o[7] = uz;
  }
  o[8] = gauge;
  o[9] = cp;
  o[10] = rho;
  o[11] = qdyn;
  o[12] = solid ? 1.0 : 0.0;
}

}  // namespace

hipError_t launch_slice_rgba(const FieldView& v, int axis, uint32_t fixed, int field, int rules, double u_ref,
                             uint32_t* rgba, hipStream_t s) {
  const uint32_t w = axis == 2 ? v.nz : v.nx;
  const uint32_t h = axis == 1 ? v.nz : v.ny;
  const uint32_t n = w * h;
  // the CPU rules take max(1e-4, uInlet) in double, the shader in f32 on the f32 uniform
  const double ur = u_ref > 1e-4 ? u_ref : 1e-4;
  const float u32v = (float)u_ref;
  const float ur32 = u32v > 1e-4f ? u32v : 1e-4f;
  hipLaunchKernelGGL(k_slice_rgba, dim3((n + 255) / 256), dim3(256), 0, s, v, axis, fixed, field, rules, ur, ur32, w,
                     h, rgba);
  return hipGetLastError();
}

uint32_t qcrit_blocks(const FieldView& v) {
  return ((v.nx + QB - 1) / QB) * ((v.ny + QTY - 1) / QTY) * ((v.nz + QTZ - 1) / QTZ);
}

hipError_t launch_qcrit(const FieldView& v, double inv_dx, double inv_2dx, float* q, float* om, float* ov,
                        double* part, double* out, hipStream_t s) {
  const uint32_t nb = qcrit_blocks(v);
  hipLaunchKernelGGL(k_qcrit, dim3(nb), dim3(QB), 0, s, v, inv_dx, inv_2dx, q, om, ov, part);
  hipLaunchKernelGGL(k_qcrit_final, dim3(1), dim3(QB), 0, s, part, nb, out);
  return hipGetLastError();
}

uint32_t mass_stats_blocks(uint32_t N) { return (N + MCHUNK - 1) / MCHUNK; }

hipError_t launch_mass_stats(const FieldView& v, double* part, double* out, hipStream_t s) {
  const uint32_t N = v.nx * v.ny * v.nz;
  const uint32_t nb = mass_stats_blocks(N);
  hipLaunchKernelGGL(k_mass_partial, dim3(nb), dim3(MB), 0, s, v, N, part);
  hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(1024), 0, s, part, nb, out);
  return hipGetLastError();
}

hipError_t launch_probe(const FieldView& v, const ProbeParams& p, const double* xyz, const uint8_t* surface,
                        uint32_t n, double* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_probe, dim3((n + 127) / 128), dim3(128), 0, s, v, p, xyz, surface, n, out);
  return hipGetLastError();
}

}  // namespace aero
