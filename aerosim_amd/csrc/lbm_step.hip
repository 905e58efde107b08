// D3Q19 wind-tunnel step kernels for gfx950: fused A-A stream-collide with regularized BGK,
// Smagorinsky LES, Bouzidi bounce-back, the reference's domain boundaries and the global mass anchor.
//
// One thread per cell. Threads at ix = 0 and ix = nx-1 do nothing: the inlet cell of a row is written by
// the thread of ix = 1 (it needs rho at ix = 1, lbmCpuReference.ts:344) and the outlet cell by the
// thread of ix = nx-2 (it needs the post-collision populations of ix = nx-2, lbmCpuReference.ts:355-371),
// so the whole step needs no barrier. The mass anchor (lbmCpuReference.ts:431-455) is a global
// reduction: each block writes a FP64 partial sum of post-boundary rho, k_finalize sums the partials and stores
// the correction factor, and the next step multiplies every value it loads by it. The sums are exact, whatever
// their order: each density summed is a float in [0.5, 1.8] (the collision clamps rho, lbmCpuReference.ts:233,
// and the inlet and outlet densities 0.85 rho + 0.15 stay inside), so a multiple of 2^-24, and FP64 holds every
// multiple of 2^-24 below 2^29, which bounds the total up to 298 M cells. Every form of the step (per-cell or
// split, any cells per thread, batched or alone) therefore finds the same anchor factor.
#include <hipcub/hipcub.hpp>

#include "lbm_kernels.h"
#include "lbm_launch.h"

namespace aero {

// Layout displacement of the neighbour x + c_k for a compile-time k, from two uniform values (the 19-entry
// table would keep 19 64-bit values live in scalar registers across the collision).
template <int K>
__device__ __forceinline__ int64_t loff_of(const Geo& g) {
  return (int64_t)cx(K) + (int64_t)cy(K) * g.lrow + (int64_t)cz(K) * g.lplane;
}

constexpr int BS = 256;
// the partials buffer holds one value per BS groups: bulk blocks are at least that large
static_assert(bulk_block_threads(2) >= BS && bulk_block_threads(4) >= BS, "bulk block size");

template <typename T>
__device__ __forceinline__ typename T::elem* fptr(const StepArgs& a) { return reinterpret_cast<typename T::elem*>(a.f); }

// ---- halo helpers -----------------------------------------------------------------------------
__device__ __forceinline__ uint32_t face_index(const Geo& g, int face, uint32_t x, uint32_t y, uint32_t z) {
  return face < 2 ? y + g.ny * z : (face < 4 ? x + g.nx * z : x + g.nx * y);
}
// Post-collision value f*_d at the clamped cell (x, y, z), which lies on `face`; d points into the
// domain from that face.
// Offset of the slot of direction d within face's block. The face index is a runtime value, so the
// arguments' arrays are read through a select chain rather than a dynamic index (which would force the
// kernel argument block into private memory).
__device__ __forceinline__ uint32_t face_base(const Geo& g, int face, int d) {
  uint32_t off, sz;
  switch (face) {
    case FX0: off = g.hoff[FX0]; sz = g.hsz[FX0]; break;
    case FX1: off = g.hoff[FX1]; sz = g.hsz[FX1]; break;
    case FY0: off = g.hoff[FY0]; sz = g.hsz[FY0]; break;
    case FY1: off = g.hoff[FY1]; sz = g.hsz[FY1]; break;
    case FZ0: off = g.hoff[FZ0]; sz = g.hsz[FZ0]; break;
    default: off = g.hoff[FZ1]; sz = g.hsz[FZ1]; break;
  }
  return off + (uint32_t)face_slot(face, d) * sz;
}
__device__ __forceinline__ float halo_get(const Geo& g, const float* h, int face, int d, uint32_t x, uint32_t y,
                                          uint32_t z) {
  return h[face_base(g, face, d) + face_index(g, face, x, y, z)];
}
template <int FACE>
__device__ __forceinline__ void halo_put_face(const Geo& g, float* h, const float* v, uint32_t x, uint32_t y,
                                              uint32_t z) {
  const uint32_t base = g.hoff[FACE] + face_index(g, FACE, x, y, z);
#pragma unroll
  for (int s = 0; s < 5; ++s) h[base + s * g.hsz[FACE]] = v[face_dir(FACE, s)];
}
// Writes the inward populations of cell (x, y, z) for every face it lies on.
__device__ __forceinline__ void halo_put(const Geo& g, float* h, const float* v, uint32_t x, uint32_t y,
                                         uint32_t z) {
  if (x == 0) halo_put_face<FX0>(g, h, v, x, y, z);
  if (x == g.nx - 1) halo_put_face<FX1>(g, h, v, x, y, z);
  if (y == 0) halo_put_face<FY0>(g, h, v, x, y, z);
  if (y == g.ny - 1) halo_put_face<FY1>(g, h, v, x, y, z);
  if (z == 0) halo_put_face<FZ0>(g, h, v, x, y, z);
  if (z == g.nz - 1) halo_put_face<FZ1>(g, h, v, x, y, z);
}
// Face of the first out-of-range coordinate of (sx, sy, sz) (x, then y, then z).
__device__ __forceinline__ int out_face(const Geo& g, int sx, int sy, int sz) {
  if (sx < 0) return FX0;
  if (sx >= (int)g.nx) return FX1;
  if (sy < 0) return FY0;
  if (sy >= (int)g.ny) return FY1;
  if (sz < 0) return FZ0;
  return FZ1;
}
__device__ __forceinline__ uint32_t clampu(int v, uint32_t n) { return v < 0 ? 0u : (v >= (int)n ? n - 1 : (uint32_t)v); }

// ---- A-A addressing ---------------------------------------------------------------------------
// Layout index of cell (x, y, z).
__device__ __forceinline__ size_t li_of(const Geo& g, uint32_t x, uint32_t y, uint32_t z) {
  return cell_li(g, x, y + g.ny * z);
}
// Incoming population k of the cell with layout index li (value f*_k(x - c_k) of the previous step).
template <typename T, bool FROM_E, int K>
__device__ __forceinline__ float ld_in(const StepArgs& a, size_t li) {
  const typename T::elem* f = fptr<T>(a);
  const size_t nx = a.g.nx;
  if (FROM_E) return T::ld(f + (li - loff_of<K>(a.g) + (size_t)opp(K) * nx), opp(K), a.ph);
  return T::ld(f + (li + (size_t)K * nx), K, a.ph);
}
// Stores f*_k of the cell with layout index li to its A-A destination (caller checks that x + c_k is inside
// for FROM_E).
template <typename T, bool FROM_E, int K>
__device__ __forceinline__ void st_out(const StepArgs& a, size_t li, float v) {
  typename T::elem* f = fptr<T>(a);
  const size_t nx = a.g.nx;
  if (FROM_E) T::st(f + (li + loff_of<K>(a.g) + (size_t)K * nx), v, K, a.ph);
  else T::st(f + (li + (size_t)opp(K) * nx), v, opp(K), a.ph);
}
// Stores all 19 post-step populations of cell (x, y, z); targets outside the domain are skipped.
template <typename T, bool FROM_E>
__device__ __forceinline__ void store_cell(const StepArgs& a, uint32_t x, uint32_t y, uint32_t z, bool check,
                                           const float* v) {
  const size_t li = li_of(a.g, x, y, z);
#define AERO_ST(K)                                                                                  \
  if (!FROM_E || !check ||                                                                          \
      ((int)x + cx(K) >= 0 && (int)x + cx(K) < (int)a.g.nx && (int)y + cy(K) >= 0 &&                \
       (int)y + cy(K) < (int)a.g.ny && (int)z + cz(K) >= 0 && (int)z + cz(K) < (int)a.g.nz))        \
    st_out<T, FROM_E, K>(a, li, v[K]);
  AERO_ST(0) AERO_ST(1) AERO_ST(2) AERO_ST(3) AERO_ST(4) AERO_ST(5) AERO_ST(6) AERO_ST(7) AERO_ST(8)
  AERO_ST(9) AERO_ST(10) AERO_ST(11) AERO_ST(12) AERO_ST(13) AERO_ST(14) AERO_ST(15) AERO_ST(16)
  AERO_ST(17) AERO_ST(18)
#undef AERO_ST
}

// f*_W(ff) with ff = x + c_D (the Bouzidi neighbour of lbmCpuReference.ts:205-212); x + c_D may be
// outside the domain, in which case the reference clamps it onto a face and the value comes from the halo.
template <typename T, bool FROM_E, int D>
__device__ __forceinline__ float ld_ff(const StepArgs& a, size_t li, uint32_t x, uint32_t y, uint32_t z) {
  constexpr int W = opp(D);
  const int fx = (int)x + cx(D), fy = (int)y + cy(D), fz = (int)z + cz(D);
  if (fx < 0 || fx >= (int)a.g.nx || fy < 0 || fy >= (int)a.g.ny || fz < 0 || fz >= (int)a.g.nz) {
    const int face = out_face(a.g, fx, fy, fz);
    return halo_get(a.g, a.halo_in, face, W, clampu(fx, a.g.nx), clampu(fy, a.g.ny), clampu(fz, a.g.nz));
  }
  const typename T::elem* f = fptr<T>(a);
  const size_t nx = a.g.nx;
  if (FROM_E) return T::ld(f + (li + loff_of<D>(a.g) + (size_t)D * nx), D, a.ph);
  return T::ld(f + (li + (size_t)W * nx), W, a.ph);
}

// Gather of a boundary or face cell: Bouzidi on links whose (clamped) pull source is solid
// (lbmCpuReference.ts:189-224), halo reads for pull sources outside the domain, in-place loads otherwise.
// Each direction loads and uses its values in turn: few registers, for kernels that visit every cell.
template <typename T, bool FROM_E>
__device__ __forceinline__ void gather_special_branchy(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                            uint8_t flag, double corr, float* f) {
  const size_t li = li_of(a.g, x, y, z);
  uint32_t mask = 0;
  int b = -1;
  if (flag == CELL_BOUNDARY) {
    b = a.bmap[idx];
    mask = a.bmask[b];
  }
  const bool sc = corr_fires(corr);
#define AERO_G(D)                                                                                         \
  {                                                                                                       \
    float v;                                                                                              \
    if (D > 0 && ((mask >> D) & 1u)) {                                                                    \
      constexpr int W = opp(D);                                                                           \
      const float qv = a.bq[(size_t)b * 18 + (W - 1)];                                                    \
      const float q = qv > 0.0f ? qv : 0.5f;                                                              \
      float fw = a.bown[(size_t)W * a.B + b];                                                             \
      if (sc) fw = apply_corr(fw, corr);                                                                  \
      if (!((mask >> W) & 1u)) {                                                                          \
        if (q < 0.5f) {                                                                                   \
          float fn = ld_ff<T, FROM_E, D>(a, li, x, y, z);                                                \
          if (sc) fn = apply_corr(fn, corr);                                                              \
          v = 2.0f * q * fw + (1.0f - 2.0f * q) * fn;                                                     \
        } else {                                                                                          \
          float fd = a.bown[(size_t)D * a.B + b];                                                         \
          if (sc) fd = apply_corr(fd, corr);                                                              \
          const float inv2q = 1.0f / (2.0f * q);                                                          \
          v = inv2q * fw + (1.0f - inv2q) * fd;                                                           \
        }                                                                                                 \
      } else {                                                                                            \
        v = fw;                                                                                           \
      }                                                                                                   \
    } else {                                                                                              \
      const int sy = (int)y - cy(D), sz = (int)z - cz(D);                                                 \
      if ((cy(D) != 0 || cz(D) != 0) && (sy < 0 || sy >= (int)a.g.ny || sz < 0 || sz >= (int)a.g.nz)) {  \
        const int face = out_face(a.g, 0, sy, sz);                                                        \
        v = halo_get(a.g, a.halo_in, face, D, (uint32_t)((int)x - cx(D)), clampu(sy, a.g.ny),             \
                     clampu(sz, a.g.nz));                                                                 \
      } else {                                                                                            \
        v = ld_in<T, FROM_E, D>(a, li);                                                                   \
      }                                                                                                   \
      if (sc) v = apply_corr(v, corr);                                                                    \
    }                                                                                                     \
    f[D] = v;                                                                                             \
  }
  AERO_G(0) AERO_G(1) AERO_G(2) AERO_G(3) AERO_G(4) AERO_G(5) AERO_G(6) AERO_G(7) AERO_G(8) AERO_G(9)
  AERO_G(10) AERO_G(11) AERO_G(12) AERO_G(13) AERO_G(14) AERO_G(15) AERO_G(16) AERO_G(17) AERO_G(18)
#undef AERO_G
}

// The same gather with every load issued first, for the kernels that visit only the listed boundary cells.
// Every value a direction may need is loaded before any is used, so the loads of all 19 directions are in flight
// together; a branch per direction around its loads and their use would wait for each direction in turn. A
// Bouzidi link loads both of its candidates (the neighbour's value, used for q < 0.5, and its own outgoing
// value, used for q >= 0.5) and selects by q afterwards; every address it forms is valid whatever q is.
template <typename T, bool FROM_E>
__device__ __forceinline__ void gather_special_eager(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                            uint8_t flag, double corr, float* f) {
  const size_t li = li_of(a.g, x, y, z);
  uint32_t mask = 0;
  int b = -1;
  if (flag == CELL_BOUNDARY) {
    b = a.bmap[idx];
    mask = a.bmask[b];
  }
  const bool sc = corr_fires(corr);
  float g0[19], g1[19], g2[19], qv[19];
#define AERO_L(D)                                                                                         \
  if (D > 0 && ((mask >> D) & 1u)) {                                                                      \
    constexpr int W = opp(D);                                                                             \
    qv[D] = a.bq[(size_t)b * 18 + (W - 1)];                                                               \
    g0[D] = a.bown[(size_t)W * a.B + b];                                                                  \
    g1[D] = ld_ff<T, FROM_E, D>(a, li, x, y, z);                                                          \
    g2[D] = a.bown[(size_t)D * a.B + b];                                                                  \
  } else {                                                                                                \
    const int sy = (int)y - cy(D), sz = (int)z - cz(D);                                                   \
    if ((cy(D) != 0 || cz(D) != 0) && (sy < 0 || sy >= (int)a.g.ny || sz < 0 || sz >= (int)a.g.nz)) {    \
      const int face = out_face(a.g, 0, sy, sz);                                                          \
      g0[D] = halo_get(a.g, a.halo_in, face, D, (uint32_t)((int)x - cx(D)), clampu(sy, a.g.ny),           \
                       clampu(sz, a.g.nz));                                                               \
    } else {                                                                                              \
      g0[D] = ld_in<T, FROM_E, D>(a, li);                                                                 \
    }                                                                                                     \
  }
  AERO_L(0) AERO_L(1) AERO_L(2) AERO_L(3) AERO_L(4) AERO_L(5) AERO_L(6) AERO_L(7) AERO_L(8) AERO_L(9)
  AERO_L(10) AERO_L(11) AERO_L(12) AERO_L(13) AERO_L(14) AERO_L(15) AERO_L(16) AERO_L(17) AERO_L(18)
#undef AERO_L
#define AERO_G(D)                                                                                         \
  {                                                                                                       \
    float v;                                                                                              \
    if (D > 0 && ((mask >> D) & 1u)) {                                                                    \
      constexpr int W = opp(D);                                                                           \
      const float q = qv[D] > 0.0f ? qv[D] : 0.5f;                                                        \
      float fw = g0[D];                                                                                   \
      if (sc) fw = apply_corr(fw, corr);                                                                  \
      if (!((mask >> W) & 1u)) {                                                                          \
        if (q < 0.5f) {                                                                                   \
          float fn = g1[D];                                                                               \
          if (sc) fn = apply_corr(fn, corr);                                                              \
          v = 2.0f * q * fw + (1.0f - 2.0f * q) * fn;                                                     \
        } else {                                                                                          \
          float fd = g2[D];                                                                               \
          if (sc) fd = apply_corr(fd, corr);                                                              \
          const float inv2q = 1.0f / (2.0f * q);                                                          \
          v = inv2q * fw + (1.0f - inv2q) * fd;                                                           \
        }                                                                                                 \
      } else {                                                                                            \
        v = fw;                                                                                           \
      }                                                                                                   \
    } else {                                                                                              \
      v = g0[D];                                                                                          \
      if (sc) v = apply_corr(v, corr);                                                                    \
    }                                                                                                     \
    f[D] = v;                                                                                             \
  }
  AERO_G(0) AERO_G(1) AERO_G(2) AERO_G(3) AERO_G(4) AERO_G(5) AERO_G(6) AERO_G(7) AERO_G(8) AERO_G(9)
  AERO_G(10) AERO_G(11) AERO_G(12) AERO_G(13) AERO_G(14) AERO_G(15) AERO_G(16) AERO_G(17) AERO_G(18)
#undef AERO_G
}

// EAGER: the load-first gather (about twice the registers), for kernels over the few listed boundary cells; the
// branchy one keeps kernels that visit every cell (the per-cell step, the learned closure's list kernel, the
// one-launch batch kernel) at their occupancy.
template <typename T, bool FROM_E, bool EAGER>
__device__ __forceinline__ void gather_special(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                            uint8_t flag, double corr, float* f) {
  if constexpr (EAGER) gather_special_eager<T, FROM_E>(a, idx, x, y, z, flag, corr, f);
  else gather_special_branchy<T, FROM_E>(a, idx, x, y, z, flag, corr, f);
}

// Pulled populations of a non-solid cell, with the previous step's mass-anchor factor applied.
template <typename T, bool FROM_E, bool EAGER = false>
__device__ __forceinline__ void gather_cell(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                            uint8_t flag, bool surf, double corr, float* f) {
  if (flag == CELL_FLUID && !surf) {
    const size_t li = li_of(a.g, x, y, z);
    f[0] = ld_in<T, FROM_E, 0>(a, li);   f[1] = ld_in<T, FROM_E, 1>(a, li);
    f[2] = ld_in<T, FROM_E, 2>(a, li);   f[3] = ld_in<T, FROM_E, 3>(a, li);
    f[4] = ld_in<T, FROM_E, 4>(a, li);   f[5] = ld_in<T, FROM_E, 5>(a, li);
    f[6] = ld_in<T, FROM_E, 6>(a, li);   f[7] = ld_in<T, FROM_E, 7>(a, li);
    f[8] = ld_in<T, FROM_E, 8>(a, li);   f[9] = ld_in<T, FROM_E, 9>(a, li);
    f[10] = ld_in<T, FROM_E, 10>(a, li); f[11] = ld_in<T, FROM_E, 11>(a, li);
    f[12] = ld_in<T, FROM_E, 12>(a, li); f[13] = ld_in<T, FROM_E, 13>(a, li);
    f[14] = ld_in<T, FROM_E, 14>(a, li); f[15] = ld_in<T, FROM_E, 15>(a, li);
    f[16] = ld_in<T, FROM_E, 16>(a, li); f[17] = ld_in<T, FROM_E, 17>(a, li);
    f[18] = ld_in<T, FROM_E, 18>(a, li);
    if (corr_fires(corr)) {
#pragma unroll
      for (int k = 0; k < 19; ++k) f[k] = apply_corr(f[k], corr);
    }
  } else {
    gather_special<T, FROM_E, EAGER>(a, idx, x, y, z, flag, corr, f);
  }
}

// Writes the final populations of a cell to every place later readers look: in-place buffer, halo for
// face cells, own-value buffer for boundary cells, readout snapshot.
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ void emit_cell(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                          uint8_t flag, bool surf, const F19 vv, float rho, float ux, float uy,
                                          float uz) {
  const float* v = vv.v;
  store_cell<T, FROM_E>(a, x, y, z, surf, v);
  if (surf) halo_put(a.g, a.halo_out, v, x, y, z);
  if (flag == CELL_BOUNDARY) {
    const int b = a.bmap[idx];
#pragma unroll
    for (int k = 0; k < 19; ++k) a.bown[(size_t)k * a.B + b] = v[k];
  }
  if (READOUT) {
    const size_t N = a.g.N;
    if (a.snap) {
#pragma unroll
      for (int k = 0; k < 19; ++k) a.snap[(size_t)k * N + idx] = v[k];
    }
    if (a.mac) {
      a.mac[idx] = ux;
      a.mac[N + idx] = uy;
      a.mac[2 * N + idx] = uz;
      a.mac[3 * N + idx] = rho;
    }
  }
}

// Outlet cell (nx-1, y, z) of a row from the pre-wall post-collision values f of (nx-2, y, z) (equilibrium
// weights when that cell is solid) and its macroscopic values (lbmCpuReference.ts:354-371). Returns its rho.
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ float emit_outlet(const StepArgs& a, uint32_t oidx, uint8_t oflag, uint32_t y, uint32_t z,
                                             bool src_solid, const float* f, float rho, float ux, float uy, float uz) {
  const float rOut = rho * 0.85f + 0.15f;
  float fe[19];
  feq_bc(fe, rOut, ux, uy, uz);
  F19 go;
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    const float fp = src_solid ? w32(k) : f[k];
    go.v[k] = 0.75f * fp + 0.25f * fe[k];
  }
  emit_cell<T, FROM_E, READOUT>(a, oidx, a.g.nx - 1, y, z, oflag, true, go, rOut, ux, uy, uz);
  return rOut;
}

// Inlet cell (0, y, z) of a row from rho of (1, y, z) (lbmCpuReference.ts:328-352). Returns its rho.
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ float emit_inlet(const StepArgs& a, uint32_t iidx, uint8_t iflag, uint32_t y, uint32_t z,
                                            float rho) {
  const long long step = a.state->step;
  const double uin = inlet_u(a.ph, step);
  const float uxIn = (float)profile_u(a, uin, y);
  float uyIn = 0.0f;
  if (step < 120) uyIn = (float)(0.015 * uin * sin(step * 0.15 + ((double)y / a.g.ny) * M_PI * 2.0));
  const float rIn = 0.85f * rho + 0.15f;
  F19 gi;
  feq_bc(gi.v, rIn, uxIn, uyIn, 0.0f);
  emit_cell<T, FROM_E, READOUT>(a, iidx, 0, y, z, iflag, true, gi, rIn, uxIn, uyIn, 0.0f);
  return rIn;
}

// Everything after the collision of cell idx: its row's outlet cell (from the pre-wall values, for the
// thread of ix = nx-2), the wall rules, the stores, and the row's inlet cell (thread of ix = 1). f holds the
// post-collision populations (unused for solid cells). Returns the post-boundary rho of the cells finalized.
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ double finish_cell(const StepArgs& a, uint32_t idx, uint32_t x, uint32_t y, uint32_t z,
                                              uint8_t flag, bool surf, float* f, float rho, float ux, float uy,
                                              float uz) {
  const Geo& g = a.g;
  double rsum = 0.0;
  {
    {
      // Outlet cell of this row (lbmCpuReference.ts:354-371), from this cell's pre-wall values.
      if (x + 2 == g.nx) {
        const uint32_t oidx = idx + 1;
        const uint8_t oflag = a.flags[oidx];
        if (oflag != CELL_SOLID) {
          rsum += (double)emit_outlet<T, FROM_E, READOUT>(a, oidx, oflag, y, z, flag == CELL_SOLID, f, rho, ux, uy, uz);
        } else if (READOUT && a.mac) {
          a.mac[oidx] = 0.f; a.mac[g.N + oidx] = 0.f; a.mac[2 * (size_t)g.N + oidx] = 0.f; a.mac[3 * (size_t)g.N + oidx] = 1.f;
        }
      }
      if (flag != CELL_SOLID) {
        // Floor, ceiling and side walls (lbmCpuReference.ts:375-429), in the reference's order.
        if (surf && flag == CELL_FLUID) {
          bool walled = false;
          if (y == 0) {
            if (a.ph.noslip) { ux = 0.0f; uz = 0.0f; }
            uy = 0.0f;
            walled = true;
          }
          if (y + 1 == g.ny) { uy = 0.0f; walled = true; }
          if (a.ph.sides && (z == 0 || z + 1 == g.nz)) { uz = 0.0f; walled = true; }
          if (walled) feq_bc(f, rho, ux, uy, uz);
        }
        F19 fo;
#pragma unroll
        for (int k = 0; k < 19; ++k) fo.v[k] = f[k];
        emit_cell<T, FROM_E, READOUT>(a, idx, x, y, z, flag, surf, fo, rho, ux, uy, uz);
        rsum += (double)rho;
      } else if (READOUT && a.mac) {
        a.mac[idx] = 0.f; a.mac[g.N + idx] = 0.f; a.mac[2 * (size_t)g.N + idx] = 0.f; a.mac[3 * (size_t)g.N + idx] = 1.f;
      }
      // Inlet cell of this row (lbmCpuReference.ts:328-352).
      if (x == 1) {
        const uint32_t iidx = idx - 1;
        const uint8_t iflag = a.flags[iidx];
        if (iflag != CELL_SOLID) {
          rsum += (double)emit_inlet<T, FROM_E, READOUT>(a, iidx, iflag, y, z, rho);
        } else if (READOUT && a.mac) {
          a.mac[iidx] = 0.f; a.mac[g.N + iidx] = 0.f; a.mac[2 * (size_t)g.N + iidx] = 0.f; a.mac[3 * (size_t)g.N + iidx] = 1.f;
        }
            }
    }
  }
  return rsum;
}

// Full update of cell idx (plus the row's inlet or outlet cell for the threads of ix = 1 and ix = nx-2).
// Returns the post-boundary rho of the cells it finalized, for the mass-anchor sum.
template <typename T, bool FROM_E, bool READOUT, bool EAGER = false>
__device__ __forceinline__ double update_cell(const StepArgs& a, uint32_t idx) {
  const Geo& g = a.g;
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  if (!(x >= 1 && x + 1 < g.nx)) return 0.0;
  const uint8_t flag = a.flags[idx];
  const bool surf = (y == 0) | (y + 1 == g.ny) | (z == 0) | (z + 1 == g.nz);
  const double corr = a.state->corr;
  float f[19];
  float rho = 1.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
  if (flag != CELL_SOLID) {
    gather_cell<T, FROM_E, EAGER>(a, idx, x, y, z, flag, surf, corr, f);
    collide(f, rho, ux, uy, uz, a.ph);
  }
  return finish_cell<T, FROM_E, READOUT>(a, idx, x, y, z, flag, surf, f, rho, ux, uy, uz);
}

// Collision with the learned closure (lbmCpuReference.ts:273-304, with tau = tau0 + 3 nu_S g: the
// Smagorinsky eddy viscosity of the cell's non-equilibrium stress scaled by the network's factor g). f holds
// the pulled populations on entry and the post-collision ones on exit. The network runs on the matrix cores
// for the whole wavefront, so every lane calls this; a lane with live false evaluates it on zero stress and
// leaves f untouched, with rho = 1 and u = 0.
__device__ __forceinline__ void collide_mlp(float* f, float& rho, float& ux, float& uy, float& uz, const StepArgs& a,
                                            bool live) {
  rho = 1.0f;
  ux = uy = uz = 0.0f;
  float P[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (live) collide_moments(f, rho, ux, uy, uz, P);
  const float ir = __builtin_amdgcn_rcpf(rho);
  const float pxx = P[0] * ir, pyy = P[1] * ir, pzz = P[2] * ir, pxy = P[3] * ir, pxz = P[4] * ir, pyz = P[5] * ir;
  float xin[MLP_INP];
  const float pn = mlp_features(pxx, pyy, pzz, pxy, pxz, pyz, a.ph.nu, a.ph.tau0, a.ph.smag, xin);
  const float yv = mlp_eval(*a.mlp, xin);
  if (live) {
    const float tau = a.ph.tau0 + (smag_tau(a.ph.tau0, a.ph.smag, pn) - a.ph.tau0) * mlp_factor(yv);
    collide_relax(f, rho, ux, uy, uz, P, __builtin_amdgcn_rcpf(fmaxf(0.5015f, fminf(1.95f, tau))));
  }
}

// The same update with the learned closure: the relaxation time comes from the MLP, which runs on the
// matrix cores for the whole wavefront, so every lane of the wave must call this (inr = false for lanes
// without a cell).
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ double update_cell_mlp(const StepArgs& a, uint32_t idx, bool inr) {
  const Geo& g = a.g;
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  const bool valid = inr && x >= 1 && x + 1 < g.nx;
  const uint8_t flag = valid ? a.flags[idx] : CELL_SOLID;
  const bool surf = (y == 0) | (y + 1 == g.ny) | (z == 0) | (z + 1 == g.nz);
  const bool fluid = valid && flag != CELL_SOLID;
  float f[19];
  if (fluid) gather_cell<T, FROM_E>(a, idx, x, y, z, flag, surf, a.state->corr, f);
  float rho, ux, uy, uz;
  collide_mlp(f, rho, ux, uy, uz, a, fluid);
  if (!valid) return 0.0;
  return finish_cell<T, FROM_E, READOUT>(a, idx, x, y, z, flag, surf, f, rho, ux, uy, uz);
}

// ---- plain fluid cells on the y and z faces --------------------------------------------------------------
// A plain fluid cell has no Bouzidi link (no solid neighbour), so each incoming population comes either from
// its A-A slot or, when the pull source lies outside the domain, from the halo (the reference's clamped
// pull, lbmCpuReference.ts:190-192). Both addresses are computed and one of them selected before any load is
// issued, so the 19 loads of a cell are in flight together; gather_special's per-direction branches would
// instead wait for each load in turn.
template <typename T, bool FROM_E, int D>
__device__ __forceinline__ float ld_face(const StepArgs& a, size_t li, uint32_t x, uint32_t y, uint32_t z) {
  if (cy(D) == 0 && cz(D) == 0) return ld_in<T, FROM_E, D>(a, li);  // x sources stay inside (1 <= x <= nx-2)
  const Geo& g = a.g;
  const int sy = (int)y - cy(D), sz = (int)z - cz(D);
  const bool oy = cy(D) != 0 && (sy < 0 || sy >= (int)g.ny);
  const bool oz = cz(D) != 0 && (sz < 0 || sz >= (int)g.nz);
  const bool out = oy || oz;
  const uint32_t hx = (uint32_t)((int)x - cx(D)), hy = clampu(sy, g.ny), hz = clampu(sz, g.nz);
  uint32_t hi = 0;
  if (cy(D) != 0) {
    constexpr int FY = cy(D) > 0 ? FY0 : FY1;
    const uint32_t h = (FY == FY0 ? g.hoff[FY0] : g.hoff[FY1]) +
                       (uint32_t)face_slot(FY, D) * (FY == FY0 ? g.hsz[FY0] : g.hsz[FY1]) + hx + g.nx * hz;
    hi = oy ? h : hi;
  }
  if (cz(D) != 0) {
    constexpr int FZ = cz(D) > 0 ? FZ0 : FZ1;
    const uint32_t h = (FZ == FZ0 ? g.hoff[FZ0] : g.hoff[FZ1]) +
                       (uint32_t)face_slot(FZ, D) * (FZ == FZ0 ? g.hsz[FZ0] : g.hsz[FZ1]) + hx + g.nx * hy;
    hi = (!oy && oz) ? h : hi;
  }
  const typename T::elem* f = fptr<T>(a);
  const size_t slot = FROM_E ? li - loff_of<D>(g) + (size_t)opp(D) * g.nx : li + (size_t)D * g.nx;
  if constexpr (sizeof(typename T::elem) == sizeof(float)) {
    // same element type: select the address, one load
    const float* p = out ? a.halo_in + hi : reinterpret_cast<const float*>(f + slot);
    return *p;
  } else {
    // the halo keeps FP32 values beside the shifted halves of the buffer: each lane loads from the one that holds
    // its value (out is the same for a whole face row, so a wave mostly takes one side)
    if (out) return a.halo_in[hi];
    return T::ld(f + slot, D, a.ph);
  }
}

template <typename T, bool FROM_E>
__device__ __forceinline__ void gather_face(const StepArgs& a, size_t li, uint32_t x, uint32_t y, uint32_t z,
                                            double corr, float* f) {
#define AERO_GF(D) f[D] = ld_face<T, FROM_E, D>(a, li, x, y, z);
  AERO_GF(0) AERO_GF(1) AERO_GF(2) AERO_GF(3) AERO_GF(4) AERO_GF(5) AERO_GF(6) AERO_GF(7) AERO_GF(8) AERO_GF(9)
  AERO_GF(10) AERO_GF(11) AERO_GF(12) AERO_GF(13) AERO_GF(14) AERO_GF(15) AERO_GF(16) AERO_GF(17) AERO_GF(18)
#undef AERO_GF
  if (corr_fires(corr)) {
#pragma unroll
    for (int k = 0; k < 19; ++k) f[k] = apply_corr(f[k], corr);
  }
}

// Update of a plain fluid cell with 1 <= x <= nx-2 anywhere in the grid (in practice: on a y or z face).
template <typename T, bool FROM_E, bool READOUT>
__device__ __forceinline__ double update_fluid_cell(const StepArgs& a, uint32_t idx) {
  const Geo& g = a.g;
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  const bool surf = (y == 0) | (y + 1 == g.ny) | (z == 0) | (z + 1 == g.nz);
  float f[19];
  gather_face<T, FROM_E>(a, cell_li(g, x, rem), x, y, z, a.state->corr, f);
  float rho, ux, uy, uz;
  collide(f, rho, ux, uy, uz, a.ph);
  return finish_cell<T, FROM_E, READOUT>(a, idx, x, y, z, CELL_FLUID, surf, f, rho, ux, uy, uz);
}

// ---- split step: bulk groups and the special-cell list -------------------------------------------
// A group of V consecutive cells of a row (x0 a multiple of V) is a bulk group when its row is off the
// y and z faces (1 <= y <= ny-2, 1 <= z <= nz-2) and every cell of the group is a plain fluid cell; for
// V = 1 the cells x = 0 and x = nx-1 are never groups of their own, and x = 1 (x = nx-2) also needs a fluid
// inlet (outlet) cell next to it. Such cells have no wall rule, no halo read and no bounce-back, and each
// direction of the group moves with one V-wide load and one V-wide store. The group holding x = 1 also
// writes the row's inlet cell and the group holding x = nx-2 the outlet cell; for V >= 2 those groups also
// hold x = 0 or x = nx-1, whose loaded values are discarded and which are left out of the stores. Keeping
// the per-cell code out of the bulk kernel keeps its register count, and so its occupancy, at what the
// streaming needs; every other cell that the step must visit goes through update_cell in a second launch
// over a compacted list (cells within one step are independent under A-A streaming, so the two launches
// may split the grid in any way).
// The two halves of the test: the row and x position (no memory access), and the flags.
template <int V>
__device__ __forceinline__ bool bulk_row_ok(const Geo& g, uint32_t x0, uint32_t y, uint32_t z) {
  if (!(y >= 1 && y + 1 < g.ny && z >= 1 && z + 1 < g.nz)) return false;
  if (V == 1) return x0 != 0 && x0 + 1 != g.nx;
  return !(x0 == 0 && x0 + V == g.nx);  // one group holding both row ends
}
// Flags of the group (and for V = 1 of the inlet/outlet cell beside x = 1 / nx-2; nx >= 4, so at most one).
template <int V>
__device__ __forceinline__ uint32_t bulk_flag_word(const Geo& g, const uint8_t* __restrict__ flags, uint32_t idx0,
                                                   uint32_t x0) {
  static_assert(CELL_FLUID == 0, "bulk test compares whole flag words with zero");
  if (V == 1) {
    const uint32_t e = x0 == 1 ? idx0 - 1 : (x0 + 2 == g.nx ? idx0 + 1 : idx0);
    return (uint32_t)flags[idx0] | (uint32_t)flags[e];
  }
  if (V == 2) return *reinterpret_cast<const uint16_t*>(flags + idx0);
  return *reinterpret_cast<const uint32_t*>(flags + idx0);
}
template <int V>
__device__ __forceinline__ bool is_bulk_group(const Geo& g, const uint8_t* __restrict__ flags, uint32_t idx0,
                                              uint32_t x0, uint32_t y, uint32_t z) {
  return bulk_row_ok<V>(g, x0, y, z) && bulk_flag_word<V>(g, flags, idx0, x0) == 0;
}

// Address of the population slot of direction k of the cell at layout index li_blk + lane + d, as a
// block-uniform base (li_blk: the block's first cell) plus a small per-lane byte offset: every direction then
// shares one offset register and differs only in its scalar base.
template <typename E>
__device__ __forceinline__ E* slot_ptr(void* f, size_t li_blk, uint32_t nx, int k, int64_t d, uint32_t lane_bytes) {
  char* base = static_cast<char*>(f) + ((int64_t)li_blk + (int64_t)k * nx + d) * (int64_t)sizeof(E);
  return reinterpret_cast<E*>(base + lane_bytes);
}

// li_blk: layout index of the block's first cell; lane_li: layout offset of the group's first cell from it.
// Issues the 19 (V-wide) population loads of a group on a row off the y and z faces into v. On such a row
// every source slot is inside the allocation, whatever the flags.
template <typename T, bool FROM_E, int V>
__device__ __forceinline__ void bulk_load(const StepArgs& a, size_t li_blk, uint32_t lane_li, float (&v)[19][V]) {
  using E = typename T::elem;
  const Geo& g = a.g;
  const uint32_t nx = g.nx;
  const uint32_t lb = lane_li * (uint32_t)sizeof(E);
  // BUF: one buffer descriptor based lmax slots below the block's first one, so each load and store is a
  // uniform 32-bit byte offset per direction plus the group's 32-bit lane offset (4 + 19 or 38 scalar
  // registers rather than 19 or 38 64-bit bases); otherwise 64-bit addresses from 19 scalar bases per pass
  constexpr bool BUF = bulk_buffer_addressing((int)sizeof(E));
  const __amdgpu_buffer_rsrc_t rs = pop_rsrc(static_cast<char*>(a.f) + ((int64_t)li_blk - g.lmax) * (int64_t)sizeof(E));
  auto boff = [&](int slot, int64_t d) { return (int)(((int64_t)slot * nx + d + g.lmax) * (int64_t)sizeof(E)); };
#define AERO_LV(K)                                                                                        \
  if constexpr (BUF) {                                                                                    \
    if (FROM_E) T::template ldv_buf<V>(rs, (int)lb, boff(opp(K), -loff_of<K>(g)), K, a.ph, v[K]);         \
    else T::template ldv_buf<V>(rs, (int)lb, boff(K, 0), K, a.ph, v[K]);                                  \
  } else {                                                                                                \
    if (FROM_E) T::template ldv<V>(slot_ptr<E>(a.f, li_blk, nx, opp(K), -loff_of<K>(g), lb), K, a.ph, v[K]); \
    else T::template ldv<V>(slot_ptr<E>(a.f, li_blk, nx, K, 0, lb), K, a.ph, v[K]);                       \
  }
  AERO_LV(0) AERO_LV(1) AERO_LV(2) AERO_LV(3) AERO_LV(4) AERO_LV(5) AERO_LV(6) AERO_LV(7) AERO_LV(8)
  AERO_LV(9) AERO_LV(10) AERO_LV(11) AERO_LV(12) AERO_LV(13) AERO_LV(14) AERO_LV(15) AERO_LV(16)
  AERO_LV(17) AERO_LV(18)
#undef AERO_LV
}

// An empty asm consuming every loaded value: it keeps the loads above a following flag branch (the
// optimizer would otherwise sink them into the branch, behind the flag load's round trip).
template <int V>
__device__ __forceinline__ void bulk_pin(const float (&v)[19][V]) {
#pragma unroll
  for (int k = 0; k < 19; ++k)
#pragma unroll
    for (int c = 0; c < V; ++c) asm volatile("" ::"v"(v[k][c]));
}

// Stores V values of consecutive cells into a plain float array: one V-wide store, or for an edge group (V >= 2
// holding x = 0 or x = nx-1) the V-1 values other than element `skip`, whose cell its row's inlet or outlet rule
// writes.
// (Non-temporal stores here made a readout step 1.29 times a plain step, against 1.14-1.16 with these.)
template <int V>
__device__ __forceinline__ void readout_run(float* p, const float* vals, bool edge, int skip) {
  if (!edge) {
    F32::template stv<V>(p, vals, 0, Phys{});
  } else {
#pragma unroll
    for (int c = 0; c < V; ++c)
      if (c != skip) p[c] = vals[c];
  }
}

// Collides and stores a bulk group whose populations bulk_load brought into v; returns its density sum. corr
// is the mass anchor's factor, read by the caller before its loads (read after them, it would be a vector
// load that waits for every load before it).
template <typename T, bool FROM_E, int V, bool READOUT>
__device__ __forceinline__ double bulk_finish(const StepArgs& a, size_t li_blk, uint32_t lane_li, uint32_t idx0,
                                              uint32_t x0, uint32_t y, uint32_t z, double corr, float (&v)[19][V]) {
  using E = typename T::elem;
  constexpr bool BUF = bulk_buffer_addressing((int)sizeof(E));
  const Geo& g = a.g;
  const uint32_t nx = g.nx;
  const uint32_t lb = lane_li * (uint32_t)sizeof(E);
  // row ends inside the group: element EI is x = 1 (inlet duty), EO is x = nx-2 (outlet duty); for V >= 2
  // element 0 of the inlet group is x = 0 and element V-1 of the outlet group is x = nx-1 (not stored)
  constexpr int EI = V == 1 ? 0 : 1, EO = V == 1 ? 0 : V - 2;
  const bool in_grp = V == 1 ? x0 == 1 : x0 == 0;
  const bool out_grp = V == 1 ? x0 + 2 == g.nx : x0 + V == g.nx;
  const bool edge = V > 1 && (in_grp || out_grp);
  const __amdgpu_buffer_rsrc_t rs = pop_rsrc(static_cast<char*>(a.f) + ((int64_t)li_blk - g.lmax) * (int64_t)sizeof(E));
  auto boff = [&](int slot, int64_t d) { return (int)(((int64_t)slot * nx + d + g.lmax) * (int64_t)sizeof(E)); };
  if (corr_fires(corr)) {
#pragma unroll
    for (int k = 0; k < 19; ++k)
#pragma unroll
      for (int c = 0; c < V; ++c) v[k][c] = apply_corr(v[k][c], corr);
  }
  double rsum = 0.0;
  float rho_c[V], ux_c[V], uy_c[V], uz_c[V];
#pragma unroll
  for (int c = 0; c < V; ++c) {
    float fc[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) fc[k] = v[k][c];
    collide(fc, rho_c[c], ux_c[c], uy_c[c], uz_c[c], a.ph);
#pragma unroll
    for (int k = 0; k < 19; ++k) v[k][c] = fc[k];
    const bool skipped = V > 1 && ((c == 0 && in_grp) || (c == V - 1 && out_grp));
    if (!skipped) rsum += (double)rho_c[c];
  }
  // stores: one V-wide store per direction, or for an edge group the V-1 cells other than x = 0 / x = nx-1
  // (the branch is taken once for all 19 directions)
  // (the store offset of direction K is the load offset of opp(K): loff[K] = -loff[opp(K)])
  // (for BUF false, writing the store address of K as the load address of opp(K) lets the compiler keep 19
  // scalar bases instead of 38)
#define AERO_SP(K)                                                                                        \
  E* const p##K = FROM_E ? slot_ptr<E>(a.f, li_blk, nx, K, loff_of<K>(g), lb)                             \
                         : slot_ptr<E>(a.f, li_blk, nx, opp(K), 0, lb);
  AERO_SP(0) AERO_SP(1) AERO_SP(2) AERO_SP(3) AERO_SP(4) AERO_SP(5) AERO_SP(6) AERO_SP(7) AERO_SP(8)
  AERO_SP(9) AERO_SP(10) AERO_SP(11) AERO_SP(12) AERO_SP(13) AERO_SP(14) AERO_SP(15) AERO_SP(16)
  AERO_SP(17) AERO_SP(18)
#undef AERO_SP
  if (!edge) {
#define AERO_SV(K)                                                                                        \
    if constexpr (BUF) T::template stv_buf<V>(rs, (int)lb, FROM_E ? boff(K, loff_of<K>(g)) : boff(opp(K), 0), \
                                              v[K], FROM_E ? K : opp(K), a.ph);                           \
    else T::template stv<V>(p##K, v[K], FROM_E ? K : opp(K), a.ph);
    AERO_SV(0) AERO_SV(1) AERO_SV(2) AERO_SV(3) AERO_SV(4) AERO_SV(5) AERO_SV(6) AERO_SV(7) AERO_SV(8)
    AERO_SV(9) AERO_SV(10) AERO_SV(11) AERO_SV(12) AERO_SV(13) AERO_SV(14) AERO_SV(15) AERO_SV(16)
    AERO_SV(17) AERO_SV(18)
#undef AERO_SV
  } else {
    // the V-1 cells other than x = 0 / x = nx-1, one element each
    const int skip = in_grp ? 0 : V - 1;
#define AERO_SE(K)                                                                                        \
    _Pragma("unroll") for (int c = 0; c < V; ++c) if (c != skip) {                                        \
      if constexpr (BUF) T::template stv_buf<1>(rs, (int)(lb + c * (uint32_t)sizeof(E)),                   \
                                                FROM_E ? boff(K, loff_of<K>(g)) : boff(opp(K), 0), &v[K][c],  \
                                                FROM_E ? K : opp(K), a.ph);                               \
      else T::st(p##K + c, v[K][c], FROM_E ? K : opp(K), a.ph);                                           \
    }
    AERO_SE(0) AERO_SE(1) AERO_SE(2) AERO_SE(3) AERO_SE(4) AERO_SE(5) AERO_SE(6) AERO_SE(7) AERO_SE(8)
    AERO_SE(9) AERO_SE(10) AERO_SE(11) AERO_SE(12) AERO_SE(13) AERO_SE(14) AERO_SE(15) AERO_SE(16)
    AERO_SE(17) AERO_SE(18)
#undef AERO_SE
  }
  if constexpr (READOUT) {
    // macroscopic fields (and, when asked for, the snapshot) of the group's cells, in plain structure of arrays
    const size_t N = g.N;
    const int skip = in_grp ? 0 : V - 1;
    if (a.mac) {
      readout_run<V>(a.mac + idx0, ux_c, edge, skip);
      readout_run<V>(a.mac + N + idx0, uy_c, edge, skip);
      readout_run<V>(a.mac + 2 * N + idx0, uz_c, edge, skip);
      readout_run<V>(a.mac + 3 * N + idx0, rho_c, edge, skip);
    }
    if (a.snap) {
#pragma unroll
      for (int k = 0; k < 19; ++k) readout_run<V>(a.snap + (size_t)k * N + idx0, v[k], edge, skip);
    }
  }
  if (out_grp) {
    float fo[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) fo[k] = v[k][EO];
    rsum += (double)emit_outlet<T, FROM_E, READOUT>(a, idx0 + EO + 1, CELL_FLUID, y, z, false, fo, rho_c[EO],
                                                   ux_c[EO], uy_c[EO], uz_c[EO]);
  }
  if (in_grp) rsum += (double)emit_inlet<T, FROM_E, READOUT>(a, idx0 + EI - 1, CELL_FLUID, y, z, rho_c[EI]);
  return rsum;
}

// One group on a row off the y and z faces: its density sum if it is a bulk group, else 0 (nothing written).
// A wave whose groups all lie in one row of plain fluid cells needs no flag loads: every flag word there is 0, and
// the row's word (rowfluid) is one scalar load. With EARLY that word is tested after the population loads are
// issued, so the wave waits for no round trip of its own before them. Other waves load their flags: with EARLY
// after the population loads, otherwise before them. Against the per-lane flag test at 512x256x256 (alternating
// runs on one GPU): FP32 step 0.9897 against 0.9990 ms, FP16 step 0.5477 against 0.5593 ms.
template <typename T, bool FROM_E, int V, bool READOUT>
__device__ __forceinline__ double bulk_group(const StepArgs& a, size_t li_blk, uint32_t lane_li, uint32_t idx0,
                                             uint32_t x0, uint32_t y, uint32_t z) {
  constexpr bool EARLY = bulk_loads_before_flags((int)sizeof(typename T::elem));
  const double corr = a.state->corr;
  const uint32_t row = y + a.g.ny * z;
  const uint32_t row0 = __builtin_amdgcn_readfirstlane(row);
  const bool one_row = __all(row == row0);
  float v[19][V];
  if constexpr (EARLY) {
    bulk_load<T, FROM_E, V>(a, li_blk, lane_li, v);
    const uint32_t fw = (one_row && a.rowfluid[row0] != 0u) ? 0u : bulk_flag_word<V>(a.g, a.flags, idx0, x0);
    bulk_pin<V>(v);
    if (fw != 0) return 0.0;
  } else {
    const uint32_t fw = (one_row && a.rowfluid[row0] != 0u) ? 0u : bulk_flag_word<V>(a.g, a.flags, idx0, x0);
    if (fw != 0) return 0.0;
    bulk_load<T, FROM_E, V>(a, li_blk, lane_li, v);
  }
  return bulk_finish<T, FROM_E, V, READOUT>(a, li_blk, lane_li, idx0, x0, y, z, corr, v);
}

// Bulk groups only; every other group is left to k_step_special. One partial per block. READOUT: also write the
// macroscopic fields (and the snapshot when a.snap is set) of the cells updated.
template <typename T, bool FROM_E, int V, bool READOUT>
__global__ __launch_bounds__(bulk_block_threads(sizeof(typename T::elem))) void k_step_bulk(StepArgs a) {
  constexpr int BSZ = bulk_block_threads(sizeof(typename T::elem));
  const uint32_t lblk = xcd_chunk_block<(int)sizeof(typename T::elem)>(blockIdx.x, gridDim.x);
  const uint32_t blk0 = lblk * BSZ * V;
  const uint32_t idx0 = blk0 + threadIdx.x * V;
  double rsum = 0.0;
  if (idx0 < a.g.N) {
    const uint32_t rem = a.g.dnx.div(idx0);
    const uint32_t x0 = idx0 - rem * a.g.nx;
    const uint32_t z = a.g.dny.div(rem);
    const uint32_t y = rem - z * a.g.ny;
    if (bulk_row_ok<V>(a.g, x0, y, z)) {  // bulk_group tests the flags
      const uint32_t rb = a.g.dnx.div(blk0);
      const size_t li_blk = cell_li(a.g, blk0 - rb * a.g.nx, rb);
      const uint32_t lane_li = (uint32_t)(cell_li(a.g, x0, rem) - li_blk);
      rsum = bulk_group<T, FROM_E, V, READOUT>(a, li_blk, lane_li, idx0, x0, y, z);
    }
  }
  const double t = block_sum<BSZ>(rsum);
  if (threadIdx.x == 0) a.partials[lblk] = t;
}

// The plain fluid cells of the list (its first nface entries); partials go to part0 + block.
template <typename T, bool FROM_E, bool READOUT>
__global__ __launch_bounds__(BS) void k_step_face(StepArgs a, uint32_t part0) {
  const uint32_t t = blockIdx.x * BS + threadIdx.x;
  double rsum = 0.0;
  if (t < a.nface) rsum = update_fluid_cell<T, FROM_E, READOUT>(a, (uint32_t)a.spec[t]);
  const double s = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[part0 + blockIdx.x] = s;
}

// The other listed cells (boundary cells next to the body, solid cells with an inlet or outlet duty) through
// the full per-cell update; partials go to part0 + block.
template <typename T, bool FROM_E, bool READOUT>
__global__ __launch_bounds__(BS) void k_step_special(StepArgs a, uint32_t part0) {
  const uint32_t t = a.nface + blockIdx.x * BS + threadIdx.x;
  double rsum = 0.0;
  if (t < a.nspec) rsum = update_cell<T, FROM_E, READOUT, true>(a, (uint32_t)a.spec[t]);
  const double s = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[part0 + blockIdx.x] = s;
}

// Every cell through the per-cell update (readout steps: macroscopic fields and snapshot for all cells).
template <typename T, bool FROM_E, bool READOUT>
__global__ __launch_bounds__(BS) void k_step(StepArgs a) {
  const uint32_t idx = blockIdx.x * BS + threadIdx.x;
  double rsum = 0.0;
  if (idx < a.g.N) rsum = update_cell<T, FROM_E, READOUT>(a, idx);
  const double t = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = t;
}

// Marks the cells outside bulk groups of V that a split step visits: cells with 1 <= x <= nx-2 outside
// bulk groups, except solid cells away from x = 1 and x = nx-2 (a solid cell at x = 1 or nx-2 still writes
// its row's inlet or outlet cell; other solid cells have nothing to do). mark_fluid gets the plain fluid
// ones, mark_other the rest. count[0], count[1], count[2] receive the number of plain fluid cells marked,
// the number of cells the bulk kernel updates and the number of other cells marked.
__global__ void k_mark_special(const uint8_t* __restrict__ flags, Geo g, int V, uint8_t* __restrict__ mark_fluid,
                               uint8_t* __restrict__ mark_other, uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t mf = 0, mo = 0, nb = 0;
  if (i < g.N) {
    const uint32_t rem = g.dnx.div(i);
    const uint32_t x = i - rem * g.nx;
    const uint32_t z = g.dny.div(rem);
    const uint32_t y = rem - z * g.ny;
    const uint32_t x0 = x - x % (uint32_t)V, idx0 = i - x % (uint32_t)V;
    const bool bulk = V == 4 ? is_bulk_group<4>(g, flags, idx0, x0, y, z)
                             : (V == 2 ? is_bulk_group<2>(g, flags, idx0, x0, y, z) : is_bulk_group<1>(g, flags, idx0, x0, y, z));
    const uint8_t f = flags[i];
    const bool inner = x >= 1 && x + 2 <= g.nx;
    const bool m = !bulk && inner && (f != CELL_SOLID || x == 1 || x + 2 == g.nx);
    mf = (m && f == CELL_FLUID) ? 1u : 0u;
    mo = (m && f != CELL_FLUID) ? 1u : 0u;
    nb = (bulk && inner) ? 1u : 0u;
    mark_fluid[i] = (uint8_t)mf;
    mark_other[i] = (uint8_t)mo;
  }
  // per-wave counts through a ballot, per-block sums in LDS, one atomic per block and counter
  __shared__ uint32_t wsum[3][4];
  const uint64_t b0 = __ballot(mf), b1 = __ballot(nb), b2 = __ballot(mo);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wsum[0][w] = (uint32_t)__popcll(b0);
    wsum[1][w] = (uint32_t)__popcll(b1);
    wsum[2][w] = (uint32_t)__popcll(b2);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const uint32_t t = wsum[threadIdx.x][0] + wsum[threadIdx.x][1] + wsum[threadIdx.x][2] + wsum[threadIdx.x][3];
    if (t) atomicAdd(count + threadIdx.x, t);
  }
}

// Step with the learned closure (one cell per thread; every lane reaches the matrix-core evaluation).
template <typename T, bool FROM_E, bool READOUT>
__global__ __launch_bounds__(BS) void k_step_mlp(StepArgs a) {
  const uint32_t idx = blockIdx.x * BS + threadIdx.x;
  const bool inr = idx < a.g.N;
  const double rsum = update_cell_mlp<T, FROM_E, READOUT>(a, inr ? idx : 0u, inr);
  const double t = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = t;
}

// Plain steps with the learned closure, split like the Smagorinsky step: bulk cells (one per thread, the
// bulk test of k_step_bulk with V = 1) here, the listed cells in k_step_list_mlp. A lane outside a bulk cell
// still reaches the matrix-core evaluation. One partial per block.
#define AERO_LD1(K) f[K] = ld_in<T, FROM_E, K>(a, li);
#define AERO_ST1(K) st_out<T, FROM_E, K>(a, li, f[K]);
template <typename T, bool FROM_E>
__global__ __launch_bounds__(BS) void k_step_bulk_mlp(StepArgs a) {
  const Geo& g = a.g;
  const uint32_t idx0 = blockIdx.x * BS + threadIdx.x;
  bool live = false;
  uint32_t x0 = 0, y = 0, z = 0, rem = 0;
  if (idx0 < g.N) {
    rem = g.dnx.div(idx0);
    x0 = idx0 - rem * g.nx;
    z = g.dny.div(rem);
    y = rem - z * g.ny;
    live = is_bulk_group<1>(g, a.flags, idx0, x0, y, z);
  }
  const size_t li = cell_li(g, x0, rem);
  float f[19];
  if (live) {
    AERO_LD1(0) AERO_LD1(1) AERO_LD1(2) AERO_LD1(3) AERO_LD1(4) AERO_LD1(5) AERO_LD1(6) AERO_LD1(7) AERO_LD1(8)
    AERO_LD1(9) AERO_LD1(10) AERO_LD1(11) AERO_LD1(12) AERO_LD1(13) AERO_LD1(14) AERO_LD1(15) AERO_LD1(16)
    AERO_LD1(17) AERO_LD1(18)
    const double corr = a.state->corr;
    if (corr_fires(corr)) {
#pragma unroll
      for (int k = 0; k < 19; ++k) f[k] = apply_corr(f[k], corr);
    }
  }
  float rho, ux, uy, uz;
  collide_mlp(f, rho, ux, uy, uz, a, live);
  double rsum = 0.0;
  if (live) {
    // a bulk cell has 1 <= x <= nx-2 on a row off the y and z faces: every A-A destination is inside
    AERO_ST1(0) AERO_ST1(1) AERO_ST1(2) AERO_ST1(3) AERO_ST1(4) AERO_ST1(5) AERO_ST1(6) AERO_ST1(7) AERO_ST1(8)
    AERO_ST1(9) AERO_ST1(10) AERO_ST1(11) AERO_ST1(12) AERO_ST1(13) AERO_ST1(14) AERO_ST1(15) AERO_ST1(16)
    AERO_ST1(17) AERO_ST1(18)
    rsum = (double)rho;
    // x = 1 also writes the row's inlet cell, x = nx-2 its outlet cell (the order of bulk_group's sum)
    if (x0 + 2 == g.nx)
      rsum += (double)emit_outlet<T, FROM_E, false>(a, idx0 + 1, CELL_FLUID, y, z, false, f, rho, ux, uy, uz);
    if (x0 == 1) rsum += (double)emit_inlet<T, FROM_E, false>(a, idx0 - 1, CELL_FLUID, y, z, rho);
  }
  const double t = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = t;
}
#undef AERO_LD1
#undef AERO_ST1

// The listed cells (face rows, cells next to the body, solid cells with an inlet/outlet duty) with the learned
// closure; partials go to part0 + block.
template <typename T, bool FROM_E>
__global__ __launch_bounds__(BS) void k_step_list_mlp(StepArgs a, uint32_t part0) {
  const uint32_t t = blockIdx.x * BS + threadIdx.x;
  const bool inr = t < a.nspec;
  const double rsum = update_cell_mlp<T, FROM_E, false>(a, inr ? (uint32_t)a.spec[t] : 0u, inr);
  const double s = block_sum<BS>(rsum);
  if (threadIdx.x == 0) a.partials[part0 + blockIdx.x] = s;
}

// ---- many tunnels per launch ------------------------------------------------------------------------
// Two forms of a batched step over all tunnels.
// - Per-kind launches (shape grid 0): the bulk kernel over tiles of consecutive cells (one block per tile; a
//   thread's group takes the bulk path when it is a bulk group and is skipped otherwise), the list kernels over
//   tiles of each tunnel's listed cells (plain fluid cells through the branch-free face path, the others through
//   update_cell), then k_finalize_batch. Every tile writes one FP64 partial into its tunnel's contiguous range.
// - One persistent launch (shape grid > 0, k_step_batch_one): every kind of tile in one kernel, split by XCD
//   and finalized by the last block to finish.
// The per-tunnel StepArgs live in device memory and are uniform per tile.

// Density sum of this thread's group in a bulk tile (0 for a skipped group).
template <typename T, bool FROM_E, int V>
__device__ __forceinline__ double batch_bulk_rsum(const StepArgs& a, uint32_t start, uint32_t count) {
  const uint32_t off = threadIdx.x * V;
  if (off >= count) return 0.0;
  const uint32_t idx0 = start + off;
  const uint32_t rem = a.g.dnx.div(idx0);
  const uint32_t x0 = idx0 - rem * a.g.nx;
  const uint32_t z = a.g.dny.div(rem);
  const uint32_t y = rem - z * a.g.ny;
  if (!bulk_row_ok<V>(a.g, x0, y, z)) return 0.0;  // bulk_group tests the flags
  const uint32_t rb = a.g.dnx.div(start);
  const size_t li_blk = cell_li(a.g, start - rb * a.g.nx, rb);
  const uint32_t lane_li = (uint32_t)(cell_li(a.g, x0, rem) - li_blk);
  return bulk_group<T, FROM_E, V, false>(a, li_blk, lane_li, idx0, x0, y, z);
}
// Density sum of this thread's entry in a list tile (0 past the tile's count).
template <typename T, bool FROM_E, bool FLUID, bool EAGER = false>
__device__ __forceinline__ double batch_list_rsum(const StepArgs& a, uint32_t start, uint32_t count) {
  if (threadIdx.x >= count) return 0.0;
  const uint32_t idx = (uint32_t)a.spec[start + threadIdx.x];
  return FLUID ? update_fluid_cell<T, FROM_E, false>(a, idx) : update_cell<T, FROM_E, false, EAGER>(a, idx);
}

template <typename T, bool FROM_E, int BSZ, int V>
__global__ __launch_bounds__(BSZ) void k_step_batch_bulk(const StepArgs* __restrict__ targs,
                                                         const BatchTile* __restrict__ tiles, uint32_t ntiles,
                                                         double* __restrict__ part) {
  if (blockIdx.x >= ntiles) return;
  const BatchTile tl = tiles[blockIdx.x];
  const double s = block_sum<BSZ>(batch_bulk_rsum<T, FROM_E, V>(targs[tl.tunnel], tl.start, tl.count));
  if (threadIdx.x == 0) part[tl.part] = s;
}

// FLUID tiles hold plain fluid entries (branch-free face path), the others the remaining entries (full per-cell
// update), which the batch runs on a second stream beside the bulk kernel.
template <typename T, bool FROM_E, bool FLUID>
__global__ __launch_bounds__(BS) void k_step_batch_list(const StepArgs* __restrict__ targs,
                                                        const BatchTile* __restrict__ tiles, uint32_t ntiles,
                                                        double* __restrict__ part) {
  if (blockIdx.x >= ntiles) return;
  const BatchTile tl = tiles[blockIdx.x];
  const double s = block_sum<BS>(batch_list_rsum<T, FROM_E, FLUID, !FLUID>(targs[tl.tunnel], tl.start, tl.count));
  if (threadIdx.x == 0) part[tl.part] = s;
}

// One batched step in one launch. Dispatched blocks go round-robin to the 8 XCDs (block b to XCD b mod 8), and the
// blocks of XCD x take the tiles of the contiguous range [xcd_begin[x], xcd_begin[x + 1]) from that XCD's work
// counter, one tile at a time, the next index fetched while the current tile runs: the blocks of one XCD sweep one
// region of one or two tunnels together and share that XCD's L2 for the neighbour rows and planes, and a block that
// drew cheap tiles takes more of them.
// Which block takes which tile depends on timing, so the densities are summed in integers: each thread adds its
// cells' density sum per tile (an exact FP64 sum of a few floats) as a 64-bit fixed-point value with 32 fraction
// bits, and the block adds the values of a run of tiles of one tunnel into that tunnel's total with one atomic.
// Integer addition is associative, so the totals are the same in every run. They are also exact, since every
// density is a float in [0.5, 1.8], a multiple of 2^-24, and so are the FP64 sums of the other step forms: every
// form finds the same total and the same anchor factor. The block that finishes last (one ticket per block)
// applies the mass-anchor rule of lbmCpuReference.ts:442-455 to every tunnel and re-arms the counters, so the step
// needs no second launch and no wait between blocks.
// sync: uint32 [0] the ticket, [16 (1 + x)] the work counter of XCD x (a line each); from byte 1024, one 64-bit
// fixed-point density total per tunnel. All zero between launches.
// The tables come as separate restrict-qualified arguments: read-only and unaliased, so the tile records and each
// tunnel's StepArgs are scalar loads and the population addresses stay uniform (through a struct of plain pointers
// they became per-lane values, every buffer access a loop over lanes: five times slower).
// Measured on the 64-tunnel sweep (133 M cells), its best shape (512 threads, 1 cell per thread, 1024 blocks) takes
// 4.77-4.81 ms per FP32 step against 4.58-4.63 ms for the per-kind launches; in FP16 storage 3.65 against 2.96-3.01
// ms, where this kernel's 121-125 registers halve the bulk tiles' occupancy. The autotuner measures both forms.
constexpr double FIX_SCALE = 4294967296.0;  // 2^32
template <typename T, bool FROM_E, int BSZ, int V>
__global__ __launch_bounds__(BSZ) void k_step_batch_one(const StepArgs* __restrict__ targs,
                                                        const BatchTileP* __restrict__ tiles,
                                                        const uint32_t* __restrict__ xcd_begin,
                                                        StepState* const* __restrict__ states, uint32_t ntun,
                                                        uint32_t* __restrict__ sync) {
  const uint32_t xcd = blockIdx.x & 7u;
  const uint32_t t0 = xcd_begin[xcd], t1 = xcd_begin[xcd + 1];
  uint32_t* const ctr = sync + 16u * (1u + xcd);
  unsigned long long* const total = reinterpret_cast<unsigned long long*>(sync + 256);
  __shared__ uint32_t s_next[2];
  __shared__ long long s_red[BSZ / 64];
  // adds the block's accumulators to tunnel tn's total
  auto flush = [&](long long v, uint32_t tn) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long sum = 0;
#pragma unroll
      for (int i = 0; i < BSZ / 64; ++i) sum += s_red[i];
      __hip_atomic_fetch_add(total + tn, (unsigned long long)sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  };
  if (threadIdx.x == 0) s_next[0] = t0 + __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  uint32_t it = 0;
  uint32_t t = __builtin_amdgcn_readfirstlane(s_next[0]);
  long long acc = 0;  // this thread's fixed-point density sum over the current run of one tunnel
  uint32_t cur = 0xffffffffu;
  while (t < t1) {
    if (threadIdx.x == 0)  // the next tile index, into the other slot while this tile runs
      s_next[(it + 1) & 1] = t0 + __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const BatchTileP tl = tiles[t];
    const uint32_t kind = tl.tk >> 30, tun = tl.tk & 0x3fffffffu;
    if (tun != cur) {
      if (cur != 0xffffffffu) flush(acc, cur);
      acc = 0;
      cur = tun;
    }
    const StepArgs& a = targs[tun];
    double r;
    if (kind == TILE_BULK) r = batch_bulk_rsum<T, FROM_E, V>(a, tl.start, tl.count);
    else if (kind == TILE_FLUID) r = batch_list_rsum<T, FROM_E, true>(a, tl.start, tl.count);
    else r = batch_list_rsum<T, FROM_E, false>(a, tl.start, tl.count);
    acc += __double2ll_rn(r * FIX_SCALE);
    __syncthreads();  // the prefetched index is in place, and every thread has read the current one
    ++it;
    t = __builtin_amdgcn_readfirstlane(s_next[it & 1]);
  }
  if (cur != 0xffffffffu) flush(acc, cur);
  __shared__ uint32_t last;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this block's totals before its ticket
    last = __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == gridDim.x;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every block's totals
  for (uint32_t tn = threadIdx.x; tn < ntun; tn += BSZ) {
    const long long q = (long long)__hip_atomic_load(total + tn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double v = (double)q * (1.0 / FIX_SCALE);
    StepState* st = states[tn];
    st->sum = v;
    double corr = 1.0;
    if (st->nfluid > 0) {
      const double mean = v / (double)st->nfluid;
      st->mean = mean;
      if (fabs(mean - 1.0) > 1e-4) {
        corr = 1.0 / mean;
        st->anchors += 1;
      }
    }
    st->corr = corr;
    st->step += 1;
    __hip_atomic_store(total + tn, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x < 8) __hip_atomic_store(sync + 16u * (1u + threadIdx.x), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (threadIdx.x == 0) __hip_atomic_store(sync, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One block per tunnel: the sum of its tiles' partials (exact, like every density sum here), then the mass-anchor
// rule.
__global__ __launch_bounds__(256) void k_finalize_batch(const double* __restrict__ tile_partials,
                                                        const uint32_t* __restrict__ tile_begin,
                                                        StepState* const* __restrict__ states) {
  __shared__ double sh[4];
  const uint32_t t = blockIdx.x;
  const uint32_t b0 = tile_begin[t], b1 = tile_begin[t + 1];
  double v = 0.0;
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += 256) v += tile_partials[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s = sh[0] + sh[1] + sh[2] + sh[3];
    StepState* st = states[t];
    st->sum = s;
    double corr = 1.0;
    if (st->nfluid > 0) {
      const double mean = s / (double)st->nfluid;
      st->mean = mean;
      if (fabs(mean - 1.0) > 1e-4) {
        corr = 1.0 / mean;
        st->anchors += 1;
      }
    }
    st->corr = corr;
    st->step += 1;
  }
}

// Sums chunks of 1024 consecutive partials (4 per thread, then a block tree): the first stage of the
// reduction when there are many blocks.
__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ in, uint32_t n, double* __restrict__ out) {
  const uint32_t base = blockIdx.x * 1024;
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = base + k * 256 + threadIdx.x;
    if (i < n) v += in[i];
  }
  const double t = block_sum<256>(v);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Sums the block partials (exactly) and applies the mass-anchor rule of
// lbmCpuReference.ts:442-455; advances the step counter.
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ partials, uint32_t n,
                                                    StepState* __restrict__ st) {
  __shared__ double sh[16];
  double v = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 1024) v += partials[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < 16; ++i) s += sh[i];
    st->sum = s;
    double corr = 1.0;
    if (st->nfluid > 0) {
      const double mean = s / (double)st->nfluid;
      st->mean = mean;
      const double c = 1.0 / mean;
      if (fabs(mean - 1.0) > 1e-4) {
        corr = c;
        st->anchors += 1;
      }
    }
    st->corr = corr;
    st->step += 1;
  }
}

// Equilibrium initialisation into A-A state E (lbmCpuReference.ts:114-143 plus seeded noise).
template <typename T>
__global__ __launch_bounds__(BS) void k_init(StepArgs a, double u0, uint64_t seed, double amp) {
  const uint32_t idx = blockIdx.x * BS + threadIdx.x;
  const Geo& g = a.g;
  if (idx >= g.N) return;
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  const uint8_t flag = a.flags[idx];
  double ux0 = u0, uy0 = 0.0, uz0 = 0.0;
  if (a.ph.profile != 0) {
    ux0 = profile_u(a, u0, y);
    if (y == 0) ux0 = 0.0;
  }
  if (flag == CELL_SOLID) {
    ux0 = 0.0;
  } else if (amp > 0.0) {
    ux0 = ux0 + amp * noise_unit(seed, idx, 0);
    uy0 = amp * noise_unit(seed, idx, 1);
    uz0 = amp * noise_unit(seed, idx, 2);
  }
  float v[19];
  feq_init(v, 1.0, ux0, uy0, uz0);
  typename T::elem* f = fptr<T>(a);
  const size_t li = cell_li(g, x, rem);
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    // state E: f*_k(x) at (x, opp(k))
    T::st(f + (li + (size_t)opp(k) * g.nx), v[k], k, a.ph);
  }
  if (flag == CELL_SOLID) return;
  const bool surf = x == 0 || x + 1 == g.nx || y == 0 || y + 1 == g.ny || z == 0 || z + 1 == g.nz;
  if (surf) halo_put(g, a.halo_out, v, x, y, z);
  if (flag == CELL_BOUNDARY) {
    const int b = a.bmap[idx];
    for (int k = 0; k < 19; ++k) a.bown[(size_t)k * a.B + b] = v[k];
  }
}

// Loads post-step populations p[19*N] (the reference's fSrc) into A-A state E.
template <typename T>
__global__ __launch_bounds__(BS) void k_load(StepArgs a, const float* __restrict__ p) {
  const uint32_t idx = blockIdx.x * BS + threadIdx.x;
  const Geo& g = a.g;
  if (idx >= g.N) return;
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  const uint8_t flag = a.flags[idx];
  float v[19];
  for (int k = 0; k < 19; ++k) v[k] = p[(size_t)k * g.N + idx];
  typename T::elem* f = fptr<T>(a);
  const size_t li = cell_li(g, x, rem);
#pragma unroll
  for (int k = 0; k < 19; ++k) {
    T::st(f + (li + (size_t)opp(k) * g.nx), v[k], k, a.ph);
  }
  if (flag == CELL_SOLID) return;
  const bool surf = x == 0 || x + 1 == g.nx || y == 0 || y + 1 == g.ny || z == 0 || z + 1 == g.nz;
  if (surf) halo_put(g, a.halo_out, v, x, y, z);
  if (flag == CELL_BOUNDARY) {
    const int b = a.bmap[idx];
    for (int k = 0; k < 19; ++k) a.bown[(size_t)k * a.B + b] = v[k];
  }
}

// Per boundary cell: bit k set when the clamped cell x - c_k is solid (the pull-source test of
// lbmCpuReference.ts:190-196; bit opp(k) is then the ff test of :205-210 and forceIntegrator.ts:179-185).
__global__ void k_links(const uint8_t* __restrict__ flags, const int32_t* __restrict__ blist, uint32_t B, Geo g,
                        uint32_t* __restrict__ bmask) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const uint32_t idx = (uint32_t)blist[b];
  const uint32_t rem = g.dnx.div(idx);
  const uint32_t x = idx - rem * g.nx;
  const uint32_t z = g.dny.div(rem);
  const uint32_t y = rem - z * g.ny;
  uint32_t m = 0;
  for (int k = 1; k < 19; ++k) {
    const uint32_t sx = clampu((int)x - cx(k), g.nx), sy = clampu((int)y - cy(k), g.ny), sz = clampu((int)z - cz(k), g.nz);
    if (flags[sx + g.nx * (sy + g.ny * sz)] == CELL_SOLID) m |= 1u << k;
  }
  bmask[b] = m;
}

// Momentum-exchange force and torque per boundary cell (forceIntegrator.ts:159-217 with the torque
// vector of forces.wgsl:95-102), block-reduced in FP64. Reads the post-step state: own values from bown,
// f*_d(x - c_d) from the in-place buffer (state given by aa_state) or the halo when x - c_d is clamped.
template <typename T>
__global__ __launch_bounds__(BS) void k_forces(StepArgs a, const int32_t* __restrict__ blist, int aa_state,
                                               const float* __restrict__ halo_cur, double* __restrict__ part) {
  const uint32_t b = blockIdx.x * BS + threadIdx.x;
  const Geo& g = a.g;
  double fx = 0, fy = 0, fz = 0, mx = 0, my = 0, mz = 0, cut = 0;
  if (b < a.B) {
    const double corr = a.state->corr;
    const uint32_t idx = (uint32_t)blist[b];
    const uint32_t rem = g.dnx.div(idx);
    const uint32_t x = idx - rem * g.nx;
    const uint32_t z = g.dny.div(rem);
    const uint32_t y = rem - z * g.ny;
    const uint32_t m = a.bmask[b];
    const typename T::elem* f = fptr<T>(a);
    const size_t li = cell_li(g, x, rem);
    double cfx = 0, cfy = 0, cfz = 0;
#pragma unroll
    for (int d = 1; d <= 18; ++d) {
      const float qf = a.bq[(size_t)b * 18 + (d - 1)];
      if (qf > 0.0f) {
        const double q = qf;
        const double fOut = (double)apply_corr(a.bown[(size_t)d * a.B + b], corr);
        double fIn = fOut;
        if (!((m >> d) & 1u)) {
          if (q < 0.5) {
            const int sx = (int)x - cx(d), sy = (int)y - cy(d), sz = (int)z - cz(d);
            float v;
            if (sx < 0 || sx >= (int)g.nx || sy < 0 || sy >= (int)g.ny || sz < 0 || sz >= (int)g.nz) {
              v = halo_get(g, halo_cur, out_face(g, sx, sy, sz), d, clampu(sx, g.nx), clampu(sy, g.ny),
                           clampu(sz, g.nz));
            } else if (aa_state == 0) {
              v = T::ld(f + (li - g.loff[d] + (size_t)opp(d) * g.nx), opp(d), a.ph);
            } else {
              v = T::ld(f + (li + (size_t)d * g.nx), d, a.ph);
            }
            const double fN = (double)apply_corr(v, corr);
            fIn = 2.0 * q * fOut + (1.0 - 2.0 * q) * fN;
          } else {
            const double fI = (double)apply_corr(a.bown[(size_t)opp(d) * a.B + b], corr);
            const double inv2q = 1.0 / (2.0 * q);
            fIn = inv2q * fOut + (1.0 - inv2q) * fI;
          }
        }
        const double mom = fOut + fIn - 2.0 * (double)w32(d);
        cfx += cx(d) * mom;
        cfy += cy(d) * mom;
        cfz += cz(d) * mom;
        if (d <= 6) cut += 1.0;
      }
    }
    fx = cfx; fy = cfy; fz = cfz;
    const double rx = (x + 0.5) - 0.5 * g.nx, ry = (y + 0.5) - 0.5 * g.ny, rz = (z + 0.5) - 0.5 * g.nz;
    mz = rx * cfy - ry * cfx;
    mx = ry * cfz - rz * cfy;
    my = rz * cfx - rx * cfz;
  }
  const double vals[7] = {fx, fy, fz, mx, my, mz, cut};
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const double t = block_sum<BS>(vals[j]);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8 + j] = t;
    __syncthreads();
  }
}

__global__ void k_forces_final(const double* __restrict__ part, uint32_t nblk, double* __restrict__ out) {
  const int j = threadIdx.x;
  if (j < 7) {
    double s = 0.0;
    for (uint32_t i = 0; i < nblk; ++i) s += part[(size_t)i * 8 + j];
    out[j] = s;
  }
}

// ---- host launchers ------------------------------------------------------------------------------
template <typename T, bool FROM_E, bool RO>
static hipError_t launch_split_e(const StepArgs& a, int V, uint32_t nb, uint32_t nf, uint32_t no, hipStream_t s,
                                 hipEvent_t mid, const SideLaunch* side) {
  hipError_t e;
  const bool fork = side && no;
  if (fork) {
    // the listed cells run beside the bulk kernel on the side stream: first the face rows (on a third stream of
    // their own they take more bandwidth from the bulk kernel: FP16 steps measured 4% slower), then the few
    // cells of the other list with their long per-cell chains (in the opposite order the FP32 step measured
    // 0.9891 against 0.9849 ms; after the bulk kernel on the main stream, 1.016 ms)
    if ((e = hipEventRecord(side->fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(side->stream, side->fork, 0)) != hipSuccess) return e;
    if (nf) hipLaunchKernelGGL((k_step_face<T, FROM_E, RO>), dim3(nf), dim3(BS), 0, side->stream, a, nb);
    hipLaunchKernelGGL((k_step_special<T, FROM_E, RO>), dim3(no), dim3(BS), 0, side->stream, a, nb + nf);
    if ((e = hipEventRecord(side->join, side->stream)) != hipSuccess) return e;
  }
  constexpr int esz = (int)sizeof(typename T::elem);
  const dim3 bb(bulk_block_threads(esz));
  if (V == 4) hipLaunchKernelGGL((k_step_bulk<T, FROM_E, 4, RO>), dim3(nb), bb, bulk_block_lds(esz), s, a);
  else if (V == 2) hipLaunchKernelGGL((k_step_bulk<T, FROM_E, 2, RO>), dim3(nb), bb, bulk_block_lds(esz), s, a);
  else hipLaunchKernelGGL((k_step_bulk<T, FROM_E, 1, RO>), dim3(nb), bb, bulk_block_lds(esz), s, a);
  if (mid && (e = hipEventRecord(mid, s)) != hipSuccess) return e;
  if (nf && !fork) hipLaunchKernelGGL((k_step_face<T, FROM_E, RO>), dim3(nf), dim3(BS), 0, s, a, nb);
  if (fork) {
    if ((e = hipStreamWaitEvent(s, side->join, 0)) != hipSuccess) return e;
  } else if (no) {
    hipLaunchKernelGGL((k_step_special<T, FROM_E, RO>), dim3(no), dim3(BS), 0, s, a, nb + nf);
  }
  return hipSuccess;
}
template <typename T>
static hipError_t launch_step_t(const StepArgs& a, int V, bool from_e, bool readout, hipStream_t s, uint32_t* nblk,
                                hipEvent_t mid, const SideLaunch* side) {
  if (V == 0 || !a.g.bufok || (readout && !a.spec)) {
    // the per-cell kernel: when asked for (V = 0), cross-sections beyond the bulk kernels' 32-bit offsets, and
    // readout steps before a cell list exists
    const uint32_t nb = (a.g.N + BS - 1) / BS;
    if (readout) {
      if (from_e) hipLaunchKernelGGL((k_step<T, true, true>), dim3(nb), dim3(BS), 0, s, a);
      else hipLaunchKernelGGL((k_step<T, false, true>), dim3(nb), dim3(BS), 0, s, a);
    } else {
      if (from_e) hipLaunchKernelGGL((k_step<T, true, false>), dim3(nb), dim3(BS), 0, s, a);
      else hipLaunchKernelGGL((k_step<T, false, false>), dim3(nb), dim3(BS), 0, s, a);
    }
    if (mid) {
      const hipError_t e = hipEventRecord(mid, s);
      if (e != hipSuccess) return e;
    }
    *nblk = nb;
    return hipGetLastError();
  }
  const uint32_t bsz = bulk_block_threads((int)sizeof(typename T::elem));
  const uint32_t nb = (a.g.N / V + bsz - 1) / bsz;
  const uint32_t nf = (a.nface + BS - 1) / BS, no = (a.nspec - a.nface + BS - 1) / BS;
  hipError_t e;
  if (readout) e = from_e ? launch_split_e<T, true, true>(a, V, nb, nf, no, s, mid, side)
                          : launch_split_e<T, false, true>(a, V, nb, nf, no, s, mid, side);
  else e = from_e ? launch_split_e<T, true, false>(a, V, nb, nf, no, s, mid, side)
                  : launch_split_e<T, false, false>(a, V, nb, nf, no, s, mid, side);
  if (e != hipSuccess) return e;
  *nblk = nb + nf + no;
  return hipGetLastError();
}
template <typename T>
static hipError_t launch_step_mlp_t(const StepArgs& a, int V, bool from_e, bool readout, hipStream_t s,
                                    uint32_t* nblk, hipEvent_t mid, const SideLaunch* side) {
  const uint32_t nb = (a.g.N + BS - 1) / BS;
  if (readout || !a.spec || V == 0) {
    // readout steps, steps before a cell list exists and steps asked to (V = 0) take every cell through the
    // per-cell kernel
    if (from_e) {
      if (readout) hipLaunchKernelGGL((k_step_mlp<T, true, true>), dim3(nb), dim3(BS), 0, s, a);
      else hipLaunchKernelGGL((k_step_mlp<T, true, false>), dim3(nb), dim3(BS), 0, s, a);
    } else {
      if (readout) hipLaunchKernelGGL((k_step_mlp<T, false, true>), dim3(nb), dim3(BS), 0, s, a);
      else hipLaunchKernelGGL((k_step_mlp<T, false, false>), dim3(nb), dim3(BS), 0, s, a);
    }
    if (mid) {
      const hipError_t e = hipEventRecord(mid, s);
      if (e != hipSuccess) return e;
    }
    *nblk = nb;
    return hipGetLastError();
  }
  const uint32_t nl = (a.nspec + BS - 1) / BS;
  hipError_t e;
  // the listed cells run beside the bulk kernel on the side stream when there is one (distinct cells touch
  // distinct A-A slots), else after it
  const bool fork = side && nl;
  const hipStream_t ls = fork ? side->stream : s;
  if (fork) {
    if ((e = hipEventRecord(side->fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(side->stream, side->fork, 0)) != hipSuccess) return e;
  }
  auto list = [&] {
    if (from_e) hipLaunchKernelGGL((k_step_list_mlp<T, true>), dim3(nl), dim3(BS), 0, ls, a, nb);
    else hipLaunchKernelGGL((k_step_list_mlp<T, false>), dim3(nl), dim3(BS), 0, ls, a, nb);
  };
  if (fork) {
    list();
    if ((e = hipEventRecord(side->join, side->stream)) != hipSuccess) return e;
  }
  if (from_e) hipLaunchKernelGGL((k_step_bulk_mlp<T, true>), dim3(nb), dim3(BS), 0, s, a);
  else hipLaunchKernelGGL((k_step_bulk_mlp<T, false>), dim3(nb), dim3(BS), 0, s, a);
  if (mid && (e = hipEventRecord(mid, s)) != hipSuccess) return e;
  if (fork) {
    if ((e = hipStreamWaitEvent(s, side->join, 0)) != hipSuccess) return e;
  } else if (nl) {
    list();
  }
  *nblk = nb + nl;
  return hipGetLastError();
}
hipError_t launch_step(const StepArgs& a, bool fp16, int V, bool from_e, bool readout, hipStream_t s,
                       uint32_t* nblk, hipEvent_t mid, const SideLaunch* side) {
  if (a.mlp) {
    // the learned closure's split step uses one-cell groups (the caller's list is built for V = 1)
    return fp16 ? launch_step_mlp_t<F16>(a, V, from_e, readout, s, nblk, mid, side)
                : launch_step_mlp_t<F32>(a, V, from_e, readout, s, nblk, mid, side);
  }
  return fp16 ? launch_step_t<F16>(a, V, from_e, readout, s, nblk, mid, side)
              : launch_step_t<F32>(a, V, from_e, readout, s, nblk, mid, side);
}
uint32_t split_blocks(uint32_t N, int V, uint32_t nfluid, uint32_t nother) {
  return (N / V + BS - 1) / BS + (nfluid + BS - 1) / BS + (nother + BS - 1) / BS;
}
hipError_t build_special_list(const uint8_t* flags, const Geo& g, int V, int32_t** list, uint32_t* nfluid,
                              uint32_t* nother, uint32_t* nbulk, hipStream_t s) {
  *list = nullptr;
  *nfluid = *nother = *nbulk = 0;
  uint8_t* mark = nullptr;
  uint32_t* dcnt = nullptr;
  int* dsel = nullptr;
  void* tmp = nullptr;
  int32_t* out = nullptr;
  hipError_t e = hipSuccess;
  auto done = [&](hipError_t r) {
    if (mark) (void)hipFree(mark);
    if (dcnt) (void)hipFree(dcnt);
    if (dsel) (void)hipFree(dsel);
    if (tmp) (void)hipFree(tmp);
    if (r != hipSuccess && out) (void)hipFree(out);
    return r;
  };
  if ((e = hipMalloc(&mark, (size_t)2 * g.N)) != hipSuccess) return done(e);
  if ((e = hipMalloc(&dcnt, 16)) != hipSuccess) return done(e);
  if ((e = hipMalloc(&dsel, 16)) != hipSuccess) return done(e);
  if ((e = hipMemsetAsync(dcnt, 0, 16, s)) != hipSuccess) return done(e);
  hipLaunchKernelGGL(k_mark_special, dim3((g.N + 255) / 256), dim3(256), 0, s, flags, g, V, mark, mark + g.N, dcnt);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  uint32_t hc[3] = {0, 0, 0};
  if ((e = hipMemcpyAsync(hc, dcnt, 12, hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(e);
  const uint32_t nf = hc[0], no = hc[2];
  *nbulk = hc[1];
  if (nf + no == 0) return done(hipSuccess);
  if ((e = hipMalloc(&out, (size_t)(nf + no) * 4)) != hipSuccess) return done(e);
  // order-preserving compactions: plain fluid cells first, then the others, each in memory order
  hipcub::CountingInputIterator<int32_t> cit(0);
  size_t tb = 0;
  if ((e = hipcub::DeviceSelect::Flagged(nullptr, tb, cit, mark, out, dsel, (int)g.N, s)) != hipSuccess) return done(e);
  if ((e = hipMalloc(&tmp, tb > 16 ? tb : 16)) != hipSuccess) return done(e);
  if ((e = hipcub::DeviceSelect::Flagged(tmp, tb, cit, mark, out, dsel, (int)g.N, s)) != hipSuccess) return done(e);
  if ((e = hipcub::DeviceSelect::Flagged(tmp, tb, cit, mark + g.N, out + nf, dsel + 1, (int)g.N, s)) != hipSuccess)
    return done(e);
  int sel[2] = {0, 0};
  if ((e = hipMemcpyAsync(sel, dsel, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(e);
  if ((uint32_t)sel[0] != nf || (uint32_t)sel[1] != no) return done(hipErrorUnknown);
  *list = out;
  *nfluid = nf;
  *nother = no;
  return done(hipSuccess);
}
__global__ __launch_bounds__(256) void k_solid_macros(const uint8_t* __restrict__ flags, uint32_t N,
                                                      float* __restrict__ mac) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N || flags[i] != CELL_SOLID) return;
  mac[i] = 0.0f;
  mac[N + i] = 0.0f;
  mac[2 * (size_t)N + i] = 0.0f;
  mac[3 * (size_t)N + i] = 1.0f;
}
hipError_t launch_solid_macros(const uint8_t* flags, uint32_t N, float* mac, hipStream_t s) {
  hipLaunchKernelGGL(k_solid_macros, dim3((N + 255) / 256), dim3(256), 0, s, flags, N, mac);
  return hipGetLastError();
}
hipError_t launch_finalize(double* partials, uint32_t n, StepState* st, hipStream_t s) {
  if (n > 4096) {
    const uint32_t m = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_reduce_chunks, dim3(m), dim3(256), 0, s, partials, n, partials + n);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, partials + n, m, st);
  } else {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, partials, n, st);
  }
  return hipGetLastError();
}
uint32_t partials_len(uint32_t nblk) { return nblk + (nblk + 1023) / 1024 + 1; }
hipError_t launch_init(const StepArgs& a, bool fp16, double u0, uint64_t seed, double amp, hipStream_t s) {
  const uint32_t nblk = (a.g.N + BS - 1) / BS;
  if (fp16) hipLaunchKernelGGL((k_init<F16>), dim3(nblk), dim3(BS), 0, s, a, u0, seed, amp);
  else hipLaunchKernelGGL((k_init<F32>), dim3(nblk), dim3(BS), 0, s, a, u0, seed, amp);
  return hipGetLastError();
}
hipError_t launch_load(const StepArgs& a, bool fp16, const float* p, hipStream_t s) {
  const uint32_t nblk = (a.g.N + BS - 1) / BS;
  if (fp16) hipLaunchKernelGGL((k_load<F16>), dim3(nblk), dim3(BS), 0, s, a, p);
  else hipLaunchKernelGGL((k_load<F32>), dim3(nblk), dim3(BS), 0, s, a, p);
  return hipGetLastError();
}
// One block per row: 1 where every cell of the row is plain fluid (all its flags are CELL_FLUID, 0).
__global__ __launch_bounds__(256) void k_row_fluid(const uint8_t* __restrict__ flags, uint32_t nx,
                                                   uint32_t* __restrict__ rowfluid) {
  const size_t base = (size_t)blockIdx.x * nx;
  uint32_t any = 0;
  for (uint32_t x = threadIdx.x; x < nx; x += 256) any |= flags[base + x];
  any = __syncthreads_or(any != 0);
  if (threadIdx.x == 0) rowfluid[blockIdx.x] = any ? 0u : 1u;
}
hipError_t launch_row_fluid(const uint8_t* flags, const Geo& g, uint32_t* rowfluid, hipStream_t s) {
  hipLaunchKernelGGL(k_row_fluid, dim3(g.ny * g.nz), dim3(256), 0, s, flags, g.nx, rowfluid);
  return hipGetLastError();
}
hipError_t launch_links(const uint8_t* flags, const int32_t* blist, uint32_t B, const Geo& g, uint32_t* bmask,
                        hipStream_t s) {
  if (B == 0) return hipSuccess;
  hipLaunchKernelGGL(k_links, dim3((B + 255) / 256), dim3(256), 0, s, flags, blist, B, g, bmask);
  return hipGetLastError();
}
uint32_t forces_blocks(uint32_t B) { return (B + BS - 1) / BS; }
hipError_t launch_forces(const StepArgs& a, bool fp16, const int32_t* blist, int aa_state, const float* halo_cur,
                         double* part, double* out, hipStream_t s) {
  const uint32_t nblk = forces_blocks(a.B);
  if (nblk > 0) {
    if (fp16) hipLaunchKernelGGL((k_forces<F16>), dim3(nblk), dim3(BS), 0, s, a, blist, aa_state, halo_cur, part);
    else hipLaunchKernelGGL((k_forces<F32>), dim3(nblk), dim3(BS), 0, s, a, blist, aa_state, halo_cur, part);
  }
  hipLaunchKernelGGL(k_forces_final, dim3(1), dim3(64), 0, s, part, nblk, out);
  return hipGetLastError();
}
uint32_t step_blocks(uint32_t N) { return (N + BS - 1) / BS; }

// Bulk tiles of 512 threads in FP32 storage take 64 KB of unused LDS, so 2 blocks (4 waves per SIMD) are resident
// per CU, as for the single-tunnel bulk kernel: on the 64-tunnel sweep the batched step measured 4.58-4.61 against
// 4.71 ms. Other shapes run uncapped: 256-thread FP32 blocks at 4 waves per SIMD (40 KB) measured 4.93 against
// 4.73 ms, and FP16 tiles gained nothing from 28, 40, 56 or 64 KB. Blocks handed out in XCD-contiguous runs, which
// speed up the single-tunnel kernel, were 1% slower here.
template <typename T, int BSZ, int V>
static void launch_batch_bulk_t(const StepArgs* targs, const BatchTile* tiles, uint32_t ntiles, double* part,
                                bool from_e, hipStream_t s) {
  constexpr int esz = (int)sizeof(typename T::elem);
  constexpr unsigned lds = (esz == 4 && BSZ == 512) ? 65536u : 0u;
  if (from_e) hipLaunchKernelGGL((k_step_batch_bulk<T, true, BSZ, V>), dim3(ntiles), dim3(BSZ), lds, s, targs, tiles, ntiles, part);
  else hipLaunchKernelGGL((k_step_batch_bulk<T, false, BSZ, V>), dim3(ntiles), dim3(BSZ), lds, s, targs, tiles, ntiles, part);
}
template <typename T, int BSZ>
static void launch_batch_bulk_v(const StepArgs* targs, const BatchTile* tiles, uint32_t ntiles, double* part,
                                bool from_e, const BatchShape& sh, hipStream_t s) {
  if (sh.cpt == 4) launch_batch_bulk_t<T, BSZ, 4>(targs, tiles, ntiles, part, from_e, s);
  else if (sh.cpt == 2) launch_batch_bulk_t<T, BSZ, 2>(targs, tiles, ntiles, part, from_e, s);
  else launch_batch_bulk_t<T, BSZ, 1>(targs, tiles, ntiles, part, from_e, s);
}
template <typename T, bool FLUID>
static void launch_batch_list_t(const StepArgs* targs, const BatchTile* list, uint32_t n, double* part, bool from_e,
                                hipStream_t s) {
  if (from_e) hipLaunchKernelGGL((k_step_batch_list<T, true, FLUID>), dim3(n), dim3(BS), 0, s, targs, list, n, part);
  else hipLaunchKernelGGL((k_step_batch_list<T, false, FLUID>), dim3(n), dim3(BS), 0, s, targs, list, n, part);
}
template <typename T>
static hipError_t launch_batch_all(const StepArgs* targs, const BatchTile* bulk, uint32_t nbulk,
                                   const BatchTile* fluid, uint32_t nfluid, const BatchTile* other, uint32_t nother,
                                   double* part, bool from_e, const BatchShape& sh, hipStream_t s,
                                   const SideLaunch* side) {
  hipError_t e;
  const bool fork = side && nother;
  if (fork) {
    if ((e = hipEventRecord(side->fork, s)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(side->stream, side->fork, 0)) != hipSuccess) return e;
    launch_batch_list_t<T, false>(targs, other, nother, part, from_e, side->stream);
    if ((e = hipEventRecord(side->join, side->stream)) != hipSuccess) return e;
  }
  if (nbulk) {
    if (sh.block == 128) launch_batch_bulk_v<T, 128>(targs, bulk, nbulk, part, from_e, sh, s);
    else if (sh.block == 512) launch_batch_bulk_v<T, 512>(targs, bulk, nbulk, part, from_e, sh, s);
    else launch_batch_bulk_v<T, 256>(targs, bulk, nbulk, part, from_e, sh, s);
  }
  if (nfluid) launch_batch_list_t<T, true>(targs, fluid, nfluid, part, from_e, s);
  if (fork) {
    if ((e = hipStreamWaitEvent(s, side->join, 0)) != hipSuccess) return e;
  } else if (nother) {
    launch_batch_list_t<T, false>(targs, other, nother, part, from_e, s);
  }
  return hipGetLastError();
}
hipError_t launch_step_batch(const StepArgs* targs, const BatchTile* bulk, uint32_t nbulk, const BatchTile* fluid,
                             uint32_t nfluid, const BatchTile* other, uint32_t nother, double* part, bool fp16,
                             bool from_e, const BatchShape& sh, hipStream_t s, const SideLaunch* side) {
  return fp16 ? launch_batch_all<F16>(targs, bulk, nbulk, fluid, nfluid, other, nother, part, from_e, sh, s, side)
              : launch_batch_all<F32>(targs, bulk, nbulk, fluid, nfluid, other, nother, part, from_e, sh, s, side);
}

template <typename T, int BSZ, int V>
static void launch_batch_one_t(const BatchOneArgs& p, bool from_e, uint32_t grid, hipStream_t s) {
  if (from_e)
    hipLaunchKernelGGL((k_step_batch_one<T, true, BSZ, V>), dim3(grid), dim3(BSZ), 0, s, p.targs, p.tiles, p.xcd_begin,
                       p.states, p.ntun, p.sync);
  else
    hipLaunchKernelGGL((k_step_batch_one<T, false, BSZ, V>), dim3(grid), dim3(BSZ), 0, s, p.targs, p.tiles, p.xcd_begin,
                       p.states, p.ntun, p.sync);
}
template <typename T, int BSZ>
static void launch_batch_one_v(const BatchOneArgs& p, bool from_e, const BatchShape& sh, hipStream_t s) {
  if (sh.cpt == 4) launch_batch_one_t<T, BSZ, 4>(p, from_e, sh.grid, s);
  else if (sh.cpt == 2) launch_batch_one_t<T, BSZ, 2>(p, from_e, sh.grid, s);
  else launch_batch_one_t<T, BSZ, 1>(p, from_e, sh.grid, s);
}
template <typename T>
static void launch_batch_one_b(const BatchOneArgs& p, bool from_e, const BatchShape& sh, hipStream_t s) {
  if (sh.block == 128) launch_batch_one_v<T, 128>(p, from_e, sh, s);
  else if (sh.block == 512) launch_batch_one_v<T, 512>(p, from_e, sh, s);
  else launch_batch_one_v<T, 256>(p, from_e, sh, s);
}
hipError_t launch_step_batch_one(const BatchOneArgs& p, bool fp16, bool from_e, const BatchShape& sh, hipStream_t s) {
  if (sh.grid == 0 || (sh.grid & 7u) != 0) return hipErrorInvalidValue;  // whole XCD rounds of blocks
  if (fp16) launch_batch_one_b<F16>(p, from_e, sh, s);
  else launch_batch_one_b<F32>(p, from_e, sh, s);
  return hipGetLastError();
}
hipError_t launch_finalize_batch(const double* part, const uint32_t* tile_begin, StepState* const* states,
                                 uint32_t ntun, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_batch, dim3(ntun), dim3(256), 0, s, part, tile_begin, states);
  return hipGetLastError();
}

}  // namespace aero
