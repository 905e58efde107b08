// Device-side building blocks of the D3Q19 A-A stream-collide step for gfx950.
//
// Memory layout in HBM
//   f       : one population buffer (A-A in-place streaming), structure of arrays f[k*N + idx], stored
//             as float or as shifted half h = (f - w_k) * 2^s.
//   A-A states: after an even number of steps (state E) the post-collision value f*_k(x) sits at
//             (x, opp(k)); after an odd number (state O) it sits at (x + c_k, k). The kernel reading
//             state E pulls from (x - c_k, opp(k)) and writes to (x + c_k, k); the kernel reading state O
//             reads (x, k) and writes (x, opp(k)). Each location is read and written by one thread only.
//   halo    : per domain face, the 5 inward-pointing post-collision populations of every face cell,
//             double-buffered by step parity, in FP32 for either storage. The reference clamps pull
//             coordinates at the domain faces (lbmCpuReference.ts:190-192), so a face cell may pull a value that
//             another face cell also consumes in the same step; those reads go to the previous parity's halo
//             instead of the in-place buffer.
//   bown    : for each boundary cell, its own post-collision populations of the previous step (the
//             Bouzidi rule reads f*_d(x) and f*_{opp d}(x), lbmCpuReference.ts:203,215, which the
//             in-place buffer no longer holds race-free).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19.h"
#include "mlp.h"

namespace aero {

// Division by a runtime constant for n < 2^31: q = (mulhi(n, m) + n) >> s.
struct FastDiv {
  uint32_t d, m, s;
  static FastDiv make(uint32_t d_) {
    FastDiv f;
    f.d = d_;
    f.s = 0;
    while ((1ull << f.s) < d_) ++f.s;
    const unsigned __int128 one = 1;
    const uint64_t mm = (uint64_t)(((one << (32 + f.s)) + d_ - 1) / d_);
    f.m = (uint32_t)(mm - (1ull << 32));
    return f;
  }
  __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return (__umulhi(n, m) + n) >> s;
#else
    return (uint32_t)((((uint64_t)n * m) >> 32) + n) >> s;
#endif
  }
};

struct StepState {
  long long step;      // steps completed since reset
  double corr;         // mass-anchor factor of the last step (1.0 when it did not fire)
  long long anchors;   // number of steps on which the anchor fired
  double sum;          // last post-boundary sum of rho over non-solid cells
  double mean;         // last mean rho
  long long nfluid;    // non-solid cell count
};

// Bulk-kernel addressing per storage element size: a buffer descriptor with 32-bit offsets (fewer VGPRs and
// SGPRs; faster for FP16 storage) or flat 64-bit addresses (faster for FP32 storage).
#ifndef AERO_BULK_BUF_F32
#define AERO_BULK_BUF_F32 0
#endif
#ifndef AERO_BULK_BUF_F16
#define AERO_BULK_BUF_F16 1
#endif
__host__ __device__ constexpr bool bulk_buffer_addressing(int elem_bytes) {
  return elem_bytes == 4 ? AERO_BULK_BUF_F32 != 0 : AERO_BULK_BUF_F16 != 0;
}

// Bulk kernel: issue the population loads before testing the flags, so a wave waits for one round trip
// rather than two. The FP32 bulk kernel gains 1%; the FP16 one gains less than the listed cells beside it then
// lose (their side stream ends after it: steps measured 1.5% slower), so FP16 tests the flags first.
__host__ __device__ constexpr bool bulk_loads_before_flags(int elem_bytes) { return elem_bytes == 4; }

// Bulk-kernel launch shape per storage element size: threads per block, and unused dynamic LDS per block, which
// caps the blocks resident per CU. FP32: 512 threads and 64 KB, so 2 blocks (16 waves) per CU; fewer
// concurrent access streams measured 1.2% faster than the full occupancy its registers allow (and 2 waves per
// SIMD 15% slower). At the same 4 waves per SIMD, 256-thread blocks (40 KB) measured 0.5% slower and 1024-thread
// blocks (96 KB) 8% slower (bulk 0.979 and 1.055 against 0.974 ms). FP16, issue-bound, runs 256-thread blocks; 28 KB caps them at 5 per CU where their registers
// allow 6, which leaves the listed-cell kernels on the side stream room to run beside it: at 512x256x256 the step
// took 0.5499 against 0.5535 ms, the side stream ending 12 against 28 us after the bulk kernel (40 KB and 52 KB,
// 4 and 3 blocks, slowed the bulk kernel by 13%).
#ifndef AERO_BULK_LDS_F32
#define AERO_BULK_LDS_F32 65536u
#endif
#ifndef AERO_BULK_LDS_F16
#define AERO_BULK_LDS_F16 28672u
#endif
#ifndef AERO_BULK_THREADS_F32
#define AERO_BULK_THREADS_F32 512
#endif
__host__ __device__ constexpr int bulk_block_threads(int elem_bytes) { return elem_bytes == 4 ? AERO_BULK_THREADS_F32 : 256; }
__host__ __device__ constexpr unsigned bulk_block_lds(int elem_bytes) {
  return elem_bytes == 4 ? AERO_BULK_LDS_F32 : AERO_BULK_LDS_F16;
}

// Logical block of dispatched block b of nb. Blocks go round-robin to the 8 XCDs (b mod 8); with XCD set, XCD x
// is given the contiguous run of logical blocks [x q + min(x, r), ... + q + (x < r)), q = nb / 8, r = nb mod 8,
// in dispatch order, so neighbouring blocks share one L2. Measured per storage element size: on for FP32
// populations (the access pattern alone: 0.916 against 0.955 ms per 512x256x256 grid), off for FP16 (0.471
// against 0.453 ms).
#ifndef AERO_BULK_XCD_F32
#define AERO_BULK_XCD_F32 1
#endif
#ifndef AERO_BULK_XCD_F16
#define AERO_BULK_XCD_F16 0
#endif
template <int ELEM_BYTES>
__device__ __forceinline__ uint32_t xcd_chunk_block(uint32_t b, uint32_t nb) {
  if (!(ELEM_BYTES == 4 ? AERO_BULK_XCD_F32 : AERO_BULK_XCD_F16)) return b;
  const uint32_t x = b & 7u, j = b >> 3, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + j;
}

struct Geo {
  uint32_t nx, ny, nz, N;
  FastDiv dnx, dny;
  int64_t loff[19];         // population-layout displacement of the neighbour x + c_k (see cell_li)
  int64_t lmax;             // largest |loff[k]|
  int64_t lrow, lplane;     // 19 nx and 19 nx ny: loff[k] = cx + cy lrow + cz lplane
  uint32_t bufok;           // the bulk kernels may run: where they use buffer offsets (2 lmax + 19 nx slots),
                            // those stay below 2^31 bytes
  uint32_t hoff[6];         // offset of each face block in a halo buffer
  uint32_t hsz[6];          // cells per face
};

// Population layout: row-blocked structure of arrays. The 19 populations of one x-row (row = y + ny z) are
// stored direction after direction, f[(row * 19 + k) * nx + x]. Within a row each direction is a contiguous
// run of nx values, so a wave's accesses coalesce as in plain structure of arrays, while the 19 runs that
// a group of cells touches lie within 19 nx values of each other instead of 19 N: far fewer DRAM pages are
// open at once. Slot of direction k of the cell with layout index li is li + k nx; the layout index of
// (x, row) is x + 19 nx row, and the neighbour x + c_k has li + loff[k], loff[k] = cx + 19 nx (cy + ny cz).
__host__ __device__ inline size_t cell_li(const Geo& g, uint32_t x, uint32_t row) {
  return (size_t)x + (size_t)19 * g.nx * row;
}

struct Phys {
  float tau0;               // 3 nu + 1/2
  float smag;               // 18 sqrt(2) Cs^2
  double u_target;
  int32_t ramp;
  int32_t profile;          // 0 uniform, 1 power law (ground), 2 log law
  int32_t noslip;
  int32_t sides;            // side walls active (nz > 4)
  double z0, zref;
  float scale, inv_scale;   // FP16 codec 2^s and 2^-s
  float nu;                 // molecular viscosity (learned-closure features)
};

struct StepArgs {
  Geo g;
  Phys ph;
  void* f;
  const uint8_t* __restrict__ flags;
  const uint32_t* __restrict__ rowfluid;  // per row y + ny z: 1 where every cell of the row is plain fluid
  const int32_t* __restrict__ bmap;
  const uint32_t* __restrict__ bmask;
  const float* __restrict__ bq;
  float* __restrict__ bown;
  uint32_t B;
  const float* __restrict__ halo_in;
  float* __restrict__ halo_out;
  const StepState* __restrict__ state;
  double* __restrict__ partials;
  float* __restrict__ mac;   // readout: ux, uy, uz, rho (4 x N) or null
  float* __restrict__ snap;  // readout: post-step populations (19 x N) or null
  const struct MlpDev* __restrict__ mlp;  // learned closure weights (null: Smagorinsky)
  const int32_t* __restrict__ spec;       // cells outside bulk groups that the step visits (split step):
  uint32_t nspec;                         // nspec in all, the first nface of them plain fluid cells
  uint32_t nface;
  const double* __restrict__ uprof;       // inflow profile factor of each row y (ny values)
};

struct F19 {
  float v[19];
};

// A run of consecutive cells of one tunnel (batched launch).
// A tile of batched work: count cells from cell start of one tunnel (bulk tiles) or count entries from entry
// start of its listed cells (list tiles); part is the tile's slot in the per-tunnel-contiguous partials.
struct BatchTile {
  uint32_t tunnel, start, count, part;
};
// A tile of the one-launch batched step (k_step_batch_one): tk = kind << 30 | tunnel (kinds TILE_*), start and count
// as in BatchTile.
struct BatchTileP {
  uint32_t tk, start, count, reserved;
};
constexpr uint32_t TILE_BULK = 0u, TILE_FLUID = 1u, TILE_OTHER = 2u;
// Tables of the one-launch batched step. XCD x takes the tiles [xcd_begin[x], xcd_begin[x + 1]); sync holds the
// block ticket, the XCDs' work counters and the tunnels' fixed-point density totals (batch_sync_bytes; zero between
// launches).
struct BatchOneArgs {
  const StepArgs* targs;
  const BatchTileP* tiles;
  const uint32_t* xcd_begin;
  StepState* const* states;
  uint32_t ntun;
  uint32_t* sync;
};
__host__ __device__ constexpr size_t batch_sync_bytes(uint32_t ntun) { return 1024 + 8 * (size_t)ntun; }
// Launch shape of the batched step: threads per block, cells per thread, and the grid: 0 for one block per tile
// (bulk, list and finalize launches per step), otherwise the block count of the one-launch persistent form.
struct BatchShape {
  int block, cpt;
  uint32_t grid;
};

// ---- storage codecs ---------------------------------------------------------------------------
// Pop<E, NT>: population element type E (float or shifted half) and whether the in-place buffer is accessed
// with non-temporal loads/stores.
template <typename E, bool NT>
struct Pop;
typedef float f32x2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float f32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned short u16x2u __attribute__((ext_vector_type(2), aligned(2)));
typedef unsigned short u16x4u __attribute__((ext_vector_type(4), aligned(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Buffer descriptor over the populations for the bulk kernel's stores: base address (block-uniform), no
// stride, and a record count above any offset the kernel forms (all offsets stay below 2^31). The last word is
// the 32-bit data format gfx9 buffer instructions expect.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_rsrc(void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, 0x7FFFFFFF, 0x00020000);
}

template <bool NT>
struct Pop<float, NT> {
  using elem = float;
  // V consecutive cells of direction k (the address needs only element alignment)
  template <int V>
  static __device__ __forceinline__ void ldv(const float* p, int k, const Phys& ph, float* out) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      out[0] = ld(p, k, ph);
    } else if constexpr (V == 2) {
      const f32x2u v = *reinterpret_cast<const f32x2u*>(p);
      out[0] = v[0]; out[1] = v[1];
    } else {
      const f32x4u v = *reinterpret_cast<const f32x4u*>(p);
      out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; out[3] = v[3];
    }
  }
  template <int V>
  static __device__ __forceinline__ void stv(float* p, const float* in, int k, const Phys& ph) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      st(p, in[0], k, ph);
    } else if constexpr (V == 2) {
      f32x2u v;
      v[0] = in[0]; v[1] = in[1];
      *reinterpret_cast<f32x2u*>(p) = v;
    } else {
      f32x4u v;
      v[0] = in[0]; v[1] = in[1]; v[2] = in[2]; v[3] = in[3];
      *reinterpret_cast<f32x4u*>(p) = v;
    }
  }
  // ldv / stv through a buffer descriptor: byte offset voff (per lane) + soff (uniform) from its base
  template <int V>
  static __device__ __forceinline__ void ldv_buf(__amdgpu_buffer_rsrc_t r, int voff, int soff, int, const Phys&,
                                                 float* out) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      out[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else if constexpr (V == 2) {
      const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
      out[0] = __uint_as_float(v[0]); out[1] = __uint_as_float(v[1]);
    } else {
      const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
      for (int c = 0; c < 4; ++c) out[c] = __uint_as_float(v[c]);
    }
  }
  template <int V>
  static __device__ __forceinline__ void stv_buf(__amdgpu_buffer_rsrc_t r, int voff, int soff, const float* in, int,
                                                 const Phys&) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(in[0]), r, voff, soff, 0);
    } else if constexpr (V == 2) {
      u32x2 v;
      v[0] = __float_as_uint(in[0]); v[1] = __float_as_uint(in[1]);
      __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
    } else {
      u32x4 v;
      v[0] = __float_as_uint(in[0]); v[1] = __float_as_uint(in[1]);
      v[2] = __float_as_uint(in[2]); v[3] = __float_as_uint(in[3]);
      __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, 0);
    }
  }
  static __device__ __forceinline__ float ld(const float* p, int, const Phys&) {
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
  }
  static __device__ __forceinline__ void st(float* p, float v, int, const Phys&) {
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
  }
};
template <bool NT>
struct Pop<__half, NT> {
  using elem = __half;
  template <int V>
  static __device__ __forceinline__ void ldv(const __half* p, int k, const Phys& ph, float* out) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      out[0] = ld(p, k, ph);
    } else if constexpr (V == 2) {
      const u16x2u v = *reinterpret_cast<const u16x2u*>(p);
      for (int c = 0; c < 2; ++c) out[c] = w32(k) + __half2float(__ushort_as_half(v[c])) * ph.inv_scale;
    } else {
      const u16x4u v = *reinterpret_cast<const u16x4u*>(p);
      for (int c = 0; c < 4; ++c) out[c] = w32(k) + __half2float(__ushort_as_half(v[c])) * ph.inv_scale;
    }
  }
  template <int V>
  static __device__ __forceinline__ void stv(__half* p, const float* in, int k, const Phys& ph) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    if constexpr (V == 1) {
      st(p, in[0], k, ph);
    } else if constexpr (V == 2) {
      u16x2u v;
      for (int c = 0; c < 2; ++c) v[c] = __half_as_ushort(__float2half_rn(fmaf(in[c], ph.scale, -(w32(k) * ph.scale))));
      *reinterpret_cast<u16x2u*>(p) = v;
    } else {
      u16x4u v;
      for (int c = 0; c < 4; ++c) v[c] = __half_as_ushort(__float2half_rn(fmaf(in[c], ph.scale, -(w32(k) * ph.scale))));
      *reinterpret_cast<u16x4u*>(p) = v;
    }
  }
  template <int V>
  static __device__ __forceinline__ void ldv_buf(__amdgpu_buffer_rsrc_t r, int voff, int soff, int k, const Phys& ph,
                                                 float* out) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    unsigned short u[V];
    if constexpr (V == 1) {
      u[0] = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    } else if constexpr (V == 2) {
      const unsigned w = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
      u[0] = (unsigned short)(w & 0xFFFFu); u[1] = (unsigned short)(w >> 16);
    } else {
      const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
      u[0] = (unsigned short)(w[0] & 0xFFFFu); u[1] = (unsigned short)(w[0] >> 16);
      u[2] = (unsigned short)(w[1] & 0xFFFFu); u[3] = (unsigned short)(w[1] >> 16);
    }
    for (int c = 0; c < V; ++c) out[c] = w32(k) + __half2float(__ushort_as_half(u[c])) * ph.inv_scale;
  }
  template <int V>
  static __device__ __forceinline__ void stv_buf(__amdgpu_buffer_rsrc_t r, int voff, int soff, const float* in, int k,
                                                 const Phys& ph) {
    static_assert(V == 1 || V == 2 || V == 4, "1, 2 or 4 cells");
    unsigned short u[V];
    for (int c = 0; c < V; ++c) u[c] = __half_as_ushort(__float2half_rn(fmaf(in[c], ph.scale, -(w32(k) * ph.scale))));
    if constexpr (V == 1) {
      __builtin_amdgcn_raw_buffer_store_b16(u[0], r, voff, soff, 0);
    } else if constexpr (V == 2) {
      __builtin_amdgcn_raw_buffer_store_b32((unsigned)u[0] | ((unsigned)u[1] << 16), r, voff, soff, 0);
    } else {
      u32x2 v;
      v[0] = (unsigned)u[0] | ((unsigned)u[1] << 16);
      v[1] = (unsigned)u[2] | ((unsigned)u[3] << 16);
      __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
    }
  }
  static __device__ __forceinline__ float ld(const __half* p, int k, const Phys& ph) {
    const unsigned short* q = reinterpret_cast<const unsigned short*>(p);
    const unsigned short u = NT ? __builtin_nontemporal_load(q) : *q;
    return w32(k) + __half2float(__ushort_as_half(u)) * ph.inv_scale;
  }
  static __device__ __forceinline__ void st(__half* p, float v, int k, const Phys& ph) {
    // (v - w) 2^s as one fma: the products by the power of two are exact, so it rounds as the difference
    const unsigned short u = __half_as_ushort(__float2half_rn(fmaf(v, ph.scale, -(w32(k) * ph.scale))));
    unsigned short* q = reinterpret_cast<unsigned short*>(p);
    if (NT) __builtin_nontemporal_store(u, q);
    else *q = u;
  }
};
using F32 = Pop<float, false>;
using F16 = Pop<__half, false>;

__device__ __forceinline__ float apply_corr(float v, double corr) { return (float)((double)v * corr); }

// corr != 1.0, tested on the bit pattern (the two agree for every value, NaN included). The factor is one
// value per step, and an integer compare keeps the test on the scalar unit, so the correction is skipped by a
// branch rather than computed on every lane and discarded by selects.
__device__ __forceinline__ bool corr_fires(double corr) {
  return __builtin_bit_cast(unsigned long long, corr) != 0x3FF0000000000000ull;
}

// Reference equilibrium with double weights (lbmConstants.ts:142-152), rounded to float: the initial state.
__device__ __forceinline__ void feq_init(float* out, double rho, double ux, double uy, double uz) {
  const double uSq = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int d = 0; d < 19; ++d) {
    const double cu = 3.0 * (cx(d) * ux + cy(d) * uy + cz(d) * uz);
    out[d] = (float)(w64(d) * rho * (1.0 + cu + 0.5 * cu * cu - 1.5 * uSq));
  }
}

// Inlet velocity of step s for row iy (lbmCpuReference.ts:334-342 with the ramp of testRunner.ts:89-96).
__device__ __forceinline__ double inlet_u(const Phys& ph, long long step) {
  double u = ph.u_target;
  if (ph.ramp > 0 && step + 1 <= ph.ramp) {
    const double s = sin(0.5 * M_PI * (double)(step + 1) / (double)ph.ramp);
    u = ph.u_target * s * s;
  }
  return u;
}
// Inflow velocity of row iy: u times the row's profile factor, which the host tabulates when the parameters are
// set (1 for uniform inflow; the ground power law and the log law of lbmCpuReference.ts:122-125, 335-338).
__device__ __forceinline__ double profile_u(const StepArgs& a, double u, uint32_t iy) { return u * a.uprof[iy]; }

// ---- collision --------------------------------------------------------------------------------------
// Regularized BGK + Smagorinsky in FP32 (lbmCpuReference.ts:233-304), written over the nine pairs
// (k, opp k), k = 1, 3, ..., 17. For a pair the equilibrium splits into an even part A = w rho (1 - 1.5 u.u +
// 4.5 (c.u)^2) and an odd part B = 3 w rho c.u (feq_k = A + B, feq_opp = A - B); the second moments of the
// non-equilibrium part need only f_k + f_opp - 2A, and the regularized term 4.5 w Q_k : Pi is the same for
// both members, so each pair is relaxed once. Signs of c are applied by selection, never by multiplying with
// 0 or +-1 (which IEEE rules would not fold away).

// c_k . (u, v, w) for a compile-time k.
__device__ __forceinline__ float cdot(int k, float u, float v, float w) {
  float c = 0.0f;
  if (cx(k)) c = cx(k) > 0 ? u : -u;
  if (cy(k)) c = cx(k) ? (cy(k) > 0 ? c + v : c - v) : (cy(k) > 0 ? v : -v);
  if (cz(k)) c = (cx(k) || cy(k)) ? (cz(k) > 0 ? c + w : c - w) : (cz(k) > 0 ? w : -w);
  return c;
}

// The same equilibrium in FP32 for the boundary rules of every step (inlet and outlet cells, walls). The bulk
// kernel runs the row ends' rules inside otherwise uniform waves, one active lane per wave, where FP64
// arithmetic would take its full issue time at half rate; FP32 moves the result by a few ulps of f.
__device__ __forceinline__ void feq_bc(float* out, float rho, float ux, float uy, float uz) {
  const float base = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
  out[0] = (w32(0) * rho) * base;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int k = 2 * p + 1;
    const float cu = 3.0f * cdot(k, ux, uy, uz);
    const float wr = w32(k) * rho;
    const float A = wr * (base + 0.5f * cu * cu);
    out[k] = A + wr * cu;
    out[k + 1] = A - wr * cu;
  }
}

// Moments of the pulled populations f: clamped rho and velocity (lbmCpuReference.ts:233-246) and the
// non-equilibrium stress P = (xx, yy, zz, xy, xz, yz) (:256-271).
// The accumulators start at -0, the identity of float addition (-0 + t = t for every t, where +0 + -0 = +0), so
// each first accumulation folds into its operand: nine additions per cell fewer than starting at +0.
__device__ __forceinline__ void collide_moments(const float* f, float& rho, float& ux, float& uy, float& uz, float* P) {
  float r = f[0], mx = -0.f, my = -0.f, mz = -0.f;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int k = 2 * p + 1;
    const float s = f[k] + f[k + 1], t = f[k] - f[k + 1];
    r += s;
    if (cx(k)) mx = cx(k) > 0 ? mx + t : mx - t;
    if (cy(k)) my = cy(k) > 0 ? my + t : my - t;
    if (cz(k)) mz = cz(k) > 0 ? mz + t : mz - t;
  }
  r = fmaxf(0.5f, fminf(1.8f, r));
  const float inv = __builtin_amdgcn_rcpf(r);
  float u = mx * inv, v = my * inv, w = mz * inv;
  const float uMag2 = u * u + v * v + w * w;
  if (uMag2 > 0.22f * 0.22f) {
    const float sc = 0.22f / sqrtf(uMag2);
    u *= sc; v *= sc; w *= sc;
  }
  const float base = 1.0f - 1.5f * (u * u + v * v + w * w);
  float pxx = -0.f, pyy = -0.f, pzz = -0.f, pxy = -0.f, pxz = -0.f, pyz = -0.f;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int k = 2 * p + 1;
    const float cu = 3.0f * cdot(k, u, v, w);
    const float A = (w32(k) * r) * (base + 0.5f * cu * cu);
    const float ne = (f[k] + f[k + 1]) - 2.0f * A;  // non-equilibrium part of f_k + f_opp
    if (cx(k)) pxx += ne;
    if (cy(k)) pyy += ne;
    if (cz(k)) pzz += ne;
    if (cx(k) && cy(k)) pxy = cx(k) * cy(k) > 0 ? pxy + ne : pxy - ne;
    if (cx(k) && cz(k)) pxz = cx(k) * cz(k) > 0 ? pxz + ne : pxz - ne;
    if (cy(k) && cz(k)) pyz = cy(k) * cz(k) > 0 ? pyz + ne : pyz - ne;
  }
  P[0] = pxx; P[1] = pyy; P[2] = pzz; P[3] = pxy; P[4] = pxz; P[5] = pyz;
  rho = r; ux = u; uy = v; uz = w;
}

// Regularized post-collision populations f*_k = feq_k + (1 - omega) 4.5 w_k Q_k : P (lbmCpuReference.ts:284-304).
__device__ __forceinline__ void collide_relax(float* f, float r, float u, float v, float w, const float* P, float omega) {
  const float base = 1.0f - 1.5f * (u * u + v * v + w * w);
  const float om1 = 1.0f - omega;
  const float tr3 = (P[0] + P[1] + P[2]) * (1.0f / 3.0f);
  f[0] = (w32(0) * r) * base - om1 * (4.5f * w32(0)) * tr3;
#pragma unroll
  for (int p = 0; p < 9; ++p) {
    const int k = 2 * p + 1;
    const float cu = 3.0f * cdot(k, u, v, w);
    const float wr = w32(k) * r;
    const float A = wr * (base + 0.5f * cu * cu);
    const float B = wr * cu;
    float qp = -tr3;  // Q_k : P with Q = c c - I/3, off-diagonal terms counted twice
    if (cx(k)) qp += P[0];
    if (cy(k)) qp += P[1];
    if (cz(k)) qp += P[2];
    if (cx(k) && cy(k)) qp = cx(k) * cy(k) > 0 ? qp + 2.0f * P[3] : qp - 2.0f * P[3];
    if (cx(k) && cz(k)) qp = cx(k) * cz(k) > 0 ? qp + 2.0f * P[4] : qp - 2.0f * P[4];
    if (cy(k) && cz(k)) qp = cy(k) * cz(k) > 0 ? qp + 2.0f * P[5] : qp - 2.0f * P[5];
    const float F = A + om1 * (4.5f * w32(k)) * qp;
    f[k] = F + B;
    f[k + 1] = F - B;
  }
}

// Full collision with the Smagorinsky relaxation time (lbmCpuReference.ts:273-282). f holds the pulled
// populations on entry and the post-collision populations on exit.
__device__ __forceinline__ void collide(float* f, float& rho, float& ux, float& uy, float& uz, const Phys& ph) {
  float P[6];
  collide_moments(f, rho, ux, uy, uz, P);
  const float piNorm = sqrtf(P[0] * P[0] + P[1] * P[1] + P[2] * P[2] + 2.0f * (P[3] * P[3] + P[4] * P[4] + P[5] * P[5]));
  const float tauEff = 0.5f * (ph.tau0 + sqrtf(ph.tau0 * ph.tau0 + ph.smag * piNorm * __builtin_amdgcn_rcpf(rho)));
  const float omega = __builtin_amdgcn_rcpf(fmaxf(0.5015f, fminf(1.95f, tauEff)));
  collide_relax(f, rho, ux, uy, uz, P, omega);
}

// Block sum of one double per thread (wave shuffles, then LDS across the waves of the block).
template <int BS>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[BS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) t += sh[i];
  }
  return t;  // valid in thread 0
}

}  // namespace aero
