// CPU oracle for the D3Q19 wind-tunnel path (TEST INFRASTRUCTURE ONLY).
//
// This file restates the reference's CPU solver so that the HIP product can be checked against it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the library built from
// it, and only as the checker; the product library never links, loads or calls it.
//
// Reference files restated (necrolingus/aerosim):
//   src/solver/lbmConstants.ts:10-66,142-152      D3Q19 tables, weights, equilibrium
//   src/solver/lbmCpuReference.ts:114-495          init, collide+stream, boundaries, mass anchor, stats
//   src/solver/forceIntegrator.ts:106-315,321-424  momentum exchange, coefficients, tracker
//   src/geometry/voxelizer.ts:65-281               voxelizer
//   src/geometry/bvh.ts:34-589                     BVH build and queries
//   src/geometry/frontalArea.ts:22-179             projected frontal area
//
// Two storage modes:
//   fp64  : every array, scratch value and weight is double ("the FP64 oracle").
//   ref32 : reference-faithful. Populations, macros, fLocal/feq scratch are float (the reference's
//           Float32Array); arithmetic is double in the reference's expression order; the collision
//           and force integrator use float-rounded weights (LBM_WEIGHTS_F32), the boundary
//           equilibria use double weights.
// Build with -ffp-contract=off so no fused multiply-add changes a rounding.
//
// Pinning (what checks each part; the checks are in tests/test_oracle.py):
//   tables/equilibrium  : moment identities (sum w = 1, sum w c = 0, sum w c c = 1/3 delta, feq moments)
//   collision           : mass/momentum conservation with clamps inactive, Pi = 0 => f* = feq,
//                         Cs = 0 => tau = tau0, brute-force single cell in Python
//   Bouzidi             : q = 0.5 reduces to half-way bounce-back, linearity in q,
//                         zero inflow => |F| < 1e-6 (phase4.test.ts:11-50)
//   voxelizer / BVH     : 1 m cube at h = 0.05 => 8000 solid cells (phase2.test.ts:42-60),
//                         q vs analytic sphere mean < 1e-4 (phase2.test.ts:84-129), link consistency
//                         (phase4.test.ts:52-117), non-manifold 3-ray vote (phase2.test.ts:62-82),
//                         brute-force point-in-mesh
//   frontal area        : sphere pi r^2 within 0.5 %, cube 1, rotated cube sqrt 2 (phase1.test.ts:11-37)
//   boundary rules      : moments of the post-step populations against the macro fields
//                         (tests/test_oracle_boundaries.py): free-slip floor and ceiling leave the equilibrium of
//                         (rho, u_x, 0, u_z), side walls of (rho, u_x, u_y, 0), the no-slip floor w_k rho; BOUNDARY
//                         cells on the floor are left as the collision wrote them (identical under free-slip and
//                         no-slip); the power-law and log-law inflows and the ground init carry their profiles;
//                         the outlet reads its neighbour's pre-wall velocity (recomputed from the pre-step state);
//                         every planted slip of these rules fails a check
//   mass anchor         : a step with and without it: one common factor 1/mean on every non-solid value, solids
//                         untouched, mean 1 afterwards, and no change below the 1e-4 threshold
//   mass statistics     : hand cases of the NaN cut-off and the solid skip
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// lbmConstants.ts:10-48 (velocity order and opposite table).
const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// lbmConstants.ts:56-63: double weights and their float-rounded copy.
double W64[19];
double W32[19];
struct WeightInit {
  WeightInit() {
    for (int d = 0; d < 19; ++d) {
      W64[d] = d == 0 ? 1.0 / 3.0 : (d <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
      W32[d] = (double)(float)W64[d];
    }
  }
} g_weight_init;

const uint8_t CELL_FLUID = 0, CELL_SOLID = 1, CELL_BOUNDARY = 2;

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// lbmConstants.ts:142-152 (double weights).
inline double feq64(int d, double rho, double ux, double uy, double uz) {
  const double cu = 3.0 * (CX[d] * ux + CY[d] * uy + CZ[d] * uz);
  const double uSq = ux * ux + uy * uy + uz * uz;
  return W64[d] * rho * (1.0 + cu + 0.5 * cu * cu - 1.5 * uSq);
}

// Counter hash for the seeded initial velocity noise (splitmix64 finaliser). The product library uses
// the same definition so both sides start from identical fields.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double noise_unit(uint64_t seed, uint64_t idx, int comp) {
  // uniform in [-1, 1): 53 random bits
  const uint64_t h = mix64(seed ^ mix64(idx * 4ull + (uint64_t)comp));
  return ((double)(h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}

// Inlet velocity profile. Mode 0: uniform; mode 1: the reference's ground power law
// (lbmCpuReference.ts:122-125, 335-338); mode 2: log law u_ref*ln(z/z0)/ln(zref/z0), z = iy + 0.5.
inline double inlet_profile(int mode, double u, int iy, int ny, double z0, double zref) {
  if (mode == 1) {
    const double yFrac = (iy + 0.5) / ny;
    return u * std::min(1.15, 0.35 + 0.8 * std::pow(yFrac, 0.3));
  }
  if (mode == 2) {
    const double z = iy + 0.5;
    const double r = std::log(z / z0) / std::log(zref / z0);
    return u * (r > 0.0 ? r : 0.0);
  }
  return u;
}

// ---------------------------------------------------------------------------------------------
// Learned closure (no counterpart in the reference; the build's own extension of lbmCpuReference.ts:280-282): a
// 6 -> 32 -> 32 -> 1 ReLU network maps invariants of the non-equilibrium stress per unit density P to a factor
// g = softplus(y + ln(e - 1)) on the Smagorinsky eddy viscosity for the same stress, tau = tau0 + (tau_S - tau0) g.
// The forward pass rounds the operands of its two matrix products to BF16 (as the GPU's matrix-core kernels do,
// the first layer's bias included as a BF16 pair) and computes in FP64 otherwise; it is pinned against the finite-difference-checked numpy oracle
// (oracle/mlp_oracle.py), the features against closed forms.
constexpr int MLP_H = 32, MLP_IN = 6, MLP_NPARAM = 1313;

inline double bf16_round(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

// Features of P = (xx, yy, zz, xy, xz, yz): magnitude, isotropic share, deviatoric share, Lode-type shape, the
// molecular viscosity and the Smagorinsky viscosity for the same stress (smag = 18 sqrt(2) Cs^2).
inline void closure_features(const double* P, double nu, double tau0, double smag, double* x) {
  const double pxx = P[0], pyy = P[1], pzz = P[2], pxy = P[3], pxz = P[4], pyz = P[5];
  const double n2 = pxx * pxx + pyy * pyy + pzz * pzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz);
  const double pn = std::sqrt(n2);
  const double tr = pxx + pyy + pzz, m = tr / 3.0;
  const double dxx = pxx - m, dyy = pyy - m, dzz = pzz - m;
  const double d2 = dxx * dxx + dyy * dyy + dzz * dzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz);
  const double det = dxx * (dyy * dzz - pyz * pyz) - pxy * (pxy * dzz - pyz * pxz) + pxz * (pxy * pyz - dyy * pxz);
  x[0] = std::log(pn + 1e-12) * 0.1 + 1.0;
  x[1] = tr / (pn + 1e-20);
  x[2] = d2 / (n2 + 1e-30);
  x[3] = std::min(1.0, std::max(-1.0, std::sqrt(6.0) * 3.0 * det / (d2 * std::sqrt(d2) + 1e-30)));
  x[4] = std::log(nu) * 0.1 + 1.0;
  const double tauS = 0.5 * (tau0 + std::sqrt(tau0 * tau0 + smag * pn));
  x[5] = std::log1p(std::max(0.0, (tauS - tau0) / 3.0) / nu) * 0.25;
}

// Network output y for features x (parameters: W1[32][6], b1[32], W2[32][32], b2[32], w3[32], b3).
inline double closure_forward(const double* prm, const double* x) {
  const double *W1 = prm, *b1 = prm + 192, *W2 = prm + 224, *b2 = prm + 1248, *w3 = prm + 1280, b3 = prm[1312];
  double h1[MLP_H], y = b3;
  for (int i = 0; i < MLP_H; ++i) {
    const double bhi = bf16_round(b1[i]);  // the bias as a BF16 pair (16 significant bits), as the GPU applies it
    double z = bhi + bf16_round(b1[i] - bhi);
    for (int k = 0; k < MLP_IN; ++k) z += bf16_round(x[k]) * bf16_round(W1[i * MLP_IN + k]);
    h1[i] = z > 0.0 ? z : 0.0;
  }
  for (int i = 0; i < MLP_H; ++i) {
    double z = b2[i];
    for (int k = 0; k < MLP_H; ++k) z += bf16_round(h1[k]) * bf16_round(W2[i * MLP_H + k]);
    y += w3[i] * (z > 0.0 ? z : 0.0);
  }
  return y;
}

inline double softplus(double y) { return y > 20.0 ? y : std::log1p(std::exp(y)); }

struct Params {
  double uInlet = 0.08, nu = 0.0005, cs = 0.12;
  int groundNoSlip = 0;
  int profile = 0;  // 0 uniform, 1 ref power law, 2 log law
  double z0 = 0.5, zref = 64.0;
  bool closure = false;               // learned closure instead of Smagorinsky
  double mlp[MLP_NPARAM] = {};        // its parameters
};

// ---------------------------------------------------------------------------------------------
// Solver (lbmCpuReference.ts:55-495) templated on storage type S (float = ref32, double = fp64).
// computeFluidMassStats (lbmCpuReference.ts:461-495) over given fields: skip solid cells, stop at the first cell
// (in index order) whose rho or velocity is NaN, and over the cells before it sum rho, count them and keep the
// largest speed. out: totalMass meanRho maxSpeed hasNaN fluidCount.
template <typename R>
void fluid_mass_stats(const uint8_t* flags, const R* rho, const R* ux, const R* uy, const R* uz, long N, double* out) {
  double total = 0, maxSpeed = 0;
  long cnt = 0;
  bool nan = false;
  for (long i = 0; i < N; ++i) {
    if (flags[i] == CELL_SOLID) continue;
    const double r = rho[i], u = ux[i], v = uy[i], w = uz[i];
    if (std::isnan(r) || std::isnan(u) || std::isnan(v) || std::isnan(w)) { nan = true; break; }
    total += r;
    ++cnt;
    const double spd = std::sqrt(u * u + v * v + w * w);
    if (spd > maxSpeed) maxSpeed = spd;
  }
  out[0] = total; out[1] = cnt > 0 ? total / cnt : 1.0; out[2] = maxSpeed; out[3] = nan ? 1.0 : 0.0;
  out[4] = (double)cnt;
}

struct SolverBase {
  virtual ~SolverBase() {}
  int nx, ny, nz;
  long N;
  Params p;
  long stepIndex = 0;
  int anchorCount = 0;
  bool anchorOn = true;  // tests switch the anchor off to compare a step with and without it
  std::vector<uint8_t> flags;
  std::vector<int32_t> bmap, blist;
  std::vector<float> q;
  int nb = 0;
  virtual void reset(uint64_t seed, double noiseAmp) = 0;
  virtual void step(int n) = 0;
  virtual void getF(double* out) const = 0;
  virtual void setF(const double* in) = 0;
  virtual void getMacro(double* ux, double* uy, double* uz, double* rho) const = 0;
  virtual void forces(double* out) const = 0;
  virtual void massStats(double* out) const = 0;
};

template <typename S>
struct Solver : SolverBase {
  std::vector<S> fSrc, fDst, ux, uy, uz, rho;
  const double* W;  // collision / force weights

  Solver(int nx_, int ny_, int nz_) {
    nx = nx_; ny = ny_; nz = nz_;
    N = (long)nx * ny * nz;
    W = (sizeof(S) == 4) ? W32 : W64;
    fSrc.assign(19 * N, 0); fDst.assign(19 * N, 0);
    ux.assign(N, 0); uy.assign(N, 0); uz.assign(N, 0); rho.assign(N, 0);
    flags.assign(N, 0); bmap.assign(N, -1);
  }

  // lbmCpuReference.ts:114-143, plus optional seeded noise on non-solid cells.
  void reset(uint64_t seed, double noiseAmp) override {
    stepIndex = 0;
    anchorCount = 0;
    for (long idx = 0; idx < N; ++idx) {
      const int iy = (int)((idx / nx) % ny);
      double ux0 = p.uInlet;
      if (p.groundNoSlip || p.profile == 2) {
        ux0 = inlet_profile(p.profile == 2 ? 2 : 1, p.uInlet, iy, ny, p.z0, p.zref);
        if (iy == 0) ux0 = 0;
      }
      double uy0 = 0.0, uz0 = 0.0;
      if (flags[idx] == CELL_SOLID) {
        ux0 = 0.0;
      } else if (noiseAmp > 0.0) {
        ux0 = ux0 + noiseAmp * noise_unit(seed, (uint64_t)idx, 0);
        uy0 = noiseAmp * noise_unit(seed, (uint64_t)idx, 1);
        uz0 = noiseAmp * noise_unit(seed, (uint64_t)idx, 2);
      }
      ux[idx] = (S)ux0; uy[idx] = (S)uy0; uz[idx] = (S)uz0; rho[idx] = (S)1.0;
      for (int d = 0; d < 19; ++d) {
        const double e = feq64(d, 1.0, ux0, uy0, uz0);
        fSrc[d * N + idx] = (S)e;
        fDst[d * N + idx] = (S)e;
      }
    }
  }

  void step(int n) override {
    for (int s = 0; s < n; ++s) {
      collideAndStream();
      applyBoundaries();
      std::swap(fSrc, fDst);
      ++stepIndex;
    }
  }

  // lbmCpuReference.ts:158-313.
  void collideAndStream() {
    const double tau0 = 3.0 * p.nu + 0.5;
    const double csSq = p.cs * p.cs;
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < nz; ++iz) {
      S fLocal[19];
      S feqS[19];
      for (int iy = 0; iy < ny; ++iy) {
        const long rowOffset = (long)nx * (iy + (long)ny * iz);
        for (int ix = 0; ix < nx; ++ix) {
          const long idx = rowOffset + ix;
          const uint8_t flag = flags[idx];
          if (flag == CELL_SOLID) {
            ux[idx] = 0; uy[idx] = 0; uz[idx] = 0; rho[idx] = (S)1.0;
            continue;
          }
          double locRho = 0.0, locUx = 0.0, locUy = 0.0, locUz = 0.0;
          const int bIdx = flag == CELL_BOUNDARY ? bmap[idx] : -1;
          const long qBase = bIdx >= 0 ? (long)bIdx * 18 : -1;
          for (int d = 0; d < 19; ++d) {
            const int sx = clampi(ix - CX[d], 0, nx - 1);
            const int sy = clampi(iy - CY[d], 0, ny - 1);
            const int sz = clampi(iz - CZ[d], 0, nz - 1);
            const long srcIdx = sx + (long)nx * (sy + (long)ny * sz);
            double val;
            if (d > 0 && flags[srcIdx] == CELL_SOLID) {
              const int wallDir = OPP[d];
              const double qv = qBase >= 0 ? (double)q[qBase + (wallDir - 1)] : 0.5;
              const double qEff = qv > 0.0 ? qv : 0.5;
              const double fOutCurr = (double)fSrc[wallDir * N + idx];
              const int ffx = clampi(ix + CX[d], 0, nx - 1);
              const int ffy = clampi(iy + CY[d], 0, ny - 1);
              const int ffz = clampi(iz + CZ[d], 0, nz - 1);
              const long ffIdx = ffx + (long)nx * (ffy + (long)ny * ffz);
              if (flags[ffIdx] != CELL_SOLID) {
                if (qEff < 0.5) {
                  const double fOutNeighbor = (double)fSrc[wallDir * N + ffIdx];
                  val = 2.0 * qEff * fOutCurr + (1.0 - 2.0 * qEff) * fOutNeighbor;
                } else {
                  const double fInCurr = (double)fSrc[d * N + idx];
                  const double inv2q = 1.0 / (2.0 * qEff);
                  val = inv2q * fOutCurr + (1.0 - inv2q) * fInCurr;
                }
              } else {
                val = fOutCurr;
              }
            } else {
              val = (double)fSrc[d * N + srcIdx];
            }
            fLocal[d] = (S)val;
            locRho += val;
            locUx += CX[d] * val;
            locUy += CY[d] * val;
            locUz += CZ[d] * val;
          }
          locRho = std::max(0.5, std::min(1.8, locRho));
          const double invRho = 1.0 / locRho;
          locUx *= invRho; locUy *= invRho; locUz *= invRho;
          const double uMagSq = locUx * locUx + locUy * locUy + locUz * locUz;
          const double maxUMag = 0.22;
          if (uMagSq > maxUMag * maxUMag) {
            const double scale = maxUMag / std::sqrt(uMagSq);
            locUx *= scale; locUy *= scale; locUz *= scale;
          }
          const double uSq = locUx * locUx + locUy * locUy + locUz * locUz;
          double piXX = 0, piYY = 0, piZZ = 0, piXY = 0, piXZ = 0, piYZ = 0;
          for (int d = 0; d < 19; ++d) {
            const int cx = CX[d], cy = CY[d], cz = CZ[d];
            const double cu = 3.0 * (cx * locUx + cy * locUy + cz * locUz);
            const double eq = W[d] * locRho * (1.0 + cu + 0.5 * cu * cu - 1.5 * uSq);
            feqS[d] = (S)eq;
            const double fNeq = (double)fLocal[d] - eq;
            piXX += cx * cx * fNeq;
            piYY += cy * cy * fNeq;
            piZZ += cz * cz * fNeq;
            piXY += cx * cy * fNeq;
            piXZ += cx * cz * fNeq;
            piYZ += cy * cz * fNeq;
          }
          const double piNorm = std::sqrt(piXX * piXX + piYY * piYY + piZZ * piZZ +
                                          2.0 * (piXY * piXY + piXZ * piXZ + piYZ * piYZ));
          double tauEff =
              0.5 * (tau0 + std::sqrt(tau0 * tau0 + (18.0 * M_SQRT2 * csSq * piNorm) / locRho));
          if (p.closure) {
            const double P[6] = {piXX / locRho, piYY / locRho, piZZ / locRho, piXY / locRho, piXZ / locRho,
                                 piYZ / locRho};
            double x[MLP_IN];
            const double smag = 18.0 * M_SQRT2 * csSq;
            closure_features(P, p.nu, tau0, smag, x);
            const double pn = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2] +
                                        2.0 * (P[3] * P[3] + P[4] * P[4] + P[5] * P[5]));
            const double tauS = 0.5 * (tau0 + std::sqrt(tau0 * tau0 + smag * pn));
            tauEff = tau0 + (tauS - tau0) * softplus(closure_forward(p.mlp, x) + std::log(std::exp(1.0) - 1.0));
          }
          const double omega = 1.0 / std::max(0.5015, std::min(1.95, tauEff));
          for (int d = 0; d < 19; ++d) {
            const int cx = CX[d], cy = CY[d], cz = CZ[d];
            const double qXX = cx * cx - 1.0 / 3.0;
            const double qYY = cy * cy - 1.0 / 3.0;
            const double qZZ = cz * cz - 1.0 / 3.0;
            const double qXY = cx * cy, qXZ = cx * cz, qYZ = cy * cz;
            const double f1Reg = 4.5 * W[d] *
                                 (qXX * piXX + qYY * piYY + qZZ * piZZ +
                                  2.0 * (qXY * piXY + qXZ * piXZ + qYZ * piYZ));
            fDst[d * N + idx] = (S)((double)feqS[d] + (1.0 - omega) * f1Reg);
          }
          ux[idx] = (S)locUx; uy[idx] = (S)locUy; uz[idx] = (S)locUz; rho[idx] = (S)locRho;
        }
      }
    }
  }

  // lbmCpuReference.ts:315-456.
  void applyBoundaries() {
    // 1. inlet (ix = 0) then outlet (ix = nx - 1), row by row.
    for (int iz = 0; iz < nz; ++iz) {
      for (int iy = 0; iy < ny; ++iy) {
        const long rowOffset = (long)nx * (iy + (long)ny * iz);
        const long inIdx = rowOffset, nextIdx = rowOffset + 1;
        if (flags[inIdx] != CELL_SOLID) {
          double uxIn = p.uInlet;
          if (p.groundNoSlip || p.profile == 2)
            uxIn = inlet_profile(p.profile == 2 ? 2 : 1, p.uInlet, iy, ny, p.z0, p.zref);
          double uyIn = 0.0;
          if (stepIndex < 120)
            uyIn = 0.015 * p.uInlet * std::sin(stepIndex * 0.15 + ((double)iy / ny) * M_PI * 2.0);
          const double rhoIn = 0.85 * (double)rho[nextIdx] + 0.15 * 1.0;
          for (int d = 0; d < 19; ++d) fDst[d * N + inIdx] = (S)feq64(d, rhoIn, uxIn, uyIn, 0.0);
          ux[inIdx] = (S)uxIn; uy[inIdx] = (S)uyIn; uz[inIdx] = 0; rho[inIdx] = (S)rhoIn;
        }
        const long outIdx = rowOffset + (nx - 1), prevIdx = outIdx - 1;
        if (flags[outIdx] != CELL_SOLID) {
          const double rhoOut = (double)rho[prevIdx] * 0.85 + 1.0 * 0.15;
          const double uxOut = ux[prevIdx], uyOut = uy[prevIdx], uzOut = uz[prevIdx];
          for (int d = 0; d < 19; ++d) {
            const double fPrev = (double)fDst[d * N + prevIdx];
            const double feqOut = feq64(d, rhoOut, uxOut, uyOut, uzOut);
            fDst[d * N + outIdx] = (S)(0.75 * fPrev + 0.25 * feqOut);
          }
          ux[outIdx] = (S)uxOut; uy[outIdx] = (S)uyOut; uz[outIdx] = (S)uzOut; rho[outIdx] = (S)rhoOut;
        }
      }
    }
    // 2. floor (iy = 0) and ceiling (iy = ny - 1) for ix in [1, nx - 2].
    for (int iz = 0; iz < nz; ++iz) {
      for (int ix = 1; ix < nx - 1; ++ix) {
        const long botIdx = ix + (long)nx * (0 + (long)ny * iz);
        if (flags[botIdx] != CELL_SOLID && flags[botIdx] != CELL_BOUNDARY) {
          const double uxWall = p.groundNoSlip ? 0.0 : (double)ux[botIdx];
          const double uzWall = p.groundNoSlip ? 0.0 : (double)uz[botIdx];
          const double rhoW = rho[botIdx];
          for (int d = 0; d < 19; ++d) fDst[d * N + botIdx] = (S)feq64(d, rhoW, uxWall, 0.0, uzWall);
          ux[botIdx] = (S)uxWall; uy[botIdx] = 0; uz[botIdx] = (S)uzWall;
        }
        const long topIdx = ix + (long)nx * (ny - 1 + (long)ny * iz);
        if (flags[topIdx] != CELL_SOLID && flags[topIdx] != CELL_BOUNDARY) {
          const double rhoW = rho[topIdx];
          for (int d = 0; d < 19; ++d)
            fDst[d * N + topIdx] = (S)feq64(d, rhoW, (double)ux[topIdx], 0.0, (double)uz[topIdx]);
          uy[topIdx] = 0;
        }
      }
    }
    // 3. side walls (iz = 0, nz - 1) when nz > 4.
    if (nz > 4) {
      for (int iy = 0; iy < ny; ++iy) {
        for (int ix = 1; ix < nx - 1; ++ix) {
          const long s0 = ix + (long)nx * (iy + (long)ny * 0);
          if (flags[s0] != CELL_SOLID && flags[s0] != CELL_BOUNDARY) {
            const double rhoW = rho[s0];
            for (int d = 0; d < 19; ++d)
              fDst[d * N + s0] = (S)feq64(d, rhoW, (double)ux[s0], (double)uy[s0], 0.0);
            uz[s0] = 0;
          }
          const long s1 = ix + (long)nx * (iy + (long)ny * (nz - 1));
          if (flags[s1] != CELL_SOLID && flags[s1] != CELL_BOUNDARY) {
            const double rhoW = rho[s1];
            for (int d = 0; d < 19; ++d)
              fDst[d * N + s1] = (S)feq64(d, rhoW, (double)ux[s1], (double)uy[s1], 0.0);
            uz[s1] = 0;
          }
        }
      }
    }
    // 4. global mass anchor (lbmCpuReference.ts:431-455).
    double sumRho = 0.0;
    long fluidCnt = 0;
    for (long idx = 0; idx < N; ++idx) {
      if (flags[idx] != CELL_SOLID) {
        sumRho += (double)rho[idx];
        ++fluidCnt;
      }
    }
    if (fluidCnt > 0 && anchorOn) {
      const double meanRho = sumRho / fluidCnt;
      const double corr = 1.0 / meanRho;
      if (std::fabs(meanRho - 1.0) > 1e-4) {
        ++anchorCount;
        for (long idx = 0; idx < N; ++idx) {
          if (flags[idx] != CELL_SOLID) {
            rho[idx] = (S)((double)rho[idx] * corr);
            for (int d = 0; d < 19; ++d) fDst[d * N + idx] = (S)((double)fDst[d * N + idx] * corr);
          }
        }
      }
    }
  }

  void getF(double* out) const override {
    for (long i = 0; i < 19 * N; ++i) out[i] = (double)fSrc[i];
  }
  void setF(const double* in) override {
    for (long i = 0; i < 19 * N; ++i) fSrc[i] = (S)in[i];
  }
  void getMacro(double* oux, double* ouy, double* ouz, double* orho) const override {
    for (long i = 0; i < N; ++i) {
      if (oux) oux[i] = ux[i];
      if (ouy) ouy[i] = uy[i];
      if (ouz) ouz[i] = uz[i];
      if (orho) orho[i] = rho[i];
    }
  }

  // forceIntegrator.ts:106-217 on fPostBoundary = fSrc (after the swap), plus the full torque vector
  // about the domain centre (forces.wgsl:95-102). out: fx fy fz mx my mz cutFaceLinks.
  void forces(double* out) const override {
    double fx = 0, fy = 0, fz = 0, mx = 0, my = 0, mz = 0;
    long cut = 0;
    for (int b = 0; b < nb; ++b) {
      const long cellIdx = blist[b];
      const int ix = (int)(cellIdx % nx);
      const long rem = (cellIdx - ix) / nx;
      const int iy = (int)(rem % ny);
      const int iz = (int)((rem - iy) / ny);
      double cfx = 0, cfy = 0, cfz = 0;
      const long qBase = (long)b * 18;
      for (int d = 1; d <= 18; ++d) {
        const double qv = q[qBase + (d - 1)];
        if (qv > 0.0) {
          const int od = OPP[d];
          const double fOut = (double)fSrc[d * N + cellIdx];
          const int ffx = clampi(ix + CX[od], 0, nx - 1);
          const int ffy = clampi(iy + CY[od], 0, ny - 1);
          const int ffz = clampi(iz + CZ[od], 0, nz - 1);
          const long ffIdx = ffx + (long)nx * (ffy + (long)ny * ffz);
          double fIn = fOut;
          if (flags[ffIdx] != CELL_SOLID) {
            if (qv < 0.5) {
              const double fOutNeighbor = (double)fSrc[d * N + ffIdx];
              fIn = 2.0 * qv * fOut + (1.0 - 2.0 * qv) * fOutNeighbor;
            } else {
              const double fInCurr = (double)fSrc[od * N + cellIdx];
              const double inv2q = 1.0 / (2.0 * qv);
              fIn = inv2q * fOut + (1.0 - inv2q) * fInCurr;
            }
          }
          const double mom = fOut + fIn - 2.0 * W[d];
          cfx += CX[d] * mom;
          cfy += CY[d] * mom;
          cfz += CZ[d] * mom;
          if (d <= 6) ++cut;
        }
      }
      fx += cfx; fy += cfy; fz += cfz;
      const double rx = (ix + 0.5) - 0.5 * nx;
      const double ry = (iy + 0.5) - 0.5 * ny;
      const double rz = (iz + 0.5) - 0.5 * nz;
      mz += rx * cfy - ry * cfx;
      mx += ry * cfz - rz * cfy;
      my += rz * cfx - rx * cfz;
    }
    out[0] = fx; out[1] = fy; out[2] = fz; out[3] = mx; out[4] = my; out[5] = mz; out[6] = (double)cut;
  }

  void massStats(double* out) const override { fluid_mass_stats(flags.data(), rho.data(), ux.data(), uy.data(), uz.data(), N, out); }
};

// ---------------------------------------------------------------------------------------------
// BVH (bvh.ts:12-590): flat nodes, median split on the largest centroid extent, leaf <= 4 triangles.
struct BVH {
  int triCount = 0;
  std::vector<double> tri;     // ax ay az e1x e1y e1z e2x e2y e2z
  std::vector<double> nodeB;   // minX minY minZ maxX maxY maxZ
  std::vector<int> nodeLF, nodeCnt, triIdx;
  double bmin[3], bmax[3];
  int nodePtr = 0;

  // pos: ntri*9 floats (vertex A, B, C). mat: 16 doubles, column-major (THREE.Matrix4.elements) or null.
  BVH(const float* pos, int ntri, const double* mat) {
    triCount = ntri;
    tri.assign((size_t)ntri * 9, 0.0);
    triIdx.resize(ntri);
    std::vector<double> cen((size_t)ntri * 3), bb((size_t)ntri * 6);
    for (int k = 0; k < 3; ++k) { bmin[k] = INFINITY; bmax[k] = -INFINITY; }
    for (int t = 0; t < ntri; ++t) {
      triIdx[t] = t;
      double v[3][3];
      for (int j = 0; j < 3; ++j) {
        double x = pos[t * 9 + j * 3], y = pos[t * 9 + j * 3 + 1], z = pos[t * 9 + j * 3 + 2];
        if (mat) {  // THREE.Vector3.applyMatrix4
          const double* e = mat;
          const double w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
          const double nx_ = (e[0] * x + e[4] * y + e[8] * z + e[12]) * w;
          const double ny_ = (e[1] * x + e[5] * y + e[9] * z + e[13]) * w;
          const double nz_ = (e[2] * x + e[6] * y + e[10] * z + e[14]) * w;
          x = nx_; y = ny_; z = nz_;
        }
        v[j][0] = x; v[j][1] = y; v[j][2] = z;
      }
      double* T = &tri[(size_t)t * 9];
      for (int k = 0; k < 3; ++k) {
        T[k] = v[0][k];
        T[3 + k] = v[1][k] - v[0][k];
        T[6 + k] = v[2][k] - v[0][k];
        cen[t * 3 + k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
        const double mn = std::min(v[0][k], std::min(v[1][k], v[2][k]));
        const double mxv = std::max(v[0][k], std::max(v[1][k], v[2][k]));
        bb[t * 6 + k] = mn;
        bb[t * 6 + 3 + k] = mxv;
        if (mn < bmin[k]) bmin[k] = mn;
        if (mxv > bmax[k]) bmax[k] = mxv;
      }
    }
    const int maxNodes = std::max(1, ntri * 2);
    nodeB.assign((size_t)maxNodes * 6, 0.0);
    nodeLF.assign(maxNodes, 0);
    nodeCnt.assign(maxNodes, 0);
    if (ntri > 0) {
      nodePtr = 1;
      build(0, 0, ntri, cen, bb);
    }
  }

  void build(int node, int first, int count, const std::vector<double>& cen, const std::vector<double>& bb) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mxv[3] = {-INFINITY, -INFINITY, -INFINITY};
    double cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = first; i < first + count; ++i) {
      const int t = triIdx[i];
      for (int k = 0; k < 3; ++k) {
        if (bb[t * 6 + k] < mn[k]) mn[k] = bb[t * 6 + k];
        if (bb[t * 6 + 3 + k] > mxv[k]) mxv[k] = bb[t * 6 + 3 + k];
        const double c = cen[t * 3 + k];
        if (c < cmn[k]) cmn[k] = c;
        if (c > cmx[k]) cmx[k] = c;
      }
    }
    for (int k = 0; k < 3; ++k) { nodeB[node * 6 + k] = mn[k]; nodeB[node * 6 + 3 + k] = mxv[k]; }
    if (count <= 4) { nodeLF[node] = first; nodeCnt[node] = count; return; }
    const double ext[3] = {cmx[0] - cmn[0], cmx[1] - cmn[1], cmx[2] - cmn[2]};
    int axis = 0;
    double split = 0.5 * (cmn[0] + cmx[0]);
    double maxExt = ext[0];
    if (ext[1] > maxExt) { axis = 1; maxExt = ext[1]; split = 0.5 * (cmn[1] + cmx[1]); }
    if (ext[2] > maxExt) { axis = 2; maxExt = ext[2]; split = 0.5 * (cmn[2] + cmx[2]); }
    if (maxExt <= 1e-12) { nodeLF[node] = first; nodeCnt[node] = count; return; }
    int i = first, j = first + count - 1;
    while (i <= j) {
      if (cen[triIdx[i] * 3 + axis] < split) ++i;
      else { std::swap(triIdx[i], triIdx[j]); --j; }
    }
    int leftCount = i - first;
    if (leftCount == 0 || leftCount == count) leftCount = count >> 1;
    const int left = nodePtr;
    nodePtr += 2;
    nodeLF[node] = left;
    nodeCnt[node] = 0;
    build(left, first, leftCount, cen, bb);
    build(left + 1, first + leftCount, count - leftCount, cen, bb);
  }

  // bvh.ts:265-367
  double intersectSegment(double ox, double oy, double oz, double dx, double dy, double dz, double maxT) const {
    if (triCount == 0) return -1;
    const double invDx = std::fabs(dx) > 1e-15 ? 1.0 / dx : (dx >= 0 ? 1e15 : -1e15);
    const double invDy = std::fabs(dy) > 1e-15 ? 1.0 / dy : (dy >= 0 ? 1e15 : -1e15);
    const double invDz = std::fabs(dz) > 1e-15 ? 1.0 / dz : (dz >= 0 ? 1e15 : -1e15);
    double closestT = maxT + 1e-7;
    bool found = false;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int node = stack[--sp];
      const double* B = &nodeB[node * 6];
      const double tx1 = (B[0] - ox) * invDx, tx2 = (B[3] - ox) * invDx;
      double tmin = std::min(tx1, tx2), tmax = std::max(tx1, tx2);
      const double ty1 = (B[1] - oy) * invDy, ty2 = (B[4] - oy) * invDy;
      tmin = std::max(tmin, std::min(ty1, ty2));
      tmax = std::min(tmax, std::max(ty1, ty2));
      const double tz1 = (B[2] - oz) * invDz, tz2 = (B[5] - oz) * invDz;
      tmin = std::max(tmin, std::min(tz1, tz2));
      tmax = std::min(tmax, std::max(tz1, tz2));
      if (tmax < std::max(0.0, tmin) || tmin > closestT) continue;
      const int cnt = nodeCnt[node];
      if (cnt > 0) {
        const int first = nodeLF[node];
        for (int i = first; i < first + cnt; ++i) {
          const double* T = &tri[(size_t)triIdx[i] * 9];
          const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
          const double e2x = T[6], e2y = T[7], e2z = T[8];
          const double px = dy * e2z - dz * e2y;
          const double py = dz * e2x - dx * e2z;
          const double pz = dx * e2y - dy * e2x;
          const double det = e1x * px + e1y * py + e1z * pz;
          if (std::fabs(det) < 1e-13) continue;
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
          if (t > 1e-7 && t <= closestT) { closestT = t; found = true; }
        }
      } else {
        const int left = nodeLF[node];
        stack[sp++] = left + 1;
        stack[sp++] = left;
      }
    }
    return found ? std::min(maxT, closestT) : -1;
  }

  // bvh.ts:375-457
  int collectXScanlineHits(double y, double z, std::vector<double>& hits) const {
    hits.clear();
    if (triCount == 0) return 0;
    const double py = y + 1.37e-8, pz = z + 2.93e-8;
    if (py < bmin[1] || py > bmax[1] || pz < bmin[2] || pz > bmax[2]) return 0;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int node = stack[--sp];
      const double* B = &nodeB[node * 6];
      if (py < B[1] || py > B[4] || pz < B[2] || pz > B[5]) continue;
      const int cnt = nodeCnt[node];
      if (cnt > 0) {
        const int first = nodeLF[node];
        for (int i = first; i < first + cnt; ++i) {
          const double* T = &tri[(size_t)triIdx[i] * 9];
          const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
          const double e2x = T[6], e2y = T[7], e2z = T[8];
          const double det = -e1y * e2z + e1z * e2y;
          if (std::fabs(det) < 1e-13) continue;
          const double invDet = 1.0 / det;
          const double ty = py - ay, tz = pz - az;
          const double u = (-ty * e2z + tz * e2y) * invDet;
          if (u < 0.0 || u > 1.0) continue;
          const double v = (ty * e1z - tz * e1y) * invDet;
          if (v < 0.0 || u + v > 1.0) continue;
          hits.push_back(ax + u * e1x + v * e2x);
        }
      } else {
        const int left = nodeLF[node];
        stack[sp++] = left + 1;
        stack[sp++] = left;
      }
    }
    std::sort(hits.begin(), hits.end());
    return (int)hits.size();
  }

  // bvh.ts:499-589
  int countRayIntersections(double ox, double oy, double oz, double dx, double dy, double dz) const {
    const double invDx = 1.0 / dx, invDy = 1.0 / dy, invDz = 1.0 / dz;
    int hitCount = 0;
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int node = stack[--sp];
      const double* B = &nodeB[node * 6];
      const double tx1 = (B[0] - ox) * invDx, tx2 = (B[3] - ox) * invDx;
      double tmin = std::min(tx1, tx2), tmax = std::max(tx1, tx2);
      const double ty1 = (B[1] - oy) * invDy, ty2 = (B[4] - oy) * invDy;
      tmin = std::max(tmin, std::min(ty1, ty2));
      tmax = std::min(tmax, std::max(ty1, ty2));
      const double tz1 = (B[2] - oz) * invDz, tz2 = (B[5] - oz) * invDz;
      tmin = std::max(tmin, std::min(tz1, tz2));
      tmax = std::min(tmax, std::max(tz1, tz2));
      if (tmax < std::max(0.0, tmin)) continue;
      const int cnt = nodeCnt[node];
      if (cnt > 0) {
        const int first = nodeLF[node];
        for (int i = first; i < first + cnt; ++i) {
          const double* T = &tri[(size_t)triIdx[i] * 9];
          const double ax = T[0], ay = T[1], az = T[2], e1x = T[3], e1y = T[4], e1z = T[5];
          const double e2x = T[6], e2y = T[7], e2z = T[8];
          const double px = dy * e2z - dz * e2y;
          const double py = dz * e2x - dx * e2z;
          const double pz = dx * e2y - dy * e2x;
          const double det = e1x * px + e1y * py + e1z * pz;
          if (std::fabs(det) < 1e-13) continue;
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
        const int left = nodeLF[node];
        stack[sp++] = left + 1;
        stack[sp++] = left;
      }
    }
    return hitCount;
  }

  // bvh.ts:464-497
  bool isPointInside(double x, double y, double z) const {
    if (x < bmin[0] || x > bmax[0] || y < bmin[1] || y > bmax[1] || z < bmin[2] || z > bmax[2]) return false;
    const double dirs[3][3] = {{1.0, 0.0137, 0.0293}, {-0.0191, 1.0, 0.0317}, {0.0239, -0.0173, 1.0}};
    int votes = 0;
    for (int d = 0; d < 3; ++d) {
      const int hits = countRayIntersections(x, y, z, dirs[d][0], dirs[d][1], dirs[d][2]);
      if ((hits & 1) == 1) ++votes;
      if (votes >= 2) return true;
      if (d == 1 && votes == 0) return false;
    }
    return votes >= 2;
  }
};

// voxelizer.ts:86-281
struct Voxels {
  int nx, ny, nz;
  double dx, ox, oy, oz;
  std::vector<uint8_t> flags;
  std::vector<int32_t> bmap, blist;
  std::vector<float> q, minQ;
  int solid = 0, boundary = 0, fluid = 0;
};

Voxels* voxelize(const float* pos, int ntri, const double* mat, double tunnelHeight, int nx, int ny, int nz) {
  Voxels* V = new Voxels();
  V->nx = nx; V->ny = ny; V->nz = nz;
  const long N = (long)nx * ny * nz;
  const double h = tunnelHeight / ny;
  const double dx = h, dy = h, dz = h;
  const double ox = -nx * h * 0.5, oy = -ny * h * 0.5, oz = -nz * h * 0.5;
  V->dx = h; V->ox = ox; V->oy = oy; V->oz = oz;
  BVH bvh(pos, ntri, mat);
  V->flags.assign(N, CELL_FLUID);
  uint8_t* flags = V->flags.data();
  const int minIx = std::max(0, (int)std::floor((bvh.bmin[0] - ox) / dx) - 1);
  const int maxIx = std::min(nx - 1, (int)std::ceil((bvh.bmax[0] - ox) / dx) + 1);
  const int minIy = std::max(0, (int)std::floor((bvh.bmin[1] - oy) / dy) - 1);
  const int maxIy = std::min(ny - 1, (int)std::ceil((bvh.bmax[1] - oy) / dy) + 1);
  const int minIz = std::max(0, (int)std::floor((bvh.bmin[2] - oz) / dz) - 1);
  const int maxIz = std::min(nz - 1, (int)std::ceil((bvh.bmax[2] - oz) / dz) + 1);
  int solid = 0;
  std::vector<double> xh;
  for (int iz = minIz; iz <= maxIz; ++iz) {
    const double z = oz + (iz + 0.5) * dz;
    for (int iy = minIy; iy <= maxIy; ++iy) {
      const double y = oy + (iy + 0.5) * dy;
      const int hc = bvh.collectXScanlineHits(y, z, xh);
      if (hc == 0) continue;
      const long row = (long)nx * (iy + (long)ny * iz);
      if ((hc & 1) == 0) {
        for (int k = 0; k < hc; k += 2) {
          const double xe = xh[k], xx = xh[k + 1];
          if (xx - xe < 1e-7) continue;
          const int s = std::max(minIx, (int)std::ceil((xe - ox) / dx - 0.5));
          const int e = std::min(maxIx, (int)std::floor((xx - ox) / dx - 0.5));
          for (int ix = s; ix <= e; ++ix) {
            if (flags[row + ix] == CELL_FLUID) { flags[row + ix] = CELL_SOLID; ++solid; }
          }
        }
      } else {
        for (int ix = minIx; ix <= maxIx; ++ix) {
          const double x = ox + (ix + 0.5) * dx;
          if (x >= xh[0] && x <= xh[hc - 1]) {
            if (bvh.isPointInside(x, y, z)) {
              if (flags[row + ix] == CELL_FLUID) { flags[row + ix] = CELL_SOLID; ++solid; }
            }
          }
        }
      }
    }
  }
  const int sMinIx = std::max(0, minIx - 1), sMaxIx = std::min(nx - 1, maxIx + 1);
  const int sMinIy = std::max(0, minIy - 1), sMaxIy = std::min(ny - 1, maxIy + 1);
  const int sMinIz = std::max(0, minIz - 1), sMaxIz = std::min(nz - 1, maxIz + 1);
  std::vector<int32_t>& bl = V->blist;
  for (int iz = sMinIz; iz <= sMaxIz; ++iz)
    for (int iy = sMinIy; iy <= sMaxIy; ++iy) {
      const long row = (long)nx * (iy + (long)ny * iz);
      for (int ix = sMinIx; ix <= sMaxIx; ++ix) {
        const long idx = row + ix;
        if (flags[idx] != CELL_FLUID) continue;
        bool has = false;
        for (int d = 1; d <= 18; ++d) {
          const int a = ix + CX[d], b = iy + CY[d], c = iz + CZ[d];
          if (a < 0 || a >= nx || b < 0 || b >= ny || c < 0 || c >= nz) continue;
          if (flags[a + (long)nx * (b + (long)ny * c)] == CELL_SOLID) { has = true; break; }
        }
        if (has) { flags[idx] = CELL_BOUNDARY; bl.push_back((int32_t)idx); }
      }
    }
  const int nb = (int)bl.size();
  V->bmap.assign(N, -1);
  V->q.assign((size_t)nb * 18, -1.0f);
  V->minQ.assign(nb, 1.0f);
  for (int b = 0; b < nb; ++b) {
    const long cellIdx = bl[b];
    V->bmap[cellIdx] = b;
    const int ix = (int)(cellIdx % nx);
    const long rem = (cellIdx - ix) / nx;
    const int iy = (int)(rem % ny);
    const int iz = (int)((rem - iy) / ny);
    const double fx = ox + (ix + 0.5) * dx, fy = oy + (iy + 0.5) * dy, fz = oz + (iz + 0.5) * dz;
    double minQ = 1.0;
    for (int d = 1; d <= 18; ++d) {
      const int a = ix + CX[d], bb = iy + CY[d], c = iz + CZ[d];
      if (a < 0 || a >= nx || bb < 0 || bb >= ny || c < 0 || c >= nz) continue;
      if (flags[a + (long)nx * (bb + (long)ny * c)] == CELL_SOLID) {
        // computeSubGridLinkFraction (voxelizer.ts:65-80)
        const double t = bvh.intersectSegment(fx, fy, fz, CX[d] * dx, CY[d] * dy, CZ[d] * dz, 1.0);
        double qv = 0.5;
        if (t > 0 && t <= 1.0) qv = std::max(1e-4, std::min(1.0, t));
        V->q[(size_t)b * 18 + (d - 1)] = (float)qv;
        if (qv < minQ) minQ = qv;
      }
    }
    V->minQ[b] = (float)minQ;
  }
  V->solid = solid;
  V->boundary = nb;
  V->fluid = (int)(N - solid - nb);
  return V;
}

// frontalArea.ts:22-179
double frontalArea(const float* pos, int ntri, const double* mat, int gridRes) {
  if (ntri <= 0) return 0;
  const int nv = ntri * 3;
  std::vector<double> pY(nv), pZ(nv);
  double minY = INFINITY, maxY = -INFINITY, minZ = INFINITY, maxZ = -INFINITY;
  for (int i = 0; i < nv; ++i) {
    double x = pos[i * 3], y = pos[i * 3 + 1], z = pos[i * 3 + 2];
    if (mat) {
      const double* e = mat;
      const double w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15]);
      const double ny_ = (e[1] * x + e[5] * y + e[9] * z + e[13]) * w;
      const double nz_ = (e[2] * x + e[6] * y + e[10] * z + e[14]) * w;
      y = ny_; z = nz_;
    }
    pY[i] = y; pZ[i] = z;
    if (y < minY) minY = y;
    if (y > maxY) maxY = y;
    if (z < minZ) minZ = z;
    if (z > maxZ) maxZ = z;
  }
  const double spanY = maxY - minY, spanZ = maxZ - minZ;
  if (spanY <= 1e-9 || spanZ <= 1e-9) return 0;
  const int ny = gridRes, nz = gridRes;
  std::vector<uint16_t> grid((size_t)ny * nz, 0);
  const double invCellY = ny / spanY, invCellZ = nz / spanZ;
  const double sub[4] = {0.125, 0.375, 0.625, 0.875};
  for (int t = 0; t < ntri; ++t) {
    const int i0 = t * 3, i1 = t * 3 + 1, i2 = t * 3 + 2;
    const double z0 = (pZ[i0] - minZ) * invCellZ, y0 = (pY[i0] - minY) * invCellY;
    const double z1 = (pZ[i1] - minZ) * invCellZ, y1 = (pY[i1] - minY) * invCellY;
    const double z2 = (pZ[i2] - minZ) * invCellZ, y2 = (pY[i2] - minY) * invCellY;
    const double area2 = (z1 - z0) * (y2 - y0) - (y1 - y0) * (z2 - z0);
    if (std::fabs(area2) < 1e-11) continue;
    double az = z0, ay = y0, bz = z1, by = y1, cz = z2, cy = y2;
    if (area2 < 0) { bz = z2; by = y2; cz = z1; cy = y1; }
    const int tMinZ = std::max(0, (int)std::floor(std::min(az, std::min(bz, cz))));
    const int tMaxZ = std::min(nz - 1, (int)std::floor(std::max(az, std::max(bz, cz))));
    const int tMinY = std::max(0, (int)std::floor(std::min(ay, std::min(by, cy))));
    const int tMaxY = std::min(ny - 1, (int)std::floor(std::max(ay, std::max(by, cy))));
    const double dz01 = bz - az, dy01 = by - ay, dz12 = cz - bz, dy12 = cy - by, dz20 = az - cz, dy20 = ay - cy;
    for (int iy = tMinY; iy <= tMaxY; ++iy) {
      for (int iz = tMinZ; iz <= tMaxZ; ++iz) {
        const size_t c = (size_t)iy * nz + iz;
        unsigned mask = grid[c];
        if (mask == 0xffff) continue;
        unsigned bit = 1;
        for (int sy = 0; sy < 4; ++sy) {
          const double py = iy + sub[sy];
          for (int sz = 0; sz < 4; ++sz) {
            if ((mask & bit) == 0) {
              const double pz = iz + sub[sz];
              const double e0 = dz01 * (py - ay) - dy01 * (pz - az);
              if (e0 >= -1e-9) {
                const double e1 = dz12 * (py - by) - dy12 * (pz - bz);
                if (e1 >= -1e-9) {
                  const double e2 = dz20 * (py - cy) - dy20 * (pz - cz);
                  if (e2 >= -1e-9) mask |= bit;
                }
              }
            }
            bit <<= 1;
          }
        }
        grid[c] = (uint16_t)(mask & 0xffff);
      }
    }
  }
  long covered = 0;
  for (size_t i = 0; i < grid.size(); ++i) covered += __builtin_popcount((unsigned)grid[i]);
  const double total = (double)ny * nz * 16;
  return ((double)covered / total) * (spanY * spanZ);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C API for the Python test harness (ctypes).
extern "C" {

void* orc_new(int nx, int ny, int nz, int ref32) {
  if (ref32) return static_cast<SolverBase*>(new Solver<float>(nx, ny, nz));
  return static_cast<SolverBase*>(new Solver<double>(nx, ny, nz));
}
void orc_free(void* h) { delete static_cast<SolverBase*>(h); }
// Threads of the parallel collide-and-stream loop (the CPU baseline times 1 thread and all of them).
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads() { return omp_get_max_threads(); }

void orc_set_geometry(void* h, const uint8_t* flags, const int32_t* bmap, const int32_t* blist, const float* q, int nb) {
  SolverBase* s = static_cast<SolverBase*>(h);
  std::memcpy(s->flags.data(), flags, (size_t)s->N);
  std::memcpy(s->bmap.data(), bmap, (size_t)s->N * 4);
  s->blist.assign(blist, blist + nb);
  s->q.assign(q, q + (size_t)nb * 18);
  s->nb = nb;
}

void orc_set_params(void* h, double uInlet, double nu, double cs, int groundNoSlip, int profile, double z0, double zref) {
  SolverBase* s = static_cast<SolverBase*>(h);
  s->p.uInlet = uInlet; s->p.nu = nu; s->p.cs = cs; s->p.groundNoSlip = groundNoSlip;
  s->p.profile = profile; s->p.z0 = z0; s->p.zref = zref;
}
void orc_set_uinlet(void* h, double u) { static_cast<SolverBase*>(h)->p.uInlet = u; }
// Learned closure parameters (1313 floats, the C ABI's order), or NULL for Smagorinsky.
void orc_set_closure(void* h, const float* prm) {
  Params& p = static_cast<SolverBase*>(h)->p;
  p.closure = prm != nullptr;
  for (int i = 0; i < MLP_NPARAM; ++i) p.mlp[i] = prm ? (double)prm[i] : 0.0;
}
void orc_closure_features(const double* P, double nu, double tau0, double smag, double* x) {
  closure_features(P, nu, tau0, smag, x);
}
// Network outputs of n feature rows (6 each) for parameters given as floats.
void orc_closure_forward(const float* prm, const double* X, long n, double* y) {
  double p[MLP_NPARAM];
  for (int i = 0; i < MLP_NPARAM; ++i) p[i] = prm[i];
  for (long s = 0; s < n; ++s) y[s] = closure_forward(p, X + 6 * s);
}
void orc_reset(void* h, uint64_t seed, double noiseAmp) { static_cast<SolverBase*>(h)->reset(seed, noiseAmp); }
void orc_step(void* h, int n) { static_cast<SolverBase*>(h)->step(n); }
long orc_step_index(void* h) { return static_cast<SolverBase*>(h)->stepIndex; }
int orc_anchor_count(void* h) { return static_cast<SolverBase*>(h)->anchorCount; }
void orc_set_anchor(void* h, int on) { static_cast<SolverBase*>(h)->anchorOn = on != 0; }
void orc_get_f(void* h, double* out) { static_cast<SolverBase*>(h)->getF(out); }
void orc_set_f(void* h, const double* in) { static_cast<SolverBase*>(h)->setF(in); }
void orc_get_macro(void* h, double* ux, double* uy, double* uz, double* rho) {
  static_cast<SolverBase*>(h)->getMacro(ux, uy, uz, rho);
}
void orc_forces(void* h, double* out) { static_cast<SolverBase*>(h)->forces(out); }
void orc_mass_stats(void* h, double* out) { static_cast<SolverBase*>(h)->massStats(out); }
// The same statistics over caller-given float fields (the GPU's readout fields, for one-to-one checks).
void orc_mass_stats_fields(const uint8_t* flags, const float* rho, const float* ux, const float* uy, const float* uz,
                           long n, double* out) {
  fluid_mass_stats(flags, rho, ux, uy, uz, n, out);
}

// Voxelizer handle.
void* orc_vox_new(const float* pos, int ntri, const double* mat, double tunnelHeight, int nx, int ny, int nz) {
  return voxelize(pos, ntri, mat, tunnelHeight, nx, ny, nz);
}
void orc_vox_counts(void* h, int* out) {
  Voxels* v = static_cast<Voxels*>(h);
  out[0] = v->solid; out[1] = v->boundary; out[2] = v->fluid;
}
void orc_vox_geom(void* h, double* out) {
  Voxels* v = static_cast<Voxels*>(h);
  out[0] = v->dx; out[1] = v->ox; out[2] = v->oy; out[3] = v->oz;
}
void orc_vox_get(void* h, uint8_t* flags, int32_t* bmap, int32_t* blist, float* q, float* minQ) {
  Voxels* v = static_cast<Voxels*>(h);
  if (flags) std::memcpy(flags, v->flags.data(), v->flags.size());
  if (bmap) std::memcpy(bmap, v->bmap.data(), v->bmap.size() * 4);
  if (blist) std::memcpy(blist, v->blist.data(), v->blist.size() * 4);
  if (q) std::memcpy(q, v->q.data(), v->q.size() * 4);
  if (minQ) std::memcpy(minQ, v->minQ.data(), v->minQ.size() * 4);
}
void orc_vox_free(void* h) { delete static_cast<Voxels*>(h); }

// BVH queries (for the point-in-mesh and segment tests).
void* orc_bvh_new(const float* pos, int ntri, const double* mat) { return new BVH(pos, ntri, mat); }
void orc_bvh_free(void* h) { delete static_cast<BVH*>(h); }
int orc_bvh_inside(void* h, double x, double y, double z) { return static_cast<BVH*>(h)->isPointInside(x, y, z) ? 1 : 0; }
double orc_bvh_segment(void* h, double ox, double oy, double oz, double dx, double dy, double dz, double maxT) {
  return static_cast<BVH*>(h)->intersectSegment(ox, oy, oz, dx, dy, dz, maxT);
}
int orc_bvh_scanline(void* h, double y, double z, double* out, int cap) {
  std::vector<double> hits;
  const int n = static_cast<BVH*>(h)->collectXScanlineHits(y, z, hits);
  for (int i = 0; i < n && i < cap; ++i) out[i] = hits[i];
  return n;
}

double orc_frontal_area(const float* pos, int ntri, const double* mat, int gridRes) {
  return frontalArea(pos, ntri, mat, gridRes);
}

// Equilibrium with double weights and the noise hash, exposed for the table/equilibrium tests.
double orc_feq64(int d, double rho, double ux, double uy, double uz) { return feq64(d, rho, ux, uy, uz); }
double orc_noise_unit(uint64_t seed, uint64_t idx, int comp) { return noise_unit(seed, idx, comp); }

}  // extern "C"
