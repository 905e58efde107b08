// Learned eddy-viscosity closure: a 6 -> 32 -> 32 -> 1 MLP over invariants of the non-equilibrium stress,
// evaluated one wavefront (64 cells or samples) at a time on the matrix cores with
// v_mfma_f32_32x32x8bf16_1k (BF16 inputs, FP32 accumulation).
//
// Transposed formulation: a layer computes H^T (32 units x 32 samples) = W (32 x K) . X^T (K x 32 samples)
// for each half of the wave's 64 samples. MFMA operand layouts for 32x32x8 (lane l, j = 0..3):
//   A (32 x 8):  row l % 32, k = 4 (l / 32) + j
//   B (8 x 32):  col l % 32, k = 4 (l / 32) + j
//   D (32 x 32): col l % 32, rows 8 (r / 4) + 4 (l / 32) + r % 4 for r = 0..15
// so the D registers r = 4c..4c+3 of one layer are exactly the B operand of k-chunk c of the next layer.
// Lane l owns sample l; samples 0..31 form group 0 and 32..63 group 1, and lane l finds its own result in
// group l / 32 at register position l % 32 of its half.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace aero {

constexpr int MLP_IN = 6, MLP_INP = 8, MLP_H = 32;
constexpr int MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1;  // 1313

// FP32 master parameters in this order (the C ABI exchanges them as 1313 floats):
//   W1[32][6], b1[32], W2[32][32], b2[32], w3[32], b3
constexpr int P_W1 = 0, P_B1 = P_W1 + MLP_H * MLP_IN, P_W2 = P_B1 + MLP_H, P_B2 = P_W2 + MLP_H * MLP_H,
              P_W3 = P_B2 + MLP_H, P_B3 = P_W3 + MLP_H;

// Device copy used by the kernels: BF16 weights in MFMA operand order plus FP32 biases. The first layer's bias
// rides in the product itself: its two padding columns hold W1[i][6] = bf16(b1[i]) and
// W1[i][7] = bf16(b1[i] - bf16(b1[i])) against inputs of 1, so the bias enters with 16 significant bits (relative
// error at most 2^-17) and costs no instruction per unit.
struct MlpDev {
  uint16_t w1[MLP_H * MLP_INP];   // W1[i][k] for k < 6, then the bias pair
  uint16_t w2[MLP_H * MLP_H];     // W2[i][k]
  uint16_t w2t[MLP_H * MLP_H];    // W2^T[i][k] = W2[k][i] (backward)
  float b2[MLP_H], w3[MLP_H];
  float b3;
  float pad[3];
};

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float x) {
  uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
  u = __float_as_uint(x);
#else
  memcpy(&u, &x, 4);
#endif
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  memcpy(&f, &u, 4);
  return f;
#endif
}

// The closure's transcendental functions use the hardware instructions (v_log_f32, v_exp_f32, v_rcp_f32,
// v_sqrt_f32: about 1 ulp of FP32). Their results feed BF16 operands (8 significant bits) or a relaxation time
// that the network itself only sets to a few digits, so the library versions' last bits buy nothing there.
__device__ __forceinline__ float fast_ln(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float softplusf(float y) { return y > 20.0f ? y : fast_ln(1.0f + fast_exp(y)); }

// Output mapping of the network: the eddy viscosity is a factor g = softplus(y + MLP_C0) times the Smagorinsky
// eddy viscosity nu_S for the same stress, so y = 0 (the zero network) is the Smagorinsky closure.
constexpr float MLP_C0 = 0.54132485461291806f;  // ln(e - 1): softplus(MLP_C0) = 1
__device__ __forceinline__ float mlp_factor(float y) { return softplusf(y + MLP_C0); }

// Smagorinsky relaxation time for a non-equilibrium stress of norm pn per unit density
// (lbmCpuReference.ts:280-281): tau_S = (tau0 + sqrt(tau0^2 + smag pn)) / 2; nu_S = (tau_S - tau0) / 3.
__device__ __forceinline__ float smag_tau(float tau0, float smag, float pn) {
  return 0.5f * (tau0 + __builtin_amdgcn_sqrtf(tau0 * tau0 + smag * pn));
}

// Features of the non-equilibrium stress per unit density P (xx, yy, zz, xy, xz, yz), the molecular
// viscosity nu and the Smagorinsky factor smag = 18 sqrt(2) Cs^2 (for the baseline-closure feature).
//   x0 = ln(|P| + 1e-12)/10 + 1      magnitude
//   x1 = tr P / |P|                   isotropic share
//   x2 = |dev P|^2 / |P|^2            deviatoric share
//   x3 = sqrt(6) tr(dev P^3) / |dev P|^3   shape (Lode-type invariant), in [-1, 1]
//   x4 = ln(nu)/10 + 1                molecular viscosity
//   x5 = ln(1 + nu_smag / nu) / 4     the Smagorinsky eddy viscosity for the same stress
// Returns |P|, which the collision's relaxation time also needs. x[6] and x[7] are not features: the forward pass
// puts the first layer's bias inputs there.
__device__ __forceinline__ float mlp_features(float pxx, float pyy, float pzz, float pxy, float pxz, float pyz,
                                              float nu, float tau0, float smag, float* x) {
  const float n2 = pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz);
  const float pn = __builtin_amdgcn_sqrtf(n2);
  const float tr = pxx + pyy + pzz;
  const float m = tr * (1.0f / 3.0f);
  const float dxx = pxx - m, dyy = pyy - m, dzz = pzz - m;
  const float d2 = dxx * dxx + dyy * dyy + dzz * dzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz);
  // tr(D^3) for symmetric D = 3 det(D) when traceless
  const float det = dxx * (dyy * dzz - pyz * pyz) - pxy * (pxy * dzz - pyz * pxz) + pxz * (pxy * pyz - dyy * pxz);
  const float d3 = 3.0f * det;
  x[0] = fast_ln(pn + 1e-12f) * 0.1f + 1.0f;
  x[1] = tr * __builtin_amdgcn_rcpf(pn + 1e-20f);
  x[2] = d2 * __builtin_amdgcn_rcpf(n2 + 1e-30f);
  const float l = 2.449489743f * d3 * __builtin_amdgcn_rcpf(d2 * __builtin_amdgcn_sqrtf(d2) + 1e-30f);
  x[3] = fminf(1.0f, fmaxf(-1.0f, l));
  x[4] = fast_ln(nu) * 0.1f + 1.0f;
  const float tauS = smag_tau(tau0, smag, pn);
  const float nuS = fmaxf(0.0f, (tauS - tau0) * (1.0f / 3.0f));
  x[5] = fast_ln(1.0f + nuS * __builtin_amdgcn_rcpf(nu)) * 0.25f;
  x[6] = 0.0f;
  x[7] = 0.0f;
  return pn;
}



// Four floats as BF16 MFMA operands, rounded to nearest even by v_cvt_pk_bf16_f32 (two per instruction; the
// same rounding as f2bf for finite values).
__device__ __forceinline__ bf16x4_t pack4(float a, float b, float c, float d) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  typedef float f32x2_t __attribute__((ext_vector_type(2)));
  const bf16x2_t lo = __builtin_convertvector((f32x2_t){a, b}, bf16x2_t);
  const bf16x2_t hi = __builtin_convertvector((f32x2_t){c, d}, bf16x2_t);
  bf16x4_t v;
  __builtin_memcpy(&v, &lo, 4);
  __builtin_memcpy(reinterpret_cast<char*>(&v) + 4, &hi, 4);
  return v;
}
__device__ __forceinline__ bf16x4_t load4(const uint16_t* p) {
  bf16x4_t v;
  v[0] = (short)p[0];
  v[1] = (short)p[1];
  v[2] = (short)p[2];
  v[3] = (short)p[3];
  return v;
}

// max(z, 0) as a signed-integer max of the bit pattern: a float with the sign bit set is a negative integer.
// One instruction; the float max would first quiet the matrix-core result (which the compiler cannot prove is
// not a signalling NaN) with a second one.
__device__ __forceinline__ float relu(float z) { return __int_as_float(max(__float_as_int(z), 0)); }

// Row of the D register r for lane l.
__device__ __forceinline__ int drow(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// Activations of one wave kept for the backward pass (per group g: 16 values per lane).
struct MlpWave {
  f32x16_t z1[2], z2[2];  // pre-activations
  f32x16_t h1[2], h2[2];  // post-ReLU
  float y;                // this lane's sample output
};

// Forward pass of the 64 samples of a wave; x is this lane's 8 features. Must be called by all 64 lanes.
__device__ __forceinline__ void mlp_forward(const MlpDev& W, const float* x, MlpWave& s) {
  const int l = threadIdx.x & 63;
  const bool lo = l < 32;
  // layer 1 operands: exchange the half of the features that the partner lane needs; inputs 6 and 7 are the
  // constant 1 that multiplies the bias pair
  const bf16x4_t own_a = pack4(x[0], x[1], x[2], x[3]);
  const bf16x4_t own_b = pack4(x[4], x[5], 1.0f, 1.0f);
  const bf16x4_t send = lo ? own_b : own_a;
  int2 sv;
  __builtin_memcpy(&sv, &send, 8);
  int2 rv;
  rv.x = __shfl_xor(sv.x, 32, 64);
  rv.y = __shfl_xor(sv.y, 32, 64);
  bf16x4_t recv;
  __builtin_memcpy(&recv, &rv, 8);
  const bf16x4_t B0 = lo ? own_a : recv;
  const bf16x4_t B1 = lo ? recv : own_b;
  const bf16x4_t A1 = load4(W.w1 + (l & 31) * MLP_INP + 4 * (l >> 5));
  const f32x16_t zero = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  s.z1[0] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1, B0, zero, 0, 0, 0);
  s.z1[1] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1, B1, zero, 0, 0, 0);
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int r = 0; r < 16; ++r) s.h1[g][r] = relu(s.z1[g][r]);
  // layer 2: four k-chunks of 8 hidden units, accumulated onto the bias
  f32x16_t b2v;
#pragma unroll
  for (int r = 0; r < 16; ++r) b2v[r] = W.b2[drow(r, l)];
  f32x16_t acc[2] = {b2v, b2v};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bf16x4_t A2 = load4(W.w2 + (l & 31) * MLP_H + 8 * c + 4 * (l >> 5));
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const bf16x4_t B = pack4(s.h1[g][4 * c], s.h1[g][4 * c + 1], s.h1[g][4 * c + 2], s.h1[g][4 * c + 3]);
      acc[g] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A2, B, acc[g], 0, 0, 0);
    }
  }
  float y[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    float p = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = drow(r, l);
      const float z = acc[g][r];
      s.z2[g][r] = z;
      const float h = relu(z);
      s.h2[g][r] = h;
      p += W.w3[i] * h;
    }
    y[g] = p + __shfl_xor(p, 32, 64) + W.b3;
  }
  s.y = lo ? y[0] : y[1];
}

// Forward only, for the collision: returns this lane's network output y.
__device__ __forceinline__ float mlp_eval(const MlpDev& W, const float* x) {
  MlpWave s;
  mlp_forward(W, x, s);
  return s.y;
}


}  // namespace aero
