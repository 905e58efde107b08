// Training of the learned closure on the GPU: training-sample generation from filtered fine-grid runs,
// forward + backward on the matrix cores (BF16 operands, FP32 accumulation), fixed-order gradient
// reduction, Adam. C ABI: aero_mlp_* (include/aerolbm.h).
//
// Sample generation (k_sgs_samples). For a coarse cell c of the fine grid's 2x2x2 blocks (coarse lattice
// units throughout):
//   u_bar, rho_bar : block means of the fine velocity and density
//   tau_ab         : block mean of u_a u_b minus u_bar_a u_bar_b (resolved sub-filter stress)
//   S_ab           : symmetric central-difference gradient of u_bar over the neighbouring blocks
//   nu_r           : -dev(tau):S / (2 S:S), the least-squares eddy viscosity of the resolved part
//   nu_f           : the fine run's own Smagorinsky viscosity, nu_S(S) / 4 (filter width 1/2: Delta^2 / 4)
//   nu_t*          : nu_r + nu_f, the eddy viscosity that best reproduces the sub-filter stress
//   P*             : -(2/3) (3 (nu + max(0, nu_t*)) + 1/2) S, the non-equilibrium stress per unit density a
//                    coarse run with that eddy viscosity would carry
//   g*             : nu_t* / nu_S(P*), the factor on the Smagorinsky viscosity that the coarse run computes
//                    from P* (the quantity the network predicts), clipped to [-8, 16]
//   w              : (nu_S(P*) |S| / nu)^2 (floored), the weight that turns the squared factor error into the
//                    squared error of the modelled stress 2 nu_S g S: the fit is then a least-squares fit of
//                    the sub-filter stress, which averages the per-cell noise of g* the way the dynamic
//                    procedure does, instead of a fit of noisy ratios
// features(P*) -> input (w in column 7), g* -> target. Only coarse cells whose 7 blocks are all fluid are used.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "mlp.h"

using namespace aero;

namespace {

constexpr int WAVES_PER_BLOCK = 4;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// sum over the 32 lanes of the same half of the wave
__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wave = 64 samples: forward, loss, backward; writes the wave's gradient (master order) and loss.
__global__ __launch_bounds__(64 * WAVES_PER_BLOCK) void k_mlp_grad(const MlpDev* __restrict__ Wp,
                                                                   const float* __restrict__ X,
                                                                   const float* __restrict__ T, int n, float scale,
                                                                   float* __restrict__ wgrad,
                                                                   float* __restrict__ wloss) {
  __shared__ uint16_t s_dz2[WAVES_PER_BLOCK][64][MLP_H];
  __shared__ uint16_t s_h1[WAVES_PER_BLOCK][64][MLP_H];
  __shared__ uint16_t s_dz1[WAVES_PER_BLOCK][64][MLP_H];
  __shared__ uint16_t s_x[WAVES_PER_BLOCK][64][MLP_INP];
  const MlpDev& W = *Wp;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wave = blockIdx.x * WAVES_PER_BLOCK + wv;
  const int p = wave * 64 + l;
  const bool valid = p < n;
  float x[MLP_INP];
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) x[k] = valid ? X[(size_t)p * MLP_INP + k] : 0.0f;
  MlpWave s;
  mlp_forward(W, x, s);
  // loss w (g(y) - t)^2 with the closure factor g(y) = softplus(y + C0), dg/dy = sigmoid(y + C0), and the
  // sample weight w from feature column 7 (which the first layer never reads; 0 stands for weight 1)
  float g = 0.0f, loss = 0.0f;
  if (valid) {
    const float w = x[7] > 0.0f ? x[7] : 1.0f;
    const float d = mlp_factor(s.y) - T[p];
    loss = w * d * d;
    const float sig = 1.0f / (1.0f + expf(-(s.y + MLP_C0)));
    g = 2.0f * w * d * sig * scale;
  }
  const bool lo = l < 32;
  const float gx = __shfl_xor(g, 32, 64);
  const float G[2] = {lo ? g : gx, lo ? gx : g};
  // output layer and dZ2
  f32x16_t dz2[2];
  float dw3[16], db2[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = drow(r, l);
    dw3[r] = 0.0f;
    db2[r] = 0.0f;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      dw3[r] += G[q] * s.h2[q][r];
      const float d = s.z2[q][r] > 0.0f ? W.w3[i] * G[q] : 0.0f;
      dz2[q][r] = d;
      db2[r] += d;
    }
  }
  // dH1 = W2^T dZ2, then dZ1
  const f32x16_t zero = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  f32x16_t dz1[2] = {zero, zero};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bf16x4_t A = load4(W.w2t + (l & 31) * MLP_H + 8 * c + 4 * (l >> 5));
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const bf16x4_t B = pack4(dz2[q][4 * c], dz2[q][4 * c + 1], dz2[q][4 * c + 2], dz2[q][4 * c + 3]);
      dz1[q] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A, B, dz1[q], 0, 0, 0);
    }
  }
  float db1[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    db1[r] = 0.0f;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float d = s.z1[q][r] > 0.0f ? dz1[q][r] : 0.0f;
      dz1[q][r] = d;
      db1[r] += d;
    }
  }
  // transpose through LDS: sample-major rows of dZ2, H1, dZ1 and X
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int smp = 32 * q + (l & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = drow(r, l);
      s_dz2[wv][smp][i] = f2bf(dz2[q][r]);
      s_h1[wv][smp][i] = f2bf(s.h1[q][r]);
      s_dz1[wv][smp][i] = f2bf(dz1[q][r]);
    }
  }
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) s_x[wv][l][k] = f2bf(x[k]);
  __syncthreads();
  // dW2[i][k] = sum_s dZ2[i][s] H1[k][s] and dW1[i][k] = sum_s dZ1[i][s] X[k][s] over the 64 samples
  f32x16_t gw2 = zero, gw1 = zero;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int s0 = 8 * q + 4 * (l >> 5);
    bf16x4_t A2, B2, A1, B1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      A2[j] = (short)s_dz2[wv][s0 + j][l & 31];
      B2[j] = (short)s_h1[wv][s0 + j][l & 31];
      A1[j] = (short)s_dz1[wv][s0 + j][l & 31];
      B1[j] = (l & 31) < MLP_INP ? (short)s_x[wv][s0 + j][l & 31] : (short)0;
    }
    gw2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A2, B2, gw2, 0, 0, 0);
    gw1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(A1, B1, gw1, 0, 0, 0);
  }
  float* out = wgrad + (size_t)wave * MLP_NPARAM;
  const int col = l & 31;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = drow(r, l);
    out[P_W2 + i * MLP_H + col] = gw2[r];
    if (col < MLP_IN) out[P_W1 + i * MLP_IN + col] = gw1[r];
  }
  // bias and output-layer gradients: sum over the 32 lanes of each half
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float a = half_sum(dw3[r]), b = half_sum(db2[r]), c = half_sum(db1[r]);
    if (col == 0) {
      const int i = drow(r, l);
      out[P_W3 + i] = a;
      out[P_B2 + i] = b;
      out[P_B1 + i] = c;
    }
  }
  const float gb3 = wave_sum(g), ls = wave_sum(loss);
  if (l == 0) {
    out[P_B3] = gb3;
    wloss[wave] = ls;
  }
}

// Fixed-order sum of the per-wave gradients and losses.
__global__ void k_mlp_reduce(const float* __restrict__ wgrad, const float* __restrict__ wloss, int nw,
                             float* __restrict__ grad, float* __restrict__ loss) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < MLP_NPARAM) {
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += wgrad[(size_t)w * MLP_NPARAM + j];
    grad[j] = s;
  }
  if (j == MLP_NPARAM) {
    float s = 0.0f;
    for (int w = 0; w < nw; ++w) s += wloss[w];
    *loss = s;
  }
}

// Adam on the FP32 master parameters, then refresh of the device operand copy.
__global__ void k_adam(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= MLP_NPARAM) return;
  const float gj = grad[j];
  const float mj = b1 * m[j] + (1.0f - b1) * gj;
  const float vj = b2 * v[j] + (1.0f - b2) * gj * gj;
  m[j] = mj;
  v[j] = vj;
  prm[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

__global__ void k_pack(const float* __restrict__ prm, MlpDev* __restrict__ W) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < MLP_H * MLP_INP) {
    const int i = j / MLP_INP, k = j % MLP_INP;
    const float b = prm[P_B1 + i];
    const uint16_t hi = f2bf(b);
    W->w1[j] = k < MLP_IN ? f2bf(prm[P_W1 + i * MLP_IN + k]) : (k == MLP_IN ? hi : f2bf(b - bf2f(hi)));
  }
  if (j < MLP_H * MLP_H) {
    const int i = j / MLP_H, k = j % MLP_H;
    W->w2[j] = f2bf(prm[P_W2 + j]);
    W->w2t[j] = f2bf(prm[P_W2 + k * MLP_H + i]);
  }
  if (j < MLP_H) {
    W->b2[j] = prm[P_B2 + j];
    W->w3[j] = prm[P_W3 + j];
  }
  if (j == 0) W->b3 = prm[P_B3];
}

__global__ __launch_bounds__(256) void k_mlp_predict(const MlpDev* __restrict__ Wp, const float* __restrict__ X, int n,
                                                     float* __restrict__ y) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float x[MLP_INP];
#pragma unroll
  for (int k = 0; k < MLP_INP; ++k) x[k] = p < n ? X[(size_t)p * MLP_INP + k] : 0.0f;
  const float v = mlp_eval(*Wp, x);
  if (p < n) y[p] = v;
}

struct SgsArgs {
  const float* mac;        // fine ux, uy, uz, rho (4 x Nf)
  const uint8_t* flags;    // fine flags
  int fx, fy, fz;          // fine dims
  int cx, cy, cz;          // coarse dims (fine / 2)
  int stride, ox, oy, oz;  // sampling lattice on the coarse grid
  int sx, sy, sz;          // samples per axis
  float nu, tau0, smag;    // coarse molecular viscosity and its closure constants
  float* X;                // out: samples x 8
  float* T;                // out: samples
  int* valid;              // out: 1 when the sample is usable
};

// Mean velocity of the 2 x 2 x 2 block of fine cells under coarse cell (i, j, k) (false if a cell is not plain
// fluid) and, when tau is given, the block's central second moments <(u_a - <u_a>)(u_b - <u_b>)> (xx yy zz xy xz
// yz), formed from the deviations: the one-pass <u_a u_b> - <u_a><u_b> would cancel some six digits in FP32 where
// the flow is nearly uniform (u ~ 0.06 against a sub-filter variance ~ 1e-8).
__device__ __forceinline__ bool block_mean(const SgsArgs& a, int i, int j, int k, float* u, float* tau) {
  const size_t Nf = (size_t)a.fx * a.fy * a.fz;
  float v[8][3];
  float su[3] = {0, 0, 0};
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    const size_t idx = (size_t)(2 * i + dx) + (size_t)a.fx * ((2 * j + dy) + (size_t)a.fy * (2 * k + dz));
    if (a.flags[idx] != 0) return false;  // only plain fluid cells
    v[c][0] = a.mac[idx]; v[c][1] = a.mac[Nf + idx]; v[c][2] = a.mac[2 * Nf + idx];
    su[0] += v[c][0]; su[1] += v[c][1]; su[2] += v[c][2];
  }
  for (int c = 0; c < 3; ++c) u[c] = su[c] * 0.125f;
  if (tau) {
    float st[6] = {0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 8; ++c) {
      const float dx = v[c][0] - u[0], dy = v[c][1] - u[1], dz = v[c][2] - u[2];
      st[0] += dx * dx; st[1] += dy * dy; st[2] += dz * dz;
      st[3] += dx * dy; st[4] += dx * dz; st[5] += dy * dz;
    }
    for (int c = 0; c < 6; ++c) tau[c] = st[c] * 0.125f;
  }
  return true;
}

__global__ void k_sgs_samples(SgsArgs a) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ns = a.sx * a.sy * a.sz;
  if (t >= ns) return;
  const int i = a.ox + a.stride * (t % a.sx);
  const int j = a.oy + a.stride * ((t / a.sx) % a.sy);
  const int k = a.oz + a.stride * (t / (a.sx * a.sy));
  a.valid[t] = 0;
  if (i < 1 || j < 1 || k < 1 || i >= a.cx - 1 || j >= a.cy - 1 || k >= a.cz - 1) return;
  float u0[3], t6[6];
  if (!block_mean(a, i, j, k, u0, t6)) return;
  float up[3][3], um[3][3];
  const int di[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int b = 0; b < 3; ++b) {
    if (!block_mean(a, i + di[b][0], j + di[b][1], k + di[b][2], up[b], nullptr)) return;
    if (!block_mean(a, i - di[b][0], j - di[b][1], k - di[b][2], um[b], nullptr)) return;
  }
  // grad[a_][b] = d u_a / d x_b
  float gr[3][3];
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b) gr[c][b] = 0.5f * (up[b][c] - um[b][c]);
  const float Sxx = gr[0][0], Syy = gr[1][1], Szz = gr[2][2];
  const float Sxy = 0.5f * (gr[0][1] + gr[1][0]), Sxz = 0.5f * (gr[0][2] + gr[2][0]), Syz = 0.5f * (gr[1][2] + gr[2][1]);
  const float SS = Sxx * Sxx + Syy * Syy + Szz * Szz + 2.0f * (Sxy * Sxy + Sxz * Sxz + Syz * Syz);
  if (!(SS > 1e-14f)) return;
  float txx = t6[0], tyy = t6[1], tzz = t6[2];
  const float txy = t6[3], txz = t6[4], tyz = t6[5];
  const float m = (txx + tyy + tzz) * (1.0f / 3.0f);
  txx -= m; tyy -= m; tzz -= m;
  const float TS = txx * Sxx + tyy * Syy + tzz * Szz + 2.0f * (txy * Sxy + txz * Sxz + tyz * Syz);
  const float nu_r = -TS / (2.0f * SS);
  // Smagorinsky viscosity of the coarse grid for strain S: tau = tau0 + (smag / 6) |S| solves the relaxation
  // time equation with |P| = (2/3) tau |S|; the fine grid's filter width is half, its viscosity a quarter
  const float nu_f = 0.25f * (a.smag / 18.0f) * sqrtf(SS);
  const float nut = nu_r + nu_f;
  const float tau = fminf(1.95f, fmaxf(0.5015f, 3.0f * (a.nu + fmaxf(0.0f, nut)) + 0.5f));
  const float f = -(2.0f / 3.0f) * tau;
  const float pxx = f * Sxx, pyy = f * Syy, pzz = f * Szz, pxy = f * Sxy, pxz = f * Sxz, pyz = f * Syz;
  const float pn = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
  const float nuS = (smag_tau(a.tau0, a.smag, pn) - a.tau0) * (1.0f / 3.0f);
  if (!(nuS > 1e-12f)) return;
  float x[MLP_INP];
  mlp_features(pxx, pyy, pzz, pxy, pxz, pyz, a.nu, a.tau0, a.smag, x);
  const float ws = nuS * sqrtf(SS) / a.nu;
  x[7] = fminf(1e6f, fmaxf(1e-12f, ws * ws));
  for (int c = 0; c < MLP_INP; ++c) a.X[(size_t)t * MLP_INP + c] = x[c];
  a.T[t] = fminf(16.0f, fmaxf(-8.0f, nut / nuS));
  a.valid[t] = 1;
}

}  // namespace

struct aero_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  float* prm = nullptr;   // FP32 master parameters
  float* m = nullptr;     // Adam moments
  float* v = nullptr;
  MlpDev* W = nullptr;    // operand copy
  float* grad = nullptr;
  float* loss = nullptr;
  float* wgrad = nullptr;
  float* wloss = nullptr;
  int wcap = 0;
  float* X = nullptr;     // dataset
  float* T = nullptr;
  int64_t n = 0, cap = 0;
  int64_t adam_t = 0;
  std::string err;
};

#define M_HIP(mm, call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      (mm)->err = std::string(#call) + ": " + hipGetErrorString(e_);           \
      return AERO_E_HIP;                                                        \
    }                                                                           \
  } while (0)

namespace {

int ensure_waves(aero_mlp* M, int nw) {
  if (nw <= M->wcap) return AERO_OK;
  if (M->wgrad) hipFree(M->wgrad);
  if (M->wloss) hipFree(M->wloss);
  M->wgrad = nullptr;
  M->wloss = nullptr;
  M_HIP(M, hipMalloc(&M->wgrad, (size_t)nw * MLP_NPARAM * 4));
  M_HIP(M, hipMalloc(&M->wloss, (size_t)nw * 4));
  M->wcap = nw;
  return AERO_OK;
}

// gradient of the mean loss over n samples at X/T (device pointers); result in M->grad, M->loss
int grad_dev(aero_mlp* M, const float* X, const float* T, int n) {
  const int nw = (n + 63) / 64;
  const int nblk = (nw + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
  int rc = ensure_waves(M, nblk * WAVES_PER_BLOCK);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(64 * WAVES_PER_BLOCK), 0, M->stream, M->W, X, T, n, 1.0f / n,
                     M->wgrad, M->wloss);
  M_HIP(M, hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, M->stream, M->wgrad, M->wloss,
                     nblk * WAVES_PER_BLOCK, M->grad, M->loss);
  M_HIP(M, hipGetLastError());
  return AERO_OK;
}

int pack(aero_mlp* M) {
  hipLaunchKernelGGL(k_pack, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, M->stream, M->prm, M->W);
  M_HIP(M, hipGetLastError());
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_mlp_create(int32_t device, aero_mlp** out) {
  if (!out) return AERO_E_ARG;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return AERO_E_HIP;
  aero_mlp* M = new aero_mlp();
  M->dev = device;
  bool ok = hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&M->prm, MLP_NPARAM * 4) == hipSuccess && hipMalloc(&M->m, MLP_NPARAM * 4) == hipSuccess &&
            hipMalloc(&M->v, MLP_NPARAM * 4) == hipSuccess && hipMalloc(&M->W, sizeof(MlpDev)) == hipSuccess &&
            hipMalloc(&M->grad, MLP_NPARAM * 4) == hipSuccess && hipMalloc(&M->loss, 4) == hipSuccess;
  if (!ok) {
    aero_mlp_destroy(M);
    return AERO_E_NOMEM;
  }
  hipMemset(M->W, 0, sizeof(MlpDev));
  hipMemset(M->prm, 0, MLP_NPARAM * 4);
  hipMemset(M->m, 0, MLP_NPARAM * 4);
  hipMemset(M->v, 0, MLP_NPARAM * 4);
  hipDeviceSynchronize();
  *out = M;
  return AERO_OK;
}

int aero_mlp_destroy(aero_mlp* M) {
  if (!M) return AERO_OK;
  hipSetDevice(M->dev);
  if (M->stream) hipStreamSynchronize(M->stream);
  float* bufs[] = {M->prm, M->m, M->v, M->grad, M->loss, M->wgrad, M->wloss, M->X, M->T};
  for (float* b : bufs)
    if (b) hipFree(b);
  if (M->W) hipFree(M->W);
  if (M->stream) hipStreamDestroy(M->stream);
  delete M;
  return AERO_OK;
}

const char* aero_mlp_last_error(const aero_mlp* M) { return M ? M->err.c_str() : ""; }

int aero_mlp_set_params(aero_mlp* M, const float* params) {
  if (!M || !params) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  for (int j = 0; j < MLP_NPARAM; ++j)
    if (!std::isfinite(params[j])) { M->err = "aero_mlp_set_params: non-finite parameter"; return AERO_E_ARG; }
  M_HIP(M, hipMemcpyAsync(M->prm, params, MLP_NPARAM * 4, hipMemcpyHostToDevice, M->stream));
  M_HIP(M, hipMemsetAsync(M->m, 0, MLP_NPARAM * 4, M->stream));
  M_HIP(M, hipMemsetAsync(M->v, 0, MLP_NPARAM * 4, M->stream));
  M->adam_t = 0;
  int rc = pack(M);
  if (rc) return rc;
  M_HIP(M, hipStreamSynchronize(M->stream));
  return AERO_OK;
}

int aero_mlp_get_params(aero_mlp* M, float* params) {
  if (!M || !params) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  M_HIP(M, hipMemcpyAsync(params, M->prm, MLP_NPARAM * 4, hipMemcpyDeviceToHost, M->stream));
  M_HIP(M, hipStreamSynchronize(M->stream));
  return AERO_OK;
}

int aero_mlp_clear_samples(aero_mlp* M) {
  if (!M) return AERO_E_ARG;
  M->n = 0;
  return AERO_OK;
}

static int reserve(aero_mlp* M, int64_t need) {
  if (need <= M->cap) return AERO_OK;
  int64_t cap = std::max<int64_t>(need, 2 * M->cap);
  float *X = nullptr, *T = nullptr;
  M_HIP(M, hipMalloc(&X, (size_t)cap * MLP_INP * 4));
  M_HIP(M, hipMalloc(&T, (size_t)cap * 4));
  if (M->n > 0) {
    M_HIP(M, hipMemcpyAsync(X, M->X, (size_t)M->n * MLP_INP * 4, hipMemcpyDeviceToDevice, M->stream));
    M_HIP(M, hipMemcpyAsync(T, M->T, (size_t)M->n * 4, hipMemcpyDeviceToDevice, M->stream));
    M_HIP(M, hipStreamSynchronize(M->stream));
  }
  if (M->X) hipFree(M->X);
  if (M->T) hipFree(M->T);
  M->X = X;
  M->T = T;
  M->cap = cap;
  return AERO_OK;
}

int aero_mlp_add_samples(aero_mlp* M, const float* X, const float* T, int64_t n) {
  if (!M || !X || !T || n < 0) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  int rc = reserve(M, M->n + n);
  if (rc) return rc;
  M_HIP(M, hipMemcpyAsync(M->X + (size_t)M->n * MLP_INP, X, (size_t)n * MLP_INP * 4, hipMemcpyHostToDevice, M->stream));
  M_HIP(M, hipMemcpyAsync(M->T + (size_t)M->n, T, (size_t)n * 4, hipMemcpyHostToDevice, M->stream));
  M_HIP(M, hipStreamSynchronize(M->stream));
  M->n += n;
  return AERO_OK;
}

int aero_mlp_num_samples(aero_mlp* M, int64_t* n) {
  if (!M || !n) return AERO_E_ARG;
  *n = M->n;
  return AERO_OK;
}

int aero_mlp_get_samples(aero_mlp* M, float* X, float* T) {
  if (!M) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  if (X) M_HIP(M, hipMemcpy(X, M->X, (size_t)M->n * MLP_INP * 4, hipMemcpyDeviceToHost));
  if (T) M_HIP(M, hipMemcpy(T, M->T, (size_t)M->n * 4, hipMemcpyDeviceToHost));
  return AERO_OK;
}

// Appends training samples taken from the last readout fields of a fine-grid context (run with
// AERO_STEP_READOUT), filtered onto the coarse grid of half its size in each axis.
int aero_mlp_sample_fine(aero_mlp* M, aero_ctx* fine, int32_t stride, uint64_t seed, double nu_coarse, double cs,
                         int64_t* added) {
  if (!M || !fine || stride < 1 || !(nu_coarse > 0)) return AERO_E_ARG;
  if (!fine->readout_valid) { M->err = "aero_mlp_sample_fine: the fine context has no readout fields"; return AERO_E_STATE; }
  if (fine->dev != M->dev) { M->err = "aero_mlp_sample_fine: device mismatch"; return AERO_E_ARG; }
  M_HIP(M, hipSetDevice(M->dev));
  M_HIP(M, hipStreamSynchronize(fine->stream));
  SgsArgs a;
  a.mac = fine->mac;
  a.flags = fine->flags;
  a.fx = fine->nx; a.fy = fine->ny; a.fz = fine->nz;
  a.cx = fine->nx / 2; a.cy = fine->ny / 2; a.cz = fine->nz / 2;
  a.stride = stride;
  a.ox = 1 + (int)(seed % stride);
  a.oy = 1 + (int)((seed / stride) % stride);
  a.oz = 1 + (int)((seed / ((uint64_t)stride * stride)) % stride);
  a.sx = (a.cx - 1 - a.ox + stride - 1) / stride;
  a.sy = (a.cy - 1 - a.oy + stride - 1) / stride;
  a.sz = (a.cz - 1 - a.oz + stride - 1) / stride;
  if (a.sx <= 0 || a.sy <= 0 || a.sz <= 0) { *added = 0; return AERO_OK; }
  a.nu = (float)nu_coarse;
  a.tau0 = (float)(3.0 * nu_coarse + 0.5);
  a.smag = (float)(18.0 * M_SQRT2 * cs * cs);
  const int ns = a.sx * a.sy * a.sz;
  float *X, *T;
  int* valid;
  M_HIP(M, hipMalloc(&X, (size_t)ns * MLP_INP * 4));
  M_HIP(M, hipMalloc(&T, (size_t)ns * 4));
  M_HIP(M, hipMalloc(&valid, (size_t)ns * 4));
  a.X = X; a.T = T; a.valid = valid;
  hipLaunchKernelGGL(k_sgs_samples, dim3((ns + 255) / 256), dim3(256), 0, M->stream, a);
  M_HIP(M, hipGetLastError());
  std::vector<float> hX((size_t)ns * MLP_INP), hT(ns);
  std::vector<int> hv(ns);
  M_HIP(M, hipMemcpyAsync(hX.data(), X, hX.size() * 4, hipMemcpyDeviceToHost, M->stream));
  M_HIP(M, hipMemcpyAsync(hT.data(), T, hT.size() * 4, hipMemcpyDeviceToHost, M->stream));
  M_HIP(M, hipMemcpyAsync(hv.data(), valid, hv.size() * 4, hipMemcpyDeviceToHost, M->stream));
  M_HIP(M, hipStreamSynchronize(M->stream));
  hipFree(X); hipFree(T); hipFree(valid);
  std::vector<float> cX, cT;
  for (int i = 0; i < ns; ++i)
    if (hv[i]) {
      cX.insert(cX.end(), hX.begin() + (size_t)i * MLP_INP, hX.begin() + (size_t)(i + 1) * MLP_INP);
      cT.push_back(hT[i]);
    }
  if (added) *added = (int64_t)cT.size();
  if (cT.empty()) return AERO_OK;
  return aero_mlp_add_samples(M, cX.data(), cT.data(), (int64_t)cT.size());
}

// Gradient of the mean loss over host samples (for checks against the CPU oracle).
int aero_mlp_grad(aero_mlp* M, const float* X, const float* T, int32_t n, float* grad, double* loss) {
  if (!M || !X || !T || n <= 0 || !grad) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  float *dX, *dT;
  M_HIP(M, hipMalloc(&dX, (size_t)n * MLP_INP * 4));
  M_HIP(M, hipMalloc(&dT, (size_t)n * 4));
  M_HIP(M, hipMemcpyAsync(dX, X, (size_t)n * MLP_INP * 4, hipMemcpyHostToDevice, M->stream));
  M_HIP(M, hipMemcpyAsync(dT, T, (size_t)n * 4, hipMemcpyHostToDevice, M->stream));
  int rc = grad_dev(M, dX, dT, n);
  float hl = 0;
  if (!rc) {
    M_HIP(M, hipMemcpyAsync(grad, M->grad, MLP_NPARAM * 4, hipMemcpyDeviceToHost, M->stream));
    M_HIP(M, hipMemcpyAsync(&hl, M->loss, 4, hipMemcpyDeviceToHost, M->stream));
    M_HIP(M, hipStreamSynchronize(M->stream));
  }
  hipFree(dX);
  hipFree(dT);
  if (loss) *loss = hl / n;
  return rc;
}

// Adam over the stored samples in contiguous mini-batches (the samples should be shuffled when added).
// loss_out (may be NULL) receives the mean training loss of every iteration.
int aero_mlp_train(aero_mlp* M, int32_t iters, int32_t batch, double lr, float* loss_out) {
  if (!M || iters < 0 || batch < 64) return AERO_E_ARG;
  if (M->n < batch) { M->err = "aero_mlp_train: fewer samples than one batch"; return AERO_E_STATE; }
  M_HIP(M, hipSetDevice(M->dev));
  std::vector<float> losses(iters, 0.0f);
  float* dl = nullptr;
  if (loss_out) M_HIP(M, hipMalloc(&dl, (size_t)std::max(1, iters) * 4));
  const int64_t nb = M->n / batch;
  for (int it = 0; it < iters; ++it) {
    const int64_t start = (it % nb) * batch;
    int rc = grad_dev(M, M->X + (size_t)start * MLP_INP, M->T + start, batch);
    if (rc) return rc;
    ++M->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)M->adam_t), c2 = 1.0f - std::pow(b2, (float)M->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, M->stream, M->prm, M->m, M->v, M->grad,
                       (float)lr, b1, b2, 1e-8f, c1, c2);
    M_HIP(M, hipGetLastError());
    rc = pack(M);
    if (rc) return rc;
    if (dl) M_HIP(M, hipMemcpyAsync(dl + it, M->loss, 4, hipMemcpyDeviceToDevice, M->stream));
  }
  if (dl) {
    M_HIP(M, hipMemcpyAsync(losses.data(), dl, (size_t)iters * 4, hipMemcpyDeviceToHost, M->stream));
    M_HIP(M, hipStreamSynchronize(M->stream));
    for (int i = 0; i < iters; ++i) loss_out[i] = losses[i] / batch;
    hipFree(dl);
  }
  M_HIP(M, hipStreamSynchronize(M->stream));
  return AERO_OK;
}

int aero_mlp_predict(aero_mlp* M, const float* X, int32_t n, float* y) {
  if (!M || !X || n <= 0 || !y) return AERO_E_ARG;
  M_HIP(M, hipSetDevice(M->dev));
  float *dX, *dy;
  M_HIP(M, hipMalloc(&dX, (size_t)n * MLP_INP * 4));
  M_HIP(M, hipMalloc(&dy, (size_t)n * 4));
  M_HIP(M, hipMemcpyAsync(dX, X, (size_t)n * MLP_INP * 4, hipMemcpyHostToDevice, M->stream));
  hipLaunchKernelGGL(k_mlp_predict, dim3((n + 255) / 256), dim3(256), 0, M->stream, M->W, dX, n, dy);
  M_HIP(M, hipGetLastError());
  M_HIP(M, hipMemcpyAsync(y, dy, (size_t)n * 4, hipMemcpyDeviceToHost, M->stream));
  M_HIP(M, hipStreamSynchronize(M->stream));
  hipFree(dX);
  hipFree(dy);
  return AERO_OK;
}

}  // extern "C"

// The solver's closure reads the operand copy of an aero_mlp (see aero_set_closure in aerolbm_api.cpp).
const void* aero_mlp_device_weights(const aero_mlp* M) { return M ? M->W : nullptr; }
int aero_mlp_device(const aero_mlp* M) { return M ? M->dev : -1; }
