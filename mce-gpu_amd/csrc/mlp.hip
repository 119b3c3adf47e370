// Learned cost model: features, FP32 training (forward, backward, Adam) and
// BF16 MFMA inference, all hand-written for gfx950. No reference counterpart:
// the model predicts log(1 + BK tree nodes) of a first-level subproblem and
// only orders work; results never depend on it.
//
// Training: one wave per row with lanes as the 64 hidden units. Layer weights
// sit in LDS (W2 rows padded to 65 floats, so the row and the column access
// are both bank-conflict free); each wave accumulates its weight gradients in
// registers and writes one partial vector; a second kernel sums the partials
// in a fixed order, so training is deterministic for a given seed.
//
// Inference: 16-row tiles through two v_mfma_f32_16x16x16_bf16 layers
// (4 + 16 MFMAs per tile), FP32 accumulation, ReLU and the final 64-wide dot
// product on the VALU. Operand layouts of the 16x16x16 instruction:
// A[i][k] in lane i + 16*(k/4), element k%4; B[k][j] in lane j + 16*(k/4),
// element k%4; D[i][j] in lane j + 16*(i/4), element i%4.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "engine.hpp"
#include "internal.hpp"
#include "rng.hpp"
#include "wave.hpp"

namespace mce {

namespace {

constexpr int kF = MCE_MLP_FEATURES, kH = MCE_MLP_HIDDEN, kP = MCE_MLP_PARAMS;
constexpr int kOffW1 = 0, kOffB1 = kF * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH;
static_assert(kOffB3 + 1 == kP, "parameter layout");

__device__ __forceinline__ float wave_sum_f32(float x) {
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

__device__ __forceinline__ float rdlane_f(float v, uint32_t l) {
  return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), l));
}

// ---- features ----------------------------------------------------------

__global__ __launch_bounds__(256) void features_kernel(uint32_t n, const uint64_t *__restrict__ rp,
                                                       const uint32_t *__restrict__ oc,
                                                       const uint64_t *__restrict__ split,
                                                       const uint32_t *__restrict__ dplus,
                                                       const uint32_t *__restrict__ prio,
                                                       const uint32_t *__restrict__ core, float degeneracy,
                                                       float rounds, float *__restrict__ F) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint32_t L = lane_id();
  for (uint32_t v = wid; v < n; v += waves) {
    const uint64_t b = rp[v], s = split[v], e = rp[v + 1];
    const uint32_t p = (uint32_t)(e - s), q = (uint32_t)(s - b), d = (uint32_t)(e - b);
    float s_out = 0, s_deg = 0, mx = 0, s_in = 0;
    for (uint64_t k = s + L; k < e; k += kWave) {
      const uint32_t u = oc[k];
      s_out += (float)dplus[u];
      s_deg += (float)(rp[u + 1] - rp[u]);
      mx = fmaxf(mx, (float)dplus[u]);
    }
    for (uint64_t k = b + L; k < s; k += kWave) s_in += (float)dplus[oc[k]];
    s_out = wave_sum_f32(s_out);
    s_deg = wave_sum_f32(s_deg);
    s_in = wave_sum_f32(s_in);
    for (int d2 = 32; d2 >= 1; d2 >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d2, 64));
    // edges inside P = N+(v): for u in P, |N+(u) & P| by binary search in P
    uint32_t ep = 0;
    const uint32_t *P = oc + s;
    for (uint64_t k = s; k < e; ++k) {
      const uint32_t u = oc[k];
      for (uint64_t t = split[u] + L; t < rp[u + 1]; t += kWave) {
        const uint32_t y = oc[t];
        const uint32_t pos = lower_bound_u32(P, p, y);
        ep += (pos < p && P[pos] == y) ? 1u : 0u;
      }
    }
    const float eP = (float)wave_sum(ep);
    const float pf = (float)p;
    const float dens = p > 1 ? 2.f * eP / (pf * (pf - 1.f)) : 0.f;
    float f[kF];
    f[0] = log1pf((float)d);
    f[1] = log1pf(pf);
    f[2] = log1pf((float)q);
    f[3] = (float)core[v] / fmaxf(degeneracy, 1.f);
    f[4] = log1pf((float)core[v]);
    f[5] = log1pf(s_out);
    f[6] = log1pf(s_in);
    f[7] = log1pf(eP);
    f[8] = dens;
    f[9] = p ? 2.f * eP / pf : 0.f;
    f[10] = log1pf(mx);
    f[11] = (float)prio[v] / fmaxf(rounds, 1.f);
    f[12] = log1pf(s_deg);
    f[13] = log1pf(pf) * dens;
    f[14] = log1pf(pf * (float)q);
    f[15] = 1.f;
    if (L < kF) {
      float val = f[0];
#pragma unroll
      for (int i = 1; i < kF; ++i)
        if ((int)L == i) val = f[i];
      F[(uint64_t)v * kF + L] = val;
    }
  }
}

// ---- training ----------------------------------------------------------

struct TrainArgs {
  const float *x;       // rows x 16
  const float *y;       // rows
  const uint32_t *idx;  // rows of this step
  uint32_t nrows;
  float inv_batch;      // 1 / rows of the step
  const float *params;  // kP
  const float *mu, *sigma;
  float *partial;       // waves x kP
  float *loss;          // waves
};

__global__ __launch_bounds__(256) void train_rows_kernel(TrainArgs a) {
  __shared__ float sW1[kF * kH];
  __shared__ float sW2[kH * (kH + 1)];
  __shared__ float sB1[kH], sB2[kH], sW3[kH];
  for (int i = threadIdx.x; i < kF * kH; i += blockDim.x) sW1[i] = a.params[kOffW1 + i];
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) sW2[(i / kH) * (kH + 1) + i % kH] = a.params[kOffW2 + i];
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    sB1[i] = a.params[kOffB1 + i];
    sB2[i] = a.params[kOffB2 + i];
    sW3[i] = a.params[kOffW3 + i];
  }
  __syncthreads();
  const float b3 = a.params[kOffB3];
  const uint32_t L = lane_id();
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  float w1c[kF];
#pragma unroll
  for (int k = 0; k < kF; ++k) w1c[k] = sW1[k * kH + L];
  const float b1 = sB1[L], b2 = sB2[L], w3 = sW3[L];
  const float mu = L < kF ? a.mu[L] : 0.f, isg = L < kF ? 1.f / a.sigma[L] : 0.f;
  float gW1[kF], gW2[kH];
#pragma unroll
  for (int k = 0; k < kF; ++k) gW1[k] = 0.f;
#pragma unroll
  for (int i = 0; i < kH; ++i) gW2[i] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, loss = 0.f;
  for (uint32_t r = wid; r < a.nrows; r += waves) {
    const uint32_t row = a.idx[r];
    const float xk = L < kF ? (a.x[(uint64_t)row * kF + L] - mu) * isg : 0.f;
    float h1 = b1;
#pragma unroll
    for (int k = 0; k < kF; ++k) h1 += rdlane_f(xk, k) * w1c[k];
    h1 = fmaxf(h1, 0.f);
    float h2 = b2;
#pragma unroll 8
    for (int i = 0; i < kH; ++i) h2 += rdlane_f(h1, i) * sW2[i * (kH + 1) + L];
    h2 = fmaxf(h2, 0.f);
    const float yhat = wave_sum_f32(h2 * w3) + b3;
    const float err = yhat - a.y[row];
    loss += err * err;
    const float dy = 2.f * err * a.inv_batch;
    const float dh2 = h2 > 0.f ? dy * w3 : 0.f;
    float dh1 = 0.f;
#pragma unroll 8
    for (int j = 0; j < kH; ++j) dh1 += rdlane_f(dh2, j) * sW2[L * (kH + 1) + j];
    dh1 = h1 > 0.f ? dh1 : 0.f;
    gw3 += dy * h2;
    gb3 += dy;
    gb2 += dh2;
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < kH; ++i) gW2[i] += rdlane_f(h1, i) * dh2;
#pragma unroll
    for (int k = 0; k < kF; ++k) gW1[k] += rdlane_f(xk, k) * dh1;
  }
  float *out = a.partial + (uint64_t)wid * kP;
#pragma unroll
  for (int k = 0; k < kF; ++k) out[kOffW1 + k * kH + L] = gW1[k];
  out[kOffB1 + L] = gb1;
#pragma unroll
  for (int i = 0; i < kH; ++i) out[kOffW2 + i * kH + L] = gW2[i];
  out[kOffB2 + L] = gb2;
  out[kOffW3 + L] = gw3;
  if (L == 0) {
    out[kOffB3] = gb3;
    a.loss[wid] = loss;
  }
}

// grad[p] = sum over waves of partial[w][p], in wave order (deterministic).
__global__ void reduce_partials_kernel(const float *__restrict__ partial, uint32_t waves, float *__restrict__ grad) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kP) return;
  float s = 0.f;
  for (uint32_t w = 0; w < waves; ++w) s += partial[(uint64_t)w * kP + p];
  grad[p] = s;
}

__global__ void adam_kernel(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                            float *__restrict__ v, float lr, float b1, float b2, float eps, float wd, float c1,
                            float c2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)kP) return;
  const bool is_bias = (p >= (uint32_t)kOffB1 && p < (uint32_t)kOffW2) || (p >= (uint32_t)kOffB2 && p < (uint32_t)kOffW3) ||
                       p == (uint32_t)kOffB3;
  const float g = grad[p] + (is_bias ? 0.f : wd * params[p]);
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---- BF16 MFMA inference -----------------------------------------------

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

__global__ __launch_bounds__(256) void predict_kernel(uint64_t rows, const float *__restrict__ x,
                                                      const float *__restrict__ params, const float *__restrict__ mu,
                                                      const float *__restrict__ sigma, float *__restrict__ out) {
  __shared__ short sW1[kF * kH];      // [k][j]
  __shared__ short sW2[kH * kH];      // [k][j]
  __shared__ float sB1[kH], sB2[kH], sW3[kH];
  __shared__ short sH[4][16][kH + 4];  // per-wave layer-1 tile, [row][k]
  for (int i = threadIdx.x; i < kF * kH; i += blockDim.x) sW1[i] = to_bf16(params[kOffW1 + i]);
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) sW2[i] = to_bf16(params[kOffW2 + i]);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    sB1[i] = params[kOffB1 + i];
    sB2[i] = params[kOffB2 + i];
    sW3[i] = params[kOffW3 + i];
  }
  __syncthreads();
  const float b3 = params[kOffB3];
  const uint32_t L = lane_id(), wig = threadIdx.x / kWave;
  const uint32_t lr = L % 16, lk = L / 16;  // row/column within a tile, k-group
  const uint64_t tiles = (rows + 15) / 16;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / kWave);
  // B operands of layer 1 do not depend on the tile
  bf16x4 bw1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int t = 0; t < 4; ++t) bw1[nt][t] = sW1[(4 * lk + t) * kH + 16 * nt + lr];
  float imu[4], isg[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    imu[t] = mu[4 * lk + t];
    isg[t] = 1.f / sigma[4 * lk + t];
  }
  for (uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x / kWave) + wig; tile < tiles; tile += waves) {
    const uint64_t r0 = tile * 16;
    // A of layer 1: row lr, features 4*lk .. 4*lk+3
    bf16x4 a1;
    const uint64_t row = r0 + lr;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float v = row < rows ? (x[row * kF + 4 * lk + t] - imu[t]) * isg[t] : 0.f;
      a1[t] = to_bf16(v);
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bw1[nt], acc, 0, 0, 0);
      const int col = 16 * nt + lr;
#pragma unroll
      for (int t = 0; t < 4; ++t) sH[wig][4 * lk + t][col] = to_bf16(fmaxf(acc[t] + sB1[col], 0.f));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        bf16x4 a2, b2;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int k = 16 * kt + 4 * lk + t;
          a2[t] = sH[wig][lr][k];
          b2[t] = sW2[k * kH + 16 * nt + lr];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, acc, 0, 0, 0);
      }
      const int col = 16 * nt + lr;
#pragma unroll
      for (int t = 0; t < 4; ++t) yp[t] += fmaxf(acc[t] + sB2[col], 0.f) * sW3[col];
    }
    // sum the 16 columns held by the lanes of each 16-lane group
#pragma unroll
    for (int t = 0; t < 4; ++t)
      for (int d = 8; d >= 1; d >>= 1) yp[t] += __shfl_xor(yp[t], d, 64);
    if (lr == 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint64_t r = r0 + 4 * lk + t;
        if (r < rows) out[r] = yp[t] + b3;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
}

// Schedule keys from predictions: key = clamp(64 * pred, 0, 4095); counts
// non-finite predictions; accumulates the predicted work sum and maximum.
__global__ void keys_kernel(uint32_t n, const float *__restrict__ pred, uint32_t *__restrict__ key,
                            unsigned int *bad, float *work) {
  float s = 0.f, mx = 0.f;
  uint32_t nb = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const float p = pred[v];
    if (!isfinite(p)) {
      ++nb;
      key[v] = 0;
      continue;
    }
    key[v] = (uint32_t)fminf(fmaxf(p * 64.f, 0.f), 4095.f);
    const float w = expm1f(fminf(fmaxf(p, 0.f), 60.f));
    s += w;
    mx = fmaxf(mx, w);
  }
  if (nb) atomicAdd(bad, nb);
  atomicAdd(&work[0], s);
  atomicMax((unsigned int *)&work[1], __float_as_uint(mx));  // non-negative floats order as their bits
}

struct DevBuf {
  std::vector<void *> ptrs;
  ~DevBuf() {
    for (void *p : ptrs) hipFree(p);
  }
  int alloc(void **p, size_t bytes) {
    if (hipMalloc(p, std::max<size_t>(bytes, 4)) != hipSuccess)
      return set_error(MCE_ENOMEM, "device allocation of %zu bytes failed", bytes);
    ptrs.push_back(*p);
    return MCE_OK;
  }
};

int set_device(int device) {
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return set_error(MCE_EDEVICE, "no HIP device available");
  if (device < 0 || device >= nd) return set_error(MCE_EINVAL, "device %d out of range", device);
  MCE_HIP(hipSetDevice(device));
  return MCE_OK;
}

bool model_ok(const mce_mlp *m) {
  for (int i = 0; i < kF; ++i)
    if (!std::isfinite(m->mu[i]) || !std::isfinite(m->sigma[i]) || !(m->sigma[i] > 0)) return false;
  for (int i = 0; i < kP; ++i)
    if (!std::isfinite(m->params[i])) return false;
  return true;
}

constexpr char kMagic[8] = {'M', 'C', 'E', 'M', 'L', 'P', '0', '1'};

}  // namespace

// Computes features of every vertex into device memory F (n x 16).
int features_device(const DevGraph &g, const OrderInfo &info, float *F, hipStream_t s) {
  if (g.n == 0) return MCE_OK;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(8192, ((uint64_t)g.n * 64 + 255) / 256);
  features_kernel<<<grid, 256, 0, s>>>(g.n, g.rp, g.oc, g.split, g.dplus, g.prio, g.core, (float)info.degeneracy,
                                       (float)info.n_rounds, F);
  MCE_HIP(hipGetLastError());
  return MCE_OK;
}

// Predictions for `rows` device feature rows into device `out`.
int predict_device(const mce_mlp &m, const float *F, uint64_t rows, float *out, hipStream_t s) {
  if (rows == 0) return MCE_OK;
  DevBuf b;
  float *p = nullptr;
  int rc = b.alloc((void **)&p, sizeof(float) * (kP + 2 * kF));
  if (rc) return rc;
  MCE_HIP(hipMemcpyAsync(p, m.params, sizeof(float) * kP, hipMemcpyHostToDevice, s));
  MCE_HIP(hipMemcpyAsync(p + kP, m.mu, sizeof(float) * kF, hipMemcpyHostToDevice, s));
  MCE_HIP(hipMemcpyAsync(p + kP + kF, m.sigma, sizeof(float) * kF, hipMemcpyHostToDevice, s));
  const uint64_t tiles = (rows + 15) / 16;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (tiles + 3) / 4));
  predict_kernel<<<grid, 256, 0, s>>>(rows, F, p, p + kP, p + kP + kF, out);
  MCE_HIP(hipGetLastError());
  MCE_HIP(hipStreamSynchronize(s));
  return MCE_OK;
}

bool mlp_valid(const mce_mlp *m) { return m && model_ok(m); }

int learned_keys(const DevGraph &g, const OrderInfo &info, const mce_mlp &m, uint32_t *keys, uint32_t *bad_out,
                 double *work_sum, double *work_max, hipStream_t s, Workspace *ws) {
  *bad_out = 0;
  *work_sum = *work_max = 0;
  if (g.n == 0) return MCE_OK;
  DevBuf b;
  float *F, *pred, *work;
  unsigned int *bad;
  int rc;
  // features | predictions | 2 work sums | bad count, 256-byte aligned parts
  const size_t fb = (sizeof(float) * (uint64_t)g.n * kF + 255) & ~(size_t)255,
               pb = (sizeof(float) * (uint64_t)g.n + 255) & ~(size_t)255;
  uint8_t *base = nullptr;
  if (ws) {
    base = (uint8_t *)ws->get(kWsFeat, fb + pb + 256);
    if (!base) return set_error(MCE_ENOMEM, "device allocation for the cost-model features failed");
  } else if ((rc = b.alloc((void **)&base, fb + pb + 256))) {
    return rc;
  }
  F = (float *)base;
  pred = (float *)(base + fb);
  work = (float *)(base + fb + pb);
  bad = (unsigned int *)(base + fb + pb + 128);
  if ((rc = features_device(g, info, F, s))) return rc;
  if ((rc = predict_device(m, F, g.n, pred, s))) return rc;
  MCE_HIP(hipMemsetAsync(work, 0, 2 * sizeof(float), s));
  MCE_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned int), s));
  keys_kernel<<<1024, 256, 0, s>>>(g.n, pred, keys, bad, work);
  float hw[2];
  unsigned int hb = 0;
  MCE_HIP(hipMemcpyAsync(hw, work, sizeof(hw), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s));
  MCE_HIP(hipStreamSynchronize(s));
  *bad_out = hb;
  *work_sum = hw[0];
  *work_max = hw[1];
  return MCE_OK;
}

}  // namespace mce

using namespace mce;

extern "C" {

int mce_mlp_init(mce_mlp *m, const float *x, uint64_t rows, uint64_t seed) {
  clear_error();
  if (!m || (rows && !x)) return set_error(MCE_EINVAL, "null argument");
  memset(m, 0, sizeof(*m));
  for (int f = 0; f < kF; ++f) {
    double s = 0, s2 = 0;
    for (uint64_t r = 0; r < rows; ++r) {
      const double v = x[r * kF + f];
      s += v;
      s2 += v * v;
    }
    const double mean = rows ? s / rows : 0.0;
    const double var = rows ? std::max(0.0, s2 / rows - mean * mean) : 1.0;
    m->mu[f] = (float)mean;
    m->sigma[f] = (float)(var > 1e-12 ? sqrt(var) : 1.0);
  }
  // He-uniform initial weights, zero biases
  auto uni = [&](uint64_t i) { return (float)(2.0 * u01(rng_at(seed, 11, i)) - 1.0); };
  const float a1 = sqrtf(6.f / kF), a2 = sqrtf(6.f / kH), a3 = sqrtf(3.f / kH);
  for (int i = 0; i < kF * kH; ++i) m->params[kOffW1 + i] = a1 * uni(i);
  for (int i = 0; i < kH * kH; ++i) m->params[kOffW2 + i] = a2 * uni(10000 + i);
  for (int i = 0; i < kH; ++i) m->params[kOffW3 + i] = a3 * uni(20000 + i);
  return MCE_OK;
}

int mce_mlp_grad(const mce_mlp *m, const float *x, const float *y, uint64_t rows, int device, float *grad,
                 double *loss_out) {
  clear_error();
  if (!m || !x || !y || !grad || rows == 0) return set_error(MCE_EINVAL, "null argument or no rows");
  if (rows > 0xFFFFFFFFull) return set_error(MCE_ELIMIT, "too many rows");
  int rc = set_device(device);
  if (rc) return rc;
  DevBuf b;
  float *dx, *dy, *dp, *dpart, *dloss, *dg;
  uint32_t *didx;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(64, (rows + 255) / 256 + 1), waves = blocks * 4;
  if ((rc = b.alloc((void **)&dx, sizeof(float) * rows * kF)) || (rc = b.alloc((void **)&dy, sizeof(float) * rows)) ||
      (rc = b.alloc((void **)&dp, sizeof(float) * (kP + 2 * kF))) || (rc = b.alloc((void **)&didx, 4 * rows)) ||
      (rc = b.alloc((void **)&dpart, sizeof(float) * kP * waves)) || (rc = b.alloc((void **)&dloss, 4 * waves)) ||
      (rc = b.alloc((void **)&dg, sizeof(float) * kP)))
    return rc;
  std::vector<uint32_t> idx(rows);
  std::iota(idx.begin(), idx.end(), 0u);
  MCE_HIP(hipMemcpy(dx, x, sizeof(float) * rows * kF, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dy, y, sizeof(float) * rows, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(didx, idx.data(), 4 * rows, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp, m->params, sizeof(float) * kP, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp + kP, m->mu, sizeof(float) * kF, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp + kP + kF, m->sigma, sizeof(float) * kF, hipMemcpyHostToDevice));
  TrainArgs a{dx, dy, didx, (uint32_t)rows, 1.f / (float)rows, dp, dp + kP, dp + kP + kF, dpart, dloss};
  train_rows_kernel<<<blocks, 256>>>(a);
  reduce_partials_kernel<<<(kP + 255) / 256, 256>>>(dpart, waves, dg);
  MCE_HIP(hipGetLastError());
  MCE_HIP(hipMemcpy(grad, dg, sizeof(float) * kP, hipMemcpyDeviceToHost));
  std::vector<float> hl(waves);
  MCE_HIP(hipMemcpy(hl.data(), dloss, 4 * waves, hipMemcpyDeviceToHost));
  if (loss_out) {
    double l = 0;
    for (float v : hl) l += v;
    *loss_out = l / (double)rows;
  }
  return MCE_OK;
}

int mce_mlp_train(mce_mlp *m, const float *x, const float *y, uint64_t rows, const mce_train_opts *o,
                  double *loss_out) {
  clear_error();
  if (!m || !x || !y || !o || rows == 0) return set_error(MCE_EINVAL, "null argument or no rows");
  if (rows > 0xFFFFFFFFull) return set_error(MCE_ELIMIT, "too many rows");
  if (o->epochs <= 0 || o->batch <= 0 || !(o->lr > 0)) return set_error(MCE_EINVAL, "bad training options");
  if (!model_ok(m)) return set_error(MCE_EMODEL, "model has non-finite values");
  int rc = set_device(o->device);
  if (rc) return rc;
  const uint32_t batch = (uint32_t)std::min<uint64_t>((uint64_t)o->batch, rows);
  const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(64, batch / 256)), waves = blocks * 4;
  DevBuf b;
  float *dx, *dy, *dp, *dpart, *dloss, *dg, *dm, *dv;
  uint32_t *didx;
  if ((rc = b.alloc((void **)&dx, sizeof(float) * rows * kF)) || (rc = b.alloc((void **)&dy, sizeof(float) * rows)) ||
      (rc = b.alloc((void **)&dp, sizeof(float) * (kP + 2 * kF))) || (rc = b.alloc((void **)&didx, 4 * rows)) ||
      (rc = b.alloc((void **)&dpart, sizeof(float) * kP * waves)) || (rc = b.alloc((void **)&dloss, 4 * waves)) ||
      (rc = b.alloc((void **)&dg, sizeof(float) * kP)) || (rc = b.alloc((void **)&dm, sizeof(float) * kP)) ||
      (rc = b.alloc((void **)&dv, sizeof(float) * kP)))
    return rc;
  MCE_HIP(hipMemcpy(dx, x, sizeof(float) * rows * kF, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dy, y, sizeof(float) * rows, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp, m->params, sizeof(float) * kP, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp + kP, m->mu, sizeof(float) * kF, hipMemcpyHostToDevice));
  MCE_HIP(hipMemcpy(dp + kP + kF, m->sigma, sizeof(float) * kF, hipMemcpyHostToDevice));
  MCE_HIP(hipMemset(dm, 0, sizeof(float) * kP));
  MCE_HIP(hipMemset(dv, 0, sizeof(float) * kP));
  std::vector<uint32_t> idx(rows);
  std::iota(idx.begin(), idx.end(), 0u);
  std::mt19937_64 gen(o->seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double last = 0;
  uint64_t step = 0;
  std::vector<float> hl(waves);
  for (int ep = 0; ep < o->epochs; ++ep) {
    std::shuffle(idx.begin(), idx.end(), gen);
    MCE_HIP(hipMemcpy(didx, idx.data(), 4 * rows, hipMemcpyHostToDevice));
    double eloss = 0;
    for (uint64_t s0 = 0; s0 < rows; s0 += batch) {
      const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, rows - s0);
      TrainArgs a{dx, dy, didx + s0, nb, 1.f / (float)nb, dp, dp + kP, dp + kP + kF, dpart, dloss};
      train_rows_kernel<<<blocks, 256>>>(a);
      reduce_partials_kernel<<<(kP + 255) / 256, 256>>>(dpart, waves, dg);
      ++step;
      const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
      adam_kernel<<<(kP + 255) / 256, 256>>>(dp, dg, dm, dv, o->lr, b1, b2, eps, o->weight_decay, c1, c2);
      MCE_HIP(hipMemcpy(hl.data(), dloss, 4 * waves, hipMemcpyDeviceToHost));
      for (float v : hl) eloss += v;
    }
    last = eloss / (double)rows;
  }
  MCE_HIP(hipGetLastError());
  MCE_HIP(hipMemcpy(m->params, dp, sizeof(float) * kP, hipMemcpyDeviceToHost));
  if (loss_out) *loss_out = last;
  if (!model_ok(m)) return set_error(MCE_EMODEL, "training diverged (non-finite weights)");
  return MCE_OK;
}

int mce_mlp_predict(const mce_mlp *m, const float *x, uint64_t rows, int device, float *out) {
  clear_error();
  if (!m || (rows && (!x || !out))) return set_error(MCE_EINVAL, "null argument");
  if (!model_ok(m)) return set_error(MCE_EMODEL, "model has non-finite values");
  int rc = set_device(device);
  if (rc) return rc;
  if (rows == 0) return MCE_OK;
  DevBuf b;
  float *dx, *dout;
  if ((rc = b.alloc((void **)&dx, sizeof(float) * rows * kF)) || (rc = b.alloc((void **)&dout, sizeof(float) * rows)))
    return rc;
  MCE_HIP(hipMemcpy(dx, x, sizeof(float) * rows * kF, hipMemcpyHostToDevice));
  hipStream_t s = nullptr;
  rc = predict_device(*m, dx, rows, dout, s);
  if (rc) return rc;
  MCE_HIP(hipMemcpy(out, dout, sizeof(float) * rows, hipMemcpyDeviceToHost));
  return MCE_OK;
}

int mce_mlp_save(const char *path, const mce_mlp *m) {
  clear_error();
  if (!path || !m) return set_error(MCE_EINVAL, "null argument");
  FILE *fp = fopen(path, "wb");
  if (!fp) return set_error(MCE_EIO, "unable to create %s", path);
  const uint32_t hdr[2] = {(uint32_t)kF, (uint32_t)kP};
  const bool ok = fwrite(kMagic, 1, 8, fp) == 8 && fwrite(hdr, 4, 2, fp) == 2 && fwrite(m, sizeof(*m), 1, fp) == 1;
  if (fclose(fp) != 0 || !ok) return set_error(MCE_EIO, "short write to %s", path);
  return MCE_OK;
}

int mce_mlp_load(const char *path, mce_mlp *m) {
  clear_error();
  if (!path || !m) return set_error(MCE_EINVAL, "null argument");
  FILE *fp = fopen(path, "rb");
  if (!fp) return set_error(MCE_EMODEL, "cost model %s not found", path);
  char magic[8];
  uint32_t hdr[2];
  mce_mlp t;
  const bool ok = fread(magic, 1, 8, fp) == 8 && memcmp(magic, kMagic, 8) == 0 && fread(hdr, 4, 2, fp) == 2 &&
                  hdr[0] == (uint32_t)kF && hdr[1] == (uint32_t)kP && fread(&t, sizeof(t), 1, fp) == 1;
  fclose(fp);
  if (!ok) return set_error(MCE_EMODEL, "cost model %s is malformed", path);
  if (!model_ok(&t)) return set_error(MCE_EMODEL, "cost model %s has non-finite values", path);
  *m = t;
  return MCE_OK;
}

}  // extern "C"
