// Small fp32 MLP kernels for the strategy cost model (no reference
// counterpart): dense layer forward/backward, MSE gradient and an Adam step.
// The model is tiny (5 -> 64 -> 64 -> 2), so these are plain per-element
// kernels with L2-resident weights; deterministic reductions (no atomics except
// the reported loss) keep training reproducible run to run.
#include "common.hpp"

namespace nbg {
namespace {

__device__ __forceinline__ float act_fwd(float z, int act) { return (act == 1 && z < 0.0f) ? 0.0f : z; }

// Y[b, o] = act(sum_i X[b, i] * W[o, i] + bias[o])
__global__ void dense_fwd_kernel(const float* __restrict__ X, const float* __restrict__ W,
                                 const float* __restrict__ bias, float* __restrict__ Y, int64_t B,
                                 int64_t I, int64_t O, int act) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * O) return;
  const int64_t r = idx / O, o = idx % O;
  const float* x = X + r * I;
  const float* w = W + o * I;
  float acc = bias ? bias[o] : 0.0f;
  for (int64_t i = 0; i < I; ++i) acc = fmaf(x[i], w[i], acc);
  Y[idx] = act_fwd(acc, act);
}

// dZ = dY * act'(.) where act' for ReLU is (Y > 0).
__global__ void dense_dz_kernel(const float* __restrict__ Y, const float* __restrict__ dY,
                                float* __restrict__ dZ, int64_t n, int act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dZ[i] = (act == 1 && !(Y[i] > 0.0f)) ? 0.0f : dY[i];
}

// dX[b, i] = sum_o dZ[b, o] * W[o, i]
__global__ void dense_dx_kernel(const float* __restrict__ dZ, const float* __restrict__ W,
                                float* __restrict__ dX, int64_t B, int64_t I, int64_t O) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * I) return;
  const int64_t r = idx / I, i = idx % I;
  float acc = 0.0f;
  for (int64_t o = 0; o < O; ++o) acc = fmaf(dZ[r * O + o], W[o * I + i], acc);
  dX[idx] = acc;
}

// dW[o, i] = sum_b dZ[b, o] * X[b, i]; db[o] = sum_b dZ[b, o]. One wave per
// output weight, lanes striding the batch, then a fixed-order wave reduction.
__global__ void dense_dw_kernel(const float* __restrict__ dZ, const float* __restrict__ X,
                                float* __restrict__ dW, float* __restrict__ db, int64_t B,
                                int64_t I, int64_t O) {
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = lane_id();
  const int64_t nweights = O * I + O;  // dW entries then db entries
  if (wave >= nweights) return;
  float acc = 0.0f;
  if (wave < O * I) {
    const int64_t o = wave / I, i = wave % I;
    for (int64_t r = lane; r < B; r += kWave) acc = fmaf(dZ[r * O + o], X[r * I + i], acc);
  } else {
    const int64_t o = wave - O * I;
    for (int64_t r = lane; r < B; r += kWave) acc += dZ[r * O + o];
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) acc += __shfl_down(acc, d, kWave);
  if (lane == 0) {
    if (wave < O * I) dW[wave] = acc;
    else db[wave - O * I] = acc;
  }
}

// dY = 2 (Y - T) / n; loss += sum (Y - T)^2 / n (reported only).
__global__ void mse_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                float* __restrict__ dY, float* __restrict__ loss, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < n) {
    const float e = Y[i] - T[i];
    dY[i] = 2.0f * e / (float)n;
    e2 = e * e / (float)n;
  }
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) e2 += __shfl_down(e2, d, kWave);
  if (loss && lane_id() == 0 && e2 != 0.0f) atomicAdd(loss, e2);
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int check_dims(const char* name, int64_t B, int64_t I, int64_t O, int act) {
  if (B < 0 || I <= 0 || O <= 0 || I > 4096 || O > 4096)
    return set_error(NBG_ERR_INVALID, "%s: bad dimensions (%lld, %lld, %lld)", name, (long long)B,
                     (long long)I, (long long)O);
  if (act != 0 && act != 1) return set_error(NBG_ERR_INVALID, "%s: act must be 0 or 1", name);
  return NBG_OK;
}

}  // namespace
}  // namespace nbg

extern "C" {

int nbg_dense_forward(const float* x, const float* w, const float* b, float* y, int64_t batch,
                      int64_t in_dim, int64_t out_dim, int act, void* stream) {
  using namespace nbg;
  if (int rc = check_dims("nbg_dense_forward", batch, in_dim, out_dim, act)) return rc;
  if (batch == 0) return NBG_OK;
  if (!x || !w || !y) return set_error(NBG_ERR_INVALID, "nbg_dense_forward: null pointer");
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(blocks_for(batch * out_dim)), dim3(256), 0,
                     (hipStream_t)stream, x, w, b, y, batch, in_dim, out_dim, act);
  return check_hip(hipGetLastError(), "dense_fwd_kernel launch");
}

int nbg_dense_backward(const float* x, const float* w, const float* y, const float* dy, float* dx,
                       float* dw, float* db, float* dz, int64_t batch, int64_t in_dim,
                       int64_t out_dim, int act, void* stream) {
  using namespace nbg;
  if (int rc = check_dims("nbg_dense_backward", batch, in_dim, out_dim, act)) return rc;
  if (!x || !w || !y || !dy || !dw || !db || !dz)
    return set_error(NBG_ERR_INVALID, "nbg_dense_backward: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (batch > 0)
    hipLaunchKernelGGL(dense_dz_kernel, dim3(blocks_for(batch * out_dim)), dim3(256), 0, s, y, dy,
                       dz, batch * out_dim, act);
  if (dx && batch > 0)
    hipLaunchKernelGGL(dense_dx_kernel, dim3(blocks_for(batch * in_dim)), dim3(256), 0, s, dz, w,
                       dx, batch, in_dim, out_dim);
  const int64_t nweights = out_dim * in_dim + out_dim;
  hipLaunchKernelGGL(dense_dw_kernel, dim3(blocks_for(nweights * kWave)), dim3(256), 0, s, dz, x,
                     dw, db, batch, in_dim, out_dim);
  return check_hip(hipGetLastError(), "dense backward launch");
}

int nbg_mse_grad(const float* y, const float* t, float* dy, float* loss, int64_t n, void* stream) {
  using namespace nbg;
  if (n < 0 || !y || !t || !dy) return set_error(NBG_ERR_INVALID, "nbg_mse_grad: bad arguments");
  if (n == 0) return NBG_OK;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, y, t,
                     dy, loss, n);
  return check_hip(hipGetLastError(), "mse_grad_kernel launch");
}

int nbg_adam_step(float* p, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                  float beta2, float eps, int64_t step, void* stream) {
  using namespace nbg;
  if (n < 0 || !p || !g || !m || !v || step < 1)
    return set_error(NBG_ERR_INVALID, "nbg_adam_step: bad arguments");
  if (n == 0) return NBG_OK;
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                     n, lr, beta1, beta2, eps, c1, c2);
  return check_hip(hipGetLastError(), "adam_kernel launch");
}

}  // extern "C"
