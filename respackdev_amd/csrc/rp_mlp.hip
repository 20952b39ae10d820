// Band-tail closure network on gfx950: a small MLP (fully connected layers,
// GELU on the hidden layers, linear output) with BF16 weights and FP32
// activations and accumulation, and its training (mean-squared error, Adam on
// FP32 master weights).
//
// Layout: activations [n][width] row-major FP32; weights W_l [out][in] BF16
// (converted from the FP32 masters before each forward pass); biases FP32.
//   forward : Z_l = A_{l-1} W_l^T + b_l,  A_l = gelu(Z_l) (hidden), Y = Z_L
//   backward: dW_l = dZ_l^T A_{l-1},  db_l = sum_n dZ_l,
//             dA_{l-1} = dZ_l W_l,  dZ_{l-1} = dA_{l-1} * gelu'(Z_{l-1})
// The weight gradient is reduced over samples in fixed chunks (partial sums,
// then a reduction kernel) so results do not depend on launch timing.
#include <algorithm>

#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

namespace {
constexpr int kRows = 32;      // samples per forward / data-gradient block
constexpr int kMaxIn = 128;    // widest layer input
constexpr int kChunk = 256;    // samples per weight-gradient partial
}  // namespace

__device__ __forceinline__ float bf16_round_to_float(unsigned short h) {
  return __builtin_bit_cast(float, (unsigned)h << 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// FP32 master -> BF16 (round to nearest even), n values.
__global__ void k_to_bf16(const float* __restrict__ src, size_t n, unsigned short* __restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    __bf16 h = (__bf16)src[i];
    dst[i] = __builtin_bit_cast(unsigned short, h);
  }
}

// Z = A W^T + b over a block of kRows samples; A_out = gelu(Z) when act.
// grid: ceil(n / kRows); block 256.  nin <= kMaxIn.
__global__ void __launch_bounds__(256) k_linear_fwd(const float* __restrict__ A, int n, int nin,
                                                   const unsigned short* __restrict__ W,
                                                   const float* __restrict__ b, int nout, int act,
                                                   float* __restrict__ Z, float* __restrict__ Aout) {
  __shared__ float sA[kRows][kMaxIn + 1];
  const int r0 = blockIdx.x * kRows;
  for (int e = threadIdx.x; e < kRows * nin; e += blockDim.x) {
    const int r = e / nin, i = e % nin;
    sA[r][i] = r0 + r < n ? A[(size_t)(r0 + r) * nin + i] : 0.0f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kRows * nout; e += blockDim.x) {
    const int r = e / nout, o = e % nout;
    if (r0 + r >= n) continue;
    const unsigned short* w = W + (size_t)o * nin;
    float acc = b[o];
    for (int i = 0; i < nin; ++i) acc = fmaf(sA[r][i], bf16_round_to_float(w[i]), acc);
    const size_t idx = (size_t)(r0 + r) * nout + o;
    Z[idx] = acc;
    if (act) Aout[idx] = gelu(acc);
  }
}

// dA = dZ W (then times gelu'(Zprev) when act): grid ceil(n / kRows); nout <= kMaxIn.
__global__ void __launch_bounds__(256) k_linear_bwd_data(const float* __restrict__ dZ, int n, int nout,
                                                        const unsigned short* __restrict__ W, int nin,
                                                        const float* __restrict__ Zprev, int act,
                                                        float* __restrict__ dA) {
  __shared__ float sD[kRows][kMaxIn + 1];
  const int r0 = blockIdx.x * kRows;
  for (int e = threadIdx.x; e < kRows * nout; e += blockDim.x) {
    const int r = e / nout, o = e % nout;
    sD[r][o] = r0 + r < n ? dZ[(size_t)(r0 + r) * nout + o] : 0.0f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kRows * nin; e += blockDim.x) {
    const int r = e / nin, i = e % nin;
    if (r0 + r >= n) continue;
    float acc = 0.0f;
    for (int o = 0; o < nout; ++o) acc = fmaf(sD[r][o], bf16_round_to_float(W[(size_t)o * nin + i]), acc);
    const size_t idx = (size_t)(r0 + r) * nin + i;
    dA[idx] = act ? acc * gelu_grad(Zprev[idx]) : acc;
  }
}

// Partial weight and bias gradients over one chunk of kChunk samples:
// part[c][o][i] = sum_{n in chunk} dZ[n][o] A[n][i], part[c][o][nin] = sum dZ[n][o].
// grid (nchunks, nout); block 256 (threads i <= nin work).
__global__ void __launch_bounds__(256) k_linear_bwd_w(const float* __restrict__ dZ, const float* __restrict__ A,
                                                     int n, int nin, int nout, float* __restrict__ part) {
  const int c = blockIdx.x, o = blockIdx.y, i = threadIdx.x;
  if (i > nin) return;
  const int s0 = c * kChunk, s1 = min(n, s0 + kChunk);
  float acc = 0.0f;
  for (int s = s0; s < s1; ++s) {
    const float d = dZ[(size_t)s * nout + o];
    acc = fmaf(d, i < nin ? A[(size_t)s * nin + i] : 1.0f, acc);
  }
  part[((size_t)c * nout + o) * (nin + 1) + i] = acc;
}

// grad W[o][i] (and b[o] at i == nin) = sum over chunks, in chunk order.
__global__ void k_reduce_w(const float* __restrict__ part, int nchunks, int nin, int nout,
                           float* __restrict__ gW, float* __restrict__ gb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nout * (nin + 1)) return;
  float acc = 0.0f;
  for (int c = 0; c < nchunks; ++c) acc += part[(size_t)c * nout * (nin + 1) + idx];
  const int o = idx / (nin + 1), i = idx % (nin + 1);
  if (i < nin) gW[(size_t)o * nin + i] = acc;
  else gb[o] = acc;
}

// Mean-squared error: dY = 2 (Y - T) / (n * nout); loss partial sums per block.
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ Y, const float* __restrict__ T, size_t m,
                                            float scale, float* __restrict__ dY, double* __restrict__ lossp) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// Adam on FP32 masters (bias-corrected step sizes a1 = lr / (1 - b1^t), a2 = 1 / (1 - b2^t)).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, size_t n, float lr_t, float b1, float b2, float c2, float eps) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr_t * a / (sqrtf(v * c2) + eps);
  }
}

// Gathers a minibatch: xb[j] = x[idx[j]], tb[j] = t[idx[j]].
__global__ void k_gather_rows(const float* __restrict__ x, const float* __restrict__ t, const int* __restrict__ idx,
                              int nb, int nin, int nout, float* __restrict__ xb, float* __restrict__ tb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const size_t r = (size_t)idx[j];
  for (int i = 0; i < nin; ++i) xb[(size_t)j * nin + i] = x[r * nin + i];
  for (int o = 0; o < nout; ++o) tb[(size_t)j * nout + o] = t[r * nout + o];
}

int mlp_max_width() { return kMaxIn; }
int mlp_chunk() { return kChunk; }

void launch_to_bf16(const float* src, size_t n, unsigned short* dst, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks), dim3(256), 0, s, src, n, dst);
  RP_HIP(hipGetLastError());
}

void launch_linear_fwd(const float* A, int n, int nin, const unsigned short* W, const float* b, int nout,
                       int act, float* Z, float* Aout, hipStream_t s) {
  RP_CHECK(nin <= kMaxIn && nout <= kMaxIn && n > 0, "MLP layer wider than 128 or empty batch");
  hipLaunchKernelGGL(k_linear_fwd, dim3((n + kRows - 1) / kRows), dim3(256), 0, s, A, n, nin, W, b, nout, act,
                     Z, Aout);
  RP_HIP(hipGetLastError());
}

void launch_linear_bwd_data(const float* dZ, int n, int nout, const unsigned short* W, int nin,
                            const float* Zprev, int act, float* dA, hipStream_t s) {
  RP_CHECK(nin <= kMaxIn && nout <= kMaxIn && n > 0, "MLP layer wider than 128 or empty batch");
  hipLaunchKernelGGL(k_linear_bwd_data, dim3((n + kRows - 1) / kRows), dim3(256), 0, s, dZ, n, nout, W, nin,
                     Zprev, act, dA);
  RP_HIP(hipGetLastError());
}

void launch_linear_bwd_w(const float* dZ, const float* A, int n, int nin, int nout, float* part, float* gW,
                         float* gb, hipStream_t s) {
  RP_CHECK(nin <= kMaxIn && nout <= kMaxIn && n > 0, "MLP layer wider than 128 or empty batch");
  const int nchunks = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_linear_bwd_w, dim3(nchunks, nout), dim3(256), 0, s, dZ, A, n, nin, nout, part);
  RP_HIP(hipGetLastError());
  const int m = nout * (nin + 1);
  hipLaunchKernelGGL(k_reduce_w, dim3((m + 255) / 256), dim3(256), 0, s, part, nchunks, nin, nout, gW, gb);
  RP_HIP(hipGetLastError());
}

void launch_mse(const float* Y, const float* T, size_t m, float scale, float* dY, double* lossp, int nblocks,
                hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nblocks), dim3(256), 0, s, Y, T, m, scale, dY, lossp);
  RP_HIP(hipGetLastError());
}

void launch_adam(float* p, const float* g, float* m1, float* m2, size_t n, float lr_t, float b1, float b2,
                 float c2, float eps, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_adam, dim3(blocks), dim3(256), 0, s, p, g, m1, m2, n, lr_t, b1, b2, c2, eps);
  RP_HIP(hipGetLastError());
}

void launch_gather_rows(const float* x, const float* t, const int* idx, int nb, int nin, int nout, float* xb,
                        float* tb, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_rows, dim3((nb + 255) / 256), dim3(256), 0, s, x, t, idx, nb, nin, nout, xb, tb);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
