// chi_GG'(q,w) = sum_K conj(A[K][G]) d[w][K] A[K][G'] on gfx950 MFMA.
//
// This is the accumulation of chiqw.F:1397-1410 restated as a rank-K update
// per frequency (K = pair*NTK + k).  The complex product is mapped onto real
// MFMA by doubling K: with a = A[K][G], b = d A[K][G'],
//   Re chi += Re a Re b + Im a Im b,   Im chi += Re a Im b - Im a Re b,
// i.e. one real GEMM whose k-slots hold (Re a | Im a) against (Re b | Im b)
// for the real part and (Im b | -Re b) for the imaginary part.
//
// Workgroup: 256 threads (2x2 waves), tile 64 G x 64 G' for one frequency;
// each wave owns 2x2 MFMA blocks of 16x16.  A K-chunk of 32 complex rows of
// both panels is staged in LDS (transposed to [G][K]) together with the 32
// weights; the weights are applied to the right panel as it is read (scale on
// load).  Precisions:
//   FP64   v_mfma_f64_16x16x4f64, FP64 accumulators.
//   FP32   v_mfma_f32_16x16x4f32; FP32 accumulators flushed into FP64 ones
//          every kFlush chunks.
// (The split-BF16 precision has its own kernel, k_chi_bf16x3.)
// The K range may be split over gridDim.z; partial tiles are summed in fixed
// order by k_chi_reduce (deterministic), which also divides by the volume.
#include <algorithm>

#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kTile = 64;   // G and G' per workgroup
constexpr int kKC = 32;     // complex K rows per LDS chunk
constexpr int kLdsRow = kKC + 2;
constexpr int kFlush = 32;  // chunks per FP32 -> FP64 flush (1024 K)

__device__ inline void stage_chunk(const float2* __restrict__ A, size_t lda, size_t K0, int g0,
                                   int h0, const double2* __restrict__ dw, float2 (*sL)[kLdsRow],
                                   float2 (*sR)[kLdsRow], double2* sD) {
  const int t = threadIdx.x;
#pragma unroll
  for (int e = 0; e < (kTile * kKC) / 256; ++e) {
    int idx = t + 256 * e;
    int kk = idx / kTile, gg = idx % kTile;
    const float2* row = A + (K0 + kk) * lda;
    sL[gg][kk] = row[g0 + gg];
    sR[gg][kk] = row[h0 + gg];
  }
  if (t < kKC) sD[t] = dw[K0 + t];
}

// One output tile, precision selected by P.  out: partial[z][w][ng][ng] with
// element (G, G') at out[... + G' * ng + G] (Fortran chiqw(ig,jg,ie) order).
template <int P>
__global__ void __launch_bounds__(256) k_chi_gemm(const float2* __restrict__ A, size_t lda,
                                                  const double2* __restrict__ dw, size_t ldd,
                                                  size_t kchunks_per_split, size_t kchunks,
                                                  int ng, int ntile, int nen,
                                                  double2* __restrict__ out) {
  __shared__ float2 sL[kTile][kLdsRow];
  __shared__ float2 sR[kTile][kLdsRow];
  __shared__ double2 sD[kKC];
  const int tile = blockIdx.x;
  const int tm = tile % ntile, tn = tile / ntile;
  const int g0 = tm * kTile, h0 = tn * kTile;
  const int w = blockIdx.y;
  const int z = blockIdx.z;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, q = lane >> 4;
  const int part = q >> 1;  // 0: real-part k-slots, 1: imaginary-part k-slots
  const double2* dww = dw + (size_t)w * ldd;
  const size_t c_begin = (size_t)z * kchunks_per_split;
  size_t c_end = c_begin + kchunks_per_split;
  if (c_end > kchunks) c_end = kchunks;

  // block validity (skip 16x16 blocks entirely outside [0, ng))
  bool rowok[2], colok[2];
  for (int b = 0; b < 2; ++b) {
    rowok[b] = g0 + wm * 32 + b * 16 < ng;
    colok[b] = h0 + wn * 32 + b * 16 < ng;
  }

  double accd_re[2][2][4], accd_im[2][2][4];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 4; ++v) accd_re[a][b][v] = accd_im[a][b][v] = 0.0;

  if constexpr (P == RP_PREC_FP64) {
    d4v acc_re[2][2], acc_im[2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc_re[a][b] = acc_im[a][b] = d4v{0, 0, 0, 0};
    const int h = q & 1;
    for (size_t c = c_begin; c < c_end; ++c) {
      stage_chunk(A, lda, c * kKC, g0, h0, dww, sL, sR, sD);
      __syncthreads();
#pragma unroll 4
      for (int s = 0; s < 16; ++s) {
        const int kl = h * 16 + s;
        double L[2], Rre[2], Rim[2];
        const double2 dd = sD[kl];
        for (int b = 0; b < 2; ++b) {
          float2 la = sL[wm * 32 + b * 16 + li][kl];
          L[b] = part ? (double)la.y : (double)la.x;
          float2 ra = sR[wn * 32 + b * 16 + li][kl];
          double bx = dd.x * (double)ra.x - dd.y * (double)ra.y;
          double by = dd.x * (double)ra.y + dd.y * (double)ra.x;
          Rre[b] = part ? by : bx;
          Rim[b] = part ? -bx : by;
        }
        for (int a = 0; a < 2; ++a) {
          if (!rowok[a]) continue;
          for (int b = 0; b < 2; ++b) {
            if (!colok[b]) continue;
            acc_re[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(L[a], Rre[b], acc_re[a][b], 0, 0, 0);
            acc_im[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(L[a], Rim[b], acc_im[a][b], 0, 0, 0);
          }
        }
      }
      __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 4; ++v) {
          accd_re[a][b][v] = acc_re[a][b][v];
          accd_im[a][b][v] = acc_im[a][b][v];
        }
  } else {
    static_assert(P == RP_PREC_FP32, "k_chi_gemm: FP64 or FP32");
    f4v acc_re[2][2], acc_im[2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) acc_re[a][b] = acc_im[a][b] = f4v{0, 0, 0, 0};
    const int h = q & 1;
    int since = 0;
    for (size_t c = c_begin; c < c_end; ++c) {
      stage_chunk(A, lda, c * kKC, g0, h0, dww, sL, sR, sD);
      __syncthreads();
#pragma unroll 4
      for (int s = 0; s < 16; ++s) {
        const int kl = h * 16 + s;
        float L[2], Rre[2], Rim[2];
        const float dx = (float)sD[kl].x, dy = (float)sD[kl].y;
        for (int b = 0; b < 2; ++b) {
          float2 la = sL[wm * 32 + b * 16 + li][kl];
          L[b] = part ? la.y : la.x;
          float2 ra = sR[wn * 32 + b * 16 + li][kl];
          float bx = dx * ra.x - dy * ra.y;
          float by = dx * ra.y + dy * ra.x;
          Rre[b] = part ? by : bx;
          Rim[b] = part ? -bx : by;
        }
        for (int a = 0; a < 2; ++a) {
          if (!rowok[a]) continue;
          for (int b = 0; b < 2; ++b) {
            if (!colok[b]) continue;
            acc_re[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(L[a], Rre[b], acc_re[a][b], 0, 0, 0);
            acc_im[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(L[a], Rim[b], acc_im[a][b], 0, 0, 0);
          }
        }
      }
      __syncthreads();
      if (++since == kFlush || c + 1 == c_end) {
        since = 0;
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            for (int v = 0; v < 4; ++v) {
              accd_re[a][b][v] += (double)acc_re[a][b][v];
              accd_im[a][b][v] += (double)acc_im[a][b][v];
            }
            acc_re[a][b] = acc_im[a][b] = f4v{0, 0, 0, 0};
          }
      }
    }
  }

  // epilogue.  Result layout of a 16x16 block (checked on gfx950): the FP32
  // and BF16 MFMAs leave C[4*(lane>>4)+v][lane&15] in element v of a lane, the
  // FP64 MFMA leaves C[(lane>>4)+4*v][lane&15].
  double2* o = out + ((size_t)z * nen + w) * (size_t)ng * ng;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int v = 0; v < 4; ++v) {
        const int row = P == RP_PREC_FP64 ? q + 4 * v : 4 * q + v;
        const int gi = g0 + wm * 32 + a * 16 + row;
        const int hj = h0 + wn * 32 + b * 16 + li;
        if (gi < ng && hj < ng)
          o[(size_t)hj * ng + gi] = make_double2(accd_re[a][b][v], accd_im[a][b][v]);
      }
}

// chi[j] = sum_z partial[z][j] / volume, fixed order over z.
__global__ void k_chi_reduce(const double2* __restrict__ part, size_t n, int nsplit, double inv_vol,
                             double2* __restrict__ chi) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double x = 0.0, y = 0.0;
    for (int z = 0; z < nsplit; ++z) {
      double2 v = part[(size_t)z * n + i];
      x += v.x;
      y += v.y;
    }
    chi[i] = make_double2(x * inv_vol, y * inv_vol);
  }
}

int chi_tile() { return kTile; }
int chi_kchunk() { return kKC; }

void launch_chi_gemm(int prec, const float2* A, size_t lda, const double2* dw, size_t ldd,
                     size_t kpad, int ng, int nen, int nsplit, double2* partial, hipStream_t s) {
  const int ntile = (ng + kTile - 1) / kTile;
  const size_t kchunks = kpad / kKC;
  const size_t per = (kchunks + nsplit - 1) / nsplit;
  dim3 grid(ntile * ntile, nen, nsplit);
  switch (prec) {
    case RP_PREC_FP64:
      hipLaunchKernelGGL(k_chi_gemm<RP_PREC_FP64>, grid, dim3(256), 0, s, A, lda, dw, ldd, per,
                         kchunks, ng, ntile, nen, partial);
      break;
    case RP_PREC_FP32:
      hipLaunchKernelGGL(k_chi_gemm<RP_PREC_FP32>, grid, dim3(256), 0, s, A, lda, dw, ldd, per,
                         kchunks, ng, ntile, nen, partial);
      break;
    default:
      throw Error("k_chi_gemm: FP64 or FP32 only (split-BF16 runs k_chi_bf16x3)");
  }
  RP_HIP(hipGetLastError());
}

// Zero fill by vector stores.  Used instead of hipMemsetAsync wherever the
// launch sequence may be captured into a graph (memset nodes are avoided there).
__global__ void k_zero(uint4* __restrict__ p, size_t n16, unsigned* __restrict__ tail,
                       size_t ntail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0u;
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  RP_CHECK(bytes % 4 == 0 && ((uintptr_t)p % 16) == 0, "zero fill needs 16-byte aligned words");
  const size_t n16 = bytes / 16, ntail = (bytes % 16) / 4;
  int blocks = (int)std::min<size_t>((n16 + 255) / 256, 4096);
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, (uint4*)p, n16,
                     (unsigned*)((char*)p + n16 * 16), ntail);
  RP_HIP(hipGetLastError());
}

void launch_chi_reduce(const double2* partial, size_t n, int nsplit, double inv_vol, double2* chi,
                       hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_chi_reduce, dim3(blocks), dim3(256), 0, s, partial, n, nsplit, inv_vol, chi);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
