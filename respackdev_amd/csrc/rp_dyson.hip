// Dyson step of calc_eps_rpa / calc_eps_rpa_q_0 (chiqw_sub.F:386-568) in FP64
// on gfx950: eps_GG' = delta - 2 chi_GG' 4pi/(|q+G||q+G'|) (head, wings and
// the Drude term at q = 0), then eps^-1 for every frequency of the batch.
//
// The inversion is Gauss-Jordan with partial pivoting by |Re|+|Im| (the pivot
// rule of LAPACK's izamax in zgetrf, which the reference calls through
// invZGE.F90:1-23).  Per pivot column one small kernel per matrix picks the
// pivot, swaps rows and scales the pivot row; one batched elementwise kernel
// applies the rank-1 elimination to every matrix; a final kernel undoes the
// row interchanges as column swaps.
#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

// eps[b][j][i] (column-major per matrix), b = dir*nen + w.
__global__ void k_eps_build(const double2* __restrict__ chi, const double* __restrict__ qn, int n,
                            int nen, int q0, int g0, const double* __restrict__ wd,
                            const double2* __restrict__ em, double delt,
                            double2* __restrict__ eps) {
  const int b = blockIdx.y;
  const int j = blockIdx.z;
  const double tpi = 2.0 * 3.141592653589793238462643383279502884;
  const double2* c = chi + ((size_t)b * n + j) * n;
  double2* e = eps + ((size_t)b * n + j) * n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double cx = 2.0 * c[i].x, cy = 2.0 * c[i].y;  // spin factor, chiqw_sub.F:413
    double vx, vy;
    if (!q0) {
      vx = cx * 2.0 * tpi / qn[i] / qn[j];
      vy = cy * 2.0 * tpi / qn[i] / qn[j];
      e[i] = i == j ? make_double2(1.0 - vx, -vy) : make_double2(-vx, -vy);
    } else {
      if (i == j && i == g0) {
        const int dir = b / nen, w = b % nen;
        double om = em[w].x;
        if (om == 0.0) om = 1.0e-5;
        const double d2 = wd[dir] * wd[dir];
        // CMPLX(x,y) without KIND: the Drude term is rounded to single precision
        const float drx = (float)(-d2 / (om * om + delt * delt));
        const float dry = (float)(delt * d2 / (om * om * om + om * delt * delt));
        e[i] = make_double2(1.0 - cx * 2.0 * tpi + (double)drx, -cy * 2.0 * tpi + (double)dry);
      } else if (i == j) {
        vx = cx * 2.0 * tpi / qn[i] / qn[j];
        vy = cy * 2.0 * tpi / qn[i] / qn[j];
        e[i] = make_double2(1.0 - vx, -vy);
      } else if (i == g0) {
        e[i] = make_double2(-cx * 2.0 * tpi / qn[j], -cy * 2.0 * tpi / qn[j]);
      } else if (j == g0) {
        e[i] = make_double2(-cx * 2.0 * tpi / qn[i], -cy * 2.0 * tpi / qn[i]);
      } else {
        e[i] = make_double2(-cx * 2.0 * tpi / qn[i] / qn[j], -cy * 2.0 * tpi / qn[i] / qn[j]);
      }
    }
  }
}

__device__ inline double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }

// One workgroup per matrix: pivot for column c, row swap, pivot-row scaling.
__global__ void __launch_bounds__(256) k_gj_pivot(double2* __restrict__ a, int n, int c,
                                                  int* __restrict__ ipiv, double2* __restrict__ rowc,
                                                  double2* __restrict__ colc,
                                                  int* __restrict__ status) {
  const int b = blockIdx.x;
  double2* m = a + (size_t)b * n * n;
  __shared__ double sv[256];
  __shared__ int si[256];
  double best = -1.0;
  int bi = n;
  for (int i = c + threadIdx.x; i < n; i += blockDim.x) {
    double v = cabs1(m[i + (size_t)n * c]);
    if (v > best) { best = v; bi = i; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      double v2 = sv[threadIdx.x + w];
      int i2 = si[threadIdx.x + w];
      if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const int p = si[0];
  if (threadIdx.x == 0) ipiv[(size_t)b * n + c] = p;
  if (p != c)
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      double2 t = m[c + (size_t)n * j];
      m[c + (size_t)n * j] = m[p + (size_t)n * j];
      m[p + (size_t)n * j] = t;
    }
  __syncthreads();
  const double2 piv = m[c + (size_t)n * c];
  if (piv.x == 0.0 && piv.y == 0.0) {
    if (threadIdx.x == 0) atomicOr(status, 4);
    return;
  }
  const zd pinv = mkz(1.0, 0.0) / mkz(piv.x, piv.y);
  double2* rc = rowc + (size_t)b * n;
  double2* cc = colc + (size_t)b * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    zd v = j == c ? pinv : mkz(m[c + (size_t)n * j].x, m[c + (size_t)n * j].y) * pinv;
    rc[j] = make_double2(v.x, v.y);
    m[c + (size_t)n * j] = rc[j];
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    cc[i] = i == c ? make_double2(0.0, 0.0) : m[i + (size_t)n * c];
}

// a[i][j] -= colc[i] rowc[j] for i != c; a[i][c] = -colc[i] rowc[c].
// grid: (ceil(n/256), n, nbatch)
__global__ void __launch_bounds__(256) k_gj_update(double2* __restrict__ a, int n, int c,
                                                   const double2* __restrict__ rowc,
                                                   const double2* __restrict__ colc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int b = blockIdx.z;
  if (i >= n || i == c) return;
  const double2 f = colc[(size_t)b * n + i];
  const double2 r = rowc[(size_t)b * n + j];
  double2* e = a + ((size_t)b * n + j) * n + i;
  if (j == c) {
    *e = make_double2(-(f.x * r.x - f.y * r.y), -(f.x * r.y + f.y * r.x));
  } else {
    double2 v = *e;
    *e = make_double2(v.x - (f.x * r.x - f.y * r.y), v.y - (f.x * r.y + f.y * r.x));
  }
}

// Undo the row interchanges: swap columns c and ipiv[c] for c = n-1 .. 0.
__global__ void __launch_bounds__(256) k_gj_colswap(double2* __restrict__ a, int n,
                                                    const int* __restrict__ ipiv) {
  const int b = blockIdx.x;
  double2* m = a + (size_t)b * n * n;
  for (int c = n - 1; c >= 0; --c) {
    const int p = ipiv[(size_t)b * n + c];
    if (p != c)
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double2 t = m[i + (size_t)n * c];
        m[i + (size_t)n * c] = m[i + (size_t)n * p];
        m[i + (size_t)n * p] = t;
      }
    __syncthreads();
  }
}

void launch_eps_build(const double2* chi, const double* qn, int n, int nbatch, int nen, int q0,
                      int g0, const double* wd, const double2* em, double delt, double2* eps,
                      hipStream_t s) {
  dim3 grid((n + 255) / 256, nbatch, n);
  hipLaunchKernelGGL(k_eps_build, grid, dim3(256), 0, s, chi, qn, n, nen, q0, g0, wd, em, delt, eps);
  RP_HIP(hipGetLastError());
}

void launch_gauss_jordan(double2* a, int n, int nbatch, int* ipiv, double2* rowc, double2* colc,
                         int* status, hipStream_t s) {
  for (int c = 0; c < n; ++c) {
    hipLaunchKernelGGL(k_gj_pivot, dim3(nbatch), dim3(256), 0, s, a, n, c, ipiv, rowc, colc, status);
    hipLaunchKernelGGL(k_gj_update, dim3((n + 255) / 256, n, nbatch), dim3(256), 0, s, a, n, c,
                       (const double2*)rowc, (const double2*)colc);
  }
  hipLaunchKernelGGL(k_gj_colswap, dim3(nbatch), dim3(256), 0, s, a, n, (const int*)ipiv);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
