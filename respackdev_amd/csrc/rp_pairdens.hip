// Pair densities of calc_InterStateMatrix (chiqw_sub.F:97-205) on gfx950.
//
// Stage 1 (once per system): every band at every full-mesh k is scattered onto
// the nwx*nwy*nwz grid and transformed with e^{+i} (fft3_bw, unscaled); the
// result is rounded to single precision exactly as the reference's CMPLX does
// (chiqw_sub.F:143,162) and kept in HBM as complex64: ur[band][k][r].
//
// Stage 2 (per q): for every (pair, k) the product conj(u_kq(r) e^{-i 2pi
// r.dG/N}) u_k(r) is formed in FP64, transformed with e^{-i} and 1/N
// (fft3_fw), read at -G for the chi sphere and rounded to single precision
// (chiqw_sub.F:192-202): A[K][lda] complex64, K = pair*NTK + k.
//
// Both transforms are LDS-resident where the FP64 grid fits one workgroup's
// LDS: each axis is done in place as independent per-thread line DFTs (the
// axis lengths are the reference's 2,3,5-smooth sizes, <= 32), so HBM sees one
// read of the inputs and one write of the outputs per transform.  The forward
// transform of stage 2 is pruned to the outputs the sphere needs (below), so
// it holds only the z planes with sphere points.  Stage 1 on grids too large
// for the LDS (e.g. 24^3 = 13824 points) runs as three passes of line DFTs
// over HBM (k_line_dft).
#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "rp_common.h"
#include "rp_fft.h"
#include "rp_kernels.h"

namespace rp {

// LDS a gfx950 workgroup may use; the FP64 grid of an LDS-resident transform
// must fit in it (10240 points).
constexpr size_t kLdsBytes = 160 * 1024;

// Twiddle tables in HBM for every line length N = 1..32, one row per output
// index: T_N[k][j] = w^{(jk) mod N} with w = exp(-2 pi i/N) (forward table) or
// its conjugate (backward table), at tab[dir*kTwHalf + off(N) + k*N + j].  A line
// DFT reads row k with offsets fixed at compile time from a wave-uniform base,
// so the twiddles arrive through scalar loads into SGPRs: no LDS traffic, no
// vector instructions and no per-term index arithmetic.
constexpr int kTwMax = 32;
constexpr int kTwHalf = (kTwMax - 1) * kTwMax * (2 * kTwMax - 1) / 6 + kTwMax * kTwMax;
__host__ __device__ constexpr int tw_off(int n) { return (n - 1) * n * (2 * n - 1) / 6; }
__host__ __device__ __forceinline__ const double2* twiddles_of(const double2* tab, int n, int sign) {
  return tab + (sign > 0 ? kTwHalf : 0) + tw_off(n);
}
__global__ void k_fill_twiddles(double2* __restrict__ tab) {
  for (int n = 1; n <= kTwMax; ++n)
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
      const int k = e / n, j = e % n, m = (j * k) % n;
      double s, c;
      sincospi(2.0 * (double)m / (double)n, &s, &c);
      tab[tw_off(n) + e] = make_double2(c, -s);
      tab[kTwHalf + tw_off(n) + e] = make_double2(c, s);
    }
}


// Length-N DFT of the lines of one axis, in place in LDS: out[k] = sum_j v[j]
// T[k][j] with T the forward or backward twiddle table of N.  Each thread
// transforms whole lines (loaded into registers first); the caller
// synchronises between axes.
template <int N>
__device__ __attribute__((always_inline)) inline void dft_axis_n(double2* g, int axis, int nx, int ny,
                                                                 int nz, const double2* tw) {
  int nlines, stride;
  if (axis == 0) { nlines = ny * nz; stride = 1; }
  else if (axis == 1) { nlines = nx * nz; stride = nx; }
  else { nlines = nx * ny; stride = nx * ny; }
  for (int l0 = 0; l0 < nlines; l0 += blockDim.x) {
    const int l = l0 + (int)threadIdx.x;
    const bool mine = l < nlines;
    int base = 0;
    if (axis == 0) base = nx * l;
    else if (axis == 1) base = (l % nx) + nx * ny * (l / nx);
    else base = l;
    // the lines of one axis are disjoint: each thread reads and rewrites its own
    if (mine) {
      double2 v[N];
#pragma unroll
      for (int j = 0; j < N; ++j) v[j] = g[base + j * stride];
#pragma unroll 1
      for (int k = 0; k < N; ++k) {
        const double2* row = tw + k * N;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double2 w = row[j];
          re = fma(v[j].x, w.x, re);
          re = fma(-v[j].y, w.y, re);
          im = fma(v[j].x, w.y, im);
          im = fma(v[j].y, w.x, im);
        }
        g[base + k * stride] = make_double2(re, im);
      }
    }
  }
}

// MAXN bounds the instantiated line lengths so that kernels for small grids
// keep a small register footprint.
template <int MAXN>
__device__ __attribute__((always_inline)) inline void dft_axis(double2* g, int axis, int n, int nx,
                                                               int ny, int nz, int sign,
                                                               const double2* tab) {
  switch (n) {
#define RP_CASE(N) \
  case N: if constexpr (N <= MAXN) dft_axis_n<N>(g, axis, nx, ny, nz, twiddles_of(tab, N, sign)); break;
    RP_CASE(1) RP_CASE(2) RP_CASE(3) RP_CASE(4) RP_CASE(5) RP_CASE(6) RP_CASE(8)
    RP_CASE(9) RP_CASE(10) RP_CASE(12) RP_CASE(15) RP_CASE(16) RP_CASE(18) RP_CASE(20)
    RP_CASE(24) RP_CASE(25) RP_CASE(27) RP_CASE(30) RP_CASE(32)
#undef RP_CASE
    default: break;  // rejected on the host before launch
  }
}

bool dft_size_supported(int n) {
  switch (n) {
    case 1: case 2: case 3: case 4: case 5: case 6: case 8: case 9: case 10: case 12:
    case 15: case 16: case 18: case 20: case 24: case 25: case 27: case 30: case 32:
      return true;
    default:
      return false;
  }
}
size_t lds_budget_bytes() { return kLdsBytes; }

// z, then y, then x (the order of fft3_fw/fft3_bw, m_fft3d_20150826.f90:1598-
// 1720), in place.  Ends synchronised.
template <int MAXN>
__device__ __attribute__((always_inline)) inline void dft3(double2* g, int nx, int ny, int nz,
                                                           int sign, const double2* tab) {
  dft_axis<MAXN>(g, 2, nz, nx, ny, nz, sign, tab);
  __syncthreads();
  dft_axis<MAXN>(g, 1, ny, nx, ny, nz, sign, tab);
  __syncthreads();
  dft_axis<MAXN>(g, 0, nx, nx, ny, nz, sign, tab);
  __syncthreads();
}

__device__ inline int wrap(int g, int n) { return ((g % n) + n) % n; }

// grid: (ntk, nband).  coef[band][k][ntg] complex128; kg[k][ntg][3]; ng0[k].
template <int MAXN>
__global__ void __launch_bounds__(256) k_band_grids(const double2* __restrict__ coef,
                                                    const int* __restrict__ kg,
                                                    const int* __restrict__ ng0, int ntk,
                                                    int ntg, int nx, int ny, int nz,
                                                    const double2* __restrict__ twtab,
                                                    float2* __restrict__ ur) {
  extern __shared__ double2 lds[];
  const int nr = nx * ny * nz;
  double2* g = lds;
  const int k = blockIdx.x, b = blockIdx.y;
  for (int r = threadIdx.x; r < nr; r += blockDim.x) g[r] = make_double2(0.0, 0.0);
  __syncthreads();
  const double2* c = coef + ((size_t)b * ntk + k) * ntg;
  const int* kk = kg + (size_t)k * ntg * 3;
  const int n = ng0[k];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int r = wrap(kk[3 * i], nx) + nx * (wrap(kk[3 * i + 1], ny) + ny * wrap(kk[3 * i + 2], nz));
    g[r] = c[i];
  }
  __syncthreads();
  dft3<MAXN>(g, nx, ny, nz, +1, twtab);
  float2* out = ur + ((size_t)b * ntk + k) * nr;
  for (int r = threadIdx.x; r < nr; r += blockDim.x)
    out[r] = make_float2((float)g[r].x, (float)g[r].y);  // CMPLX(): single precision
}

// ------------------------------------------ grids beyond the LDS: HBM passes
// In place along one axis of nb grids g[b][r] (complex FP64): every line is
// one thread's, its N points in registers; sign as the twiddle table tw.
template <int N>
__global__ void __launch_bounds__(256) k_line_dft(double2* __restrict__ g, long long nlines, int axis,
                                                  int nx, int ny, int nz,
                                                  const double2* __restrict__ tw) {
  const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (idx >= nlines) return;
  const int nr = nx * ny * nz, per = nr / N;
  const long long b = idx / per;
  const int l = (int)(idx % per);
  int base, stride;
  if (axis == 0) { base = nx * l; stride = 1; }
  else if (axis == 1) { base = (l % nx) + nx * ny * (l / nx); stride = nx; }
  else { base = l; stride = nx * ny; }
  double2* p = g + (size_t)b * nr + base;
  double2 v[N];
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = p[j * stride];
  if constexpr (N <= 16) {
    double2 y[N];
    fft_reg<N, 1, N>(v, y, tw + N);  // row 1 of the table: w^j
#pragma unroll
    for (int k = 0; k < N; ++k) p[k * stride] = y[k];
  } else {
#pragma unroll 1
    for (int k = 0; k < N; ++k) {
      const double2* row = tw + k * N;
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double2 w = row[j];
        re = fma(v[j].x, w.x, re);
        re = fma(-v[j].y, w.y, re);
        im = fma(v[j].x, w.y, im);
        im = fma(v[j].y, w.x, im);
      }
      p[k * stride] = make_double2(re, im);
    }
  }
}

// Coefficients of the grids (b, k) = gb0 .. gb0+nb-1 (flattened b*ntk + k)
// scattered onto zeroed grids g[gb - gb0][r].  grid (ceil(ntg/256), nb).
__global__ void __launch_bounds__(256) k_scatter_coef(const double2* __restrict__ coef,
                                                      const int* __restrict__ kg,
                                                      const int* __restrict__ ng0, int ntk, int ntg,
                                                      int nx, int ny, int nz, long long gb0,
                                                      double2* __restrict__ g) {
  const long long gb = gb0 + blockIdx.y;
  const int k = (int)(gb % ntk);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ng0[k]) return;
  const int* kk = kg + (size_t)k * ntg * 3;
  const int r = wrap(kk[3 * i], nx) + nx * (wrap(kk[3 * i + 1], ny) + ny * wrap(kk[3 * i + 2], nz));
  g[(size_t)blockIdx.y * nx * ny * nz + r] = coef[(size_t)gb * ntg + i];
}

// CMPLX(): the FP64 grids rounded to single precision into ur.
__global__ void k_round_grids(const double2* __restrict__ g, size_t n, float2* __restrict__ ur) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    ur[i] = make_float2((float)g[i].x, (float)g[i].y);
}

// ----------------------------------------------- pruned forward transform
// Only the chi sphere's points of the forward transform are read, so each axis
// stores only the outputs that a later axis or the final read needs:
//   z axis: straight from the band grids in registers (no LDS pass for the
//           product), only the z planes that hold sphere points, stored
//           compacted: g[x + nx (y + ny iz)], iz the plane's rank;
//   y axis: the lines (x, iz) of those planes, only the y rows that hold
//           sphere points in plane iz;
//   x axis: only the lines (y, iz) that hold sphere points, only their x.
// The plan (per q, built on the host from the sphere): [0] nzz, [1] nxl,
// [2] zmask, [3] 0, ymask[nzz], xline[nxl] = y | iz << 8, xmask[nxl],
// pos[ng] = position of sphere point i in the compacted grid.
std::vector<int> pairdens_plan(const std::vector<int>& lg, int nx, int ny, int nz) {
  const int ng = (int)(lg.size() / 3);
  std::vector<int> px(ng), py(ng), pz(ng);
  unsigned zmask = 0;
  for (int i = 0; i < ng; ++i) {
    px[i] = ((-lg[3 * i]) % nx + nx) % nx;
    py[i] = ((-lg[3 * i + 1]) % ny + ny) % ny;
    pz[i] = ((-lg[3 * i + 2]) % nz + nz) % nz;
    zmask |= 1u << pz[i];
  }
  std::vector<int> izof(nz, -1);
  int nzz = 0;
  for (int z = 0; z < nz; ++z)
    if ((zmask >> z) & 1u) izof[z] = nzz++;
  std::vector<unsigned> ymask(nzz, 0u), xm((size_t)ny * nzz, 0u);
  for (int i = 0; i < ng; ++i) {
    const int iz = izof[pz[i]];
    ymask[iz] |= 1u << py[i];
    xm[(size_t)iz * ny + py[i]] |= 1u << px[i];
  }
  std::vector<int> xline, xmask;
  for (int iz = 0; iz < nzz; ++iz)
    for (int y = 0; y < ny; ++y)
      if (xm[(size_t)iz * ny + y]) {
        xline.push_back(y | (iz << 8));
        xmask.push_back((int)xm[(size_t)iz * ny + y]);
      }
  const int nxl = (int)xline.size();
  std::vector<int> plan;
  plan.reserve(4 + nzz + 2 * nxl + ng);
  plan.push_back(nzz);
  plan.push_back(nxl);
  plan.push_back((int)zmask);
  plan.push_back(0);
  for (int iz = 0; iz < nzz; ++iz) plan.push_back((int)ymask[iz]);
  plan.insert(plan.end(), xline.begin(), xline.end());
  plan.insert(plan.end(), xmask.begin(), xmask.end());
  for (int i = 0; i < ng; ++i) plan.push_back(px[i] + nx * (py[i] + ny * izof[pz[i]]));
  return plan;
}

// Outputs k of mask of one line v (twiddle rows of tw, wave-uniform), to
// dst[k * stride] (z axis: the compacted plane of k).  Lines of up to 16
// points take the register FFT (all N outputs at about a quarter of the
// multiply-adds of the masked direct sum), longer ones the direct sum of the
// masked outputs (the FFT's registers would spill there).
template <int N, bool ZPLANES>
__device__ __forceinline__ void masked_line(const double2 (&v)[N], unsigned mask, const double2* tw,
                                            double2* dst, int stride) {
  if constexpr (N <= 16) {
    double2 y[N];
    fft_reg<N, 1, N>(v, y, tw + N);  // row 1 of the table: w^j
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (!((mask >> k) & 1u)) continue;
      const int slot = ZPLANES ? __builtin_popcount(mask & ((1u << k) - 1u)) : k;
      dst[slot * stride] = y[k];
    }
    return;
  }
#pragma unroll 1
  for (int k = 0; k < N; ++k) {
    if (!((mask >> k) & 1u)) continue;
    const double2* row = tw + k * N;
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double2 w = row[j];
      re = fma(v[j].x, w.x, re);
      re = fma(-v[j].y, w.y, re);
      im = fma(v[j].x, w.y, im);
      im = fma(v[j].y, w.x, im);
    }
    const int slot = ZPLANES ? __builtin_popcount(mask & ((1u << k) - 1u)) : k;
    dst[slot * stride] = make_double2(re, im);
  }
}

template <int N>
__device__ __forceinline__ void pd_zaxis(const float2* __restrict__ uk, const float2* __restrict__ ukq,
                                         const double2* ph, int nx, int ny, unsigned zmask,
                                         const double2* tw, double2* g) {
  const int nl = nx * ny;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    double2 v[N];
    if (ph) {
      const int r1 = l % nx, r2 = l / nx;
      const double2 px = ph[r1], py = ph[nx + r2];
      const double qx = px.x * py.x - px.y * py.y, qy = px.x * py.y + px.y * py.x;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double2 pz = ph[nx + ny + j];
        const double fx = qx * pz.x - qy * pz.y, fy = qx * pz.y + qy * pz.x;
        const float2 a = uk[l + nl * j], b = ukq[l + nl * j];
        const double bx = (double)b.x * fx - (double)b.y * fy;
        const double by = (double)b.x * fy + (double)b.y * fx;
        const double ax2 = a.x, ay2 = a.y;
        v[j] = make_double2(bx * ax2 + by * ay2, bx * ay2 - by * ax2);
      }
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const float2 a = uk[l + nl * j], b = ukq[l + nl * j];
        const double bx = b.x, by = b.y, ax = a.x, ay = a.y;
        v[j] = make_double2(bx * ax + by * ay, bx * ay - by * ax);
      }
    }
    masked_line<N, true>(v, zmask, tw, g + l, nl);
  }
}

template <int N>
__device__ __forceinline__ void pd_yaxis(double2* g, int nx, int nzz, const int* __restrict__ ymask,
                                         const double2* tw) {
  const int nl = nx * nzz;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    const int x = l % nx, iz = l / nx;
    double2* base = g + x + nx * N * iz;
    double2 v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = base[j * nx];
    masked_line<N, false>(v, (unsigned)ymask[iz], tw, base, nx);
  }
}

template <int N>
__device__ __forceinline__ void pd_xaxis(double2* g, int ny, int nxl, const int* __restrict__ xline,
                                         const int* __restrict__ xmask, const double2* tw) {
  for (int m = threadIdx.x; m < nxl; m += blockDim.x) {
    const int y = xline[m] & 255, iz = xline[m] >> 8;
    double2* base = g + N * (y + ny * iz);
    double2 v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = base[j];
    masked_line<N, false>(v, (unsigned)xmask[m], tw, base, 1);
  }
}

#define RP_PD_SWITCH(n, MAXN, CALL)                                                            \
  switch (n) {                                                                                 \
    RP_PD_CASE(1, MAXN, CALL) RP_PD_CASE(2, MAXN, CALL) RP_PD_CASE(3, MAXN, CALL)              \
    RP_PD_CASE(4, MAXN, CALL) RP_PD_CASE(5, MAXN, CALL) RP_PD_CASE(6, MAXN, CALL)              \
    RP_PD_CASE(8, MAXN, CALL) RP_PD_CASE(9, MAXN, CALL) RP_PD_CASE(10, MAXN, CALL)             \
    RP_PD_CASE(12, MAXN, CALL) RP_PD_CASE(15, MAXN, CALL) RP_PD_CASE(16, MAXN, CALL)           \
    RP_PD_CASE(18, MAXN, CALL) RP_PD_CASE(20, MAXN, CALL) RP_PD_CASE(24, MAXN, CALL)           \
    RP_PD_CASE(25, MAXN, CALL) RP_PD_CASE(27, MAXN, CALL) RP_PD_CASE(30, MAXN, CALL)           \
    RP_PD_CASE(32, MAXN, CALL)                                                                 \
    default: break;                                                                            \
  }
#define RP_PD_CASE(NN, MAXN, CALL) \
  case NN: {                       \
    constexpr int N = NN;          \
    if constexpr (N <= MAXN) CALL; \
  } break;

// grid: K = npair*ntk blocks (flattened).  u_occ = ur of band pocc[p] at k,
// u_vir = ur of band pvir[p] at ikq[k] (both index the same band array), the
// product conj(u_vir(r) e^{-i 2pi r.dG/N}) u_occ(r) in FP64, transformed with
// e^{-i} and 1/N and read at -G for the chi sphere, rounded to single precision
// (chiqw_sub.F:192-202): A[K][lda] complex64.
template <int MAXN>
__global__ void __launch_bounds__(256) k_pair_densities_pruned(
    const float2* __restrict__ ur, int ntk, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const int* __restrict__ shift,
    int nx, int ny, int nz, const int* __restrict__ plan, int ng, size_t lda, int k_offset,
    const double2* __restrict__ twtab, float2* __restrict__ A) {
  extern __shared__ double2 lds[];
  const int nzz = plan[0], nxl = plan[1];
  const unsigned zmask = (unsigned)plan[2];
  const int* ymask = plan + 4;
  const int* xline = ymask + nzz;
  const int* xmask = xline + nxl;
  const int* pos = xmask + nxl;
  double2* g = lds;                           // [nzz][ny][nx]
  double2* ph = g + (size_t)nx * ny * nzz;    // umklapp phase factors per axis
  const int K = blockIdx.x + k_offset;
  const int p = K / ntk, k = K % ntk;
  const int kq = ikq[k];
  const int nr = nx * ny * nz;
  const float2* uk = ur + ((size_t)pocc[p] * ntk + k) * nr;
  const float2* ukq = ur + ((size_t)pvir[p] * ntk + kq) * nr;
  const int d1 = shift[3 * k], d2 = shift[3 * k + 1], d3 = shift[3 * k + 2];
  const bool umklapp = d1 != 0 || d2 != 0 || d3 != 0;  // uniform over the block
  if (umklapp) {
    // exp(-i 2 pi r.dG/N) = phx(r1) phy(r2) phz(r3) (chiqw_sub.F:165-185)
    const double tpi = 2.0 * 3.141592653589793238462643383279502884;
    for (int i = threadIdx.x; i < nx + ny + nz; i += blockDim.x) {
      const int ax = i < nx ? 0 : (i < nx + ny ? 1 : 2);
      const int ri = ax == 0 ? i : (ax == 1 ? i - nx : i - nx - ny);
      const int n = ax == 0 ? nx : (ax == 1 ? ny : nz);
      const int dg = ax == 0 ? d1 : (ax == 1 ? d2 : d3);
      double sn, cs;
      sincos(tpi * ((double)ri * (double)dg / (double)n), &sn, &cs);
      ph[i] = make_double2(cs, -sn);
    }
    __syncthreads();
  }
  const double2* phz = umklapp ? ph : nullptr;
  // z, then y, then x (the order of fft3_fw, m_fft3d_20150826.f90:1598-1720)
  RP_PD_SWITCH(nz, MAXN, (pd_zaxis<N>(uk, ukq, phz, nx, ny, zmask, twiddles_of(twtab, N, -1), g)))
  __syncthreads();
  RP_PD_SWITCH(ny, MAXN, (pd_yaxis<N>(g, nx, nzz, ymask, twiddles_of(twtab, N, -1))))
  __syncthreads();
  RP_PD_SWITCH(nx, MAXN, (pd_xaxis<N>(g, ny, nxl, xline, xmask, twiddles_of(twtab, N, -1))))
  __syncthreads();
  const double scl = 1.0 / (double)nr;
  float2* out = A + (size_t)K * lda;
  for (int i = threadIdx.x; i < (int)lda; i += blockDim.x) {
    if (i < ng) {
      const double2 v = g[pos[i]];
      out[i] = make_float2((float)(v.x * scl), (float)(v.y * scl));
    } else {
      out[i] = make_float2(0.f, 0.f);
    }
  }
}
// ------------------------------------- Wannier pair densities summed over k
// calc_int needs the pair densities only through their k sums
//   rho_p(q+G) = (1/NTK) sum_k w_k M_p,k(G) / |q+G|
// (calc_intW.F:647-767 with w_k = 1; calc_intJ.F:618-720 with
// w_k = e^{-i 2pi k.R}), and M_p,k(G) is the forward transform of the product
// conj(u_iw,k+q(r) e^{-i 2pi r.dG_k/N}) u_jw,k(r).  The transform is linear, so
// the products are summed over k on the real-space grid first (FP64), and one
// pruned transform per (q, pair) replaces NTK of them.  The only arithmetic the
// reference does differently is the single-precision rounding of each M_k
// before the sum (CMPLX, calc_int_sub.F:214): about 1e-8 relative in rho.
//
// k_pair_rsum, grid (ceil(nr/256), nq, ngroups), 256 threads, one grid point r
// per thread:
//   FULL = false: pairs p = w (iw = jw = w), w < nwf <= NI, one group;
//   FULL = true:  pairs p = iw + nwf*jw, group z takes jw in {2z, 2z + 1}.
// NI (8 or 16) bounds the Wannier functions held per thread.
// P[q][p][r] complex FP64.  ikq[q][k], shift[q][k][3] (the umklapp vectors),
// wk[k] (may be null: 1).
constexpr int kRsumJB = 2;
template <bool FULL, int NI>
__global__ void __launch_bounds__(256) k_pair_rsum(const float2* __restrict__ ur, int ntk, int nwf,
                                                   int nx, int ny, int nz, const int* __restrict__ ikq,
                                                   const int* __restrict__ shift,
                                                   const double2* __restrict__ wk,
                                                   double2* __restrict__ P) {
  __shared__ double2 tab[3 * kTwMax];  // e^{-i 2pi m/n} per axis (n = nx, ny, nz)
  const double tpi = 2.0 * 3.141592653589793238462643383279502884;
  for (int i = threadIdx.x; i < nx + ny + nz; i += blockDim.x) {
    const int ax = i < nx ? 0 : (i < nx + ny ? 1 : 2);
    const int m = ax == 0 ? i : (ax == 1 ? i - nx : i - nx - ny);
    const int n = ax == 0 ? nx : (ax == 1 ? ny : nz);
    double sn, cs;
    sincos(tpi * ((double)m / (double)n), &sn, &cs);
    tab[i] = make_double2(cs, -sn);
  }
  __syncthreads();
  const int nr = nx * ny * nz;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y, z = blockIdx.z;
  const int npair = FULL ? nwf * nwf : nwf;
  if (r >= nr) return;
  const int x = r % nx, y = (r / nx) % ny, zz = r / (nx * ny);
  constexpr int NA = FULL ? kRsumJB : NI;  // accumulator rows (jw, or w on the diagonal)
  double2 acc[NA][FULL ? NI : 1];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < (FULL ? NI : 1); ++b) acc[a][b] = make_double2(0.0, 0.0);
  const int* kq_of = ikq + (size_t)q * ntk;
  const int* sh_of = shift + (size_t)3 * q * ntk;
  for (int k = 0; k < ntk; ++k) {
    const int kq = kq_of[k];
    const int d1 = sh_of[3 * k], d2 = sh_of[3 * k + 1], d3 = sh_of[3 * k + 2];
    double2 ph = make_double2(1.0, 0.0);
    if (d1 != 0 || d2 != 0 || d3 != 0) {  // uniform over the grid
      const double2 px = tab[((d1 * x) % nx + nx) % nx];
      const double2 py = tab[nx + ((d2 * y) % ny + ny) % ny];
      const double2 pz = tab[nx + ny + ((d3 * zz) % nz + nz) % nz];
      const double2 t = make_double2(px.x * py.x - px.y * py.y, px.x * py.y + px.y * py.x);
      ph = make_double2(t.x * pz.x - t.y * pz.y, t.x * pz.y + t.y * pz.x);
    }
    const double2 wv = wk ? wk[k] : make_double2(1.0, 0.0);
    // b_iw = u_iw,k+q(r) e^{-i 2pi r.dG/N} (FP64 product of the single-precision grid)
    double2 bq[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      bq[i] = make_double2(0.0, 0.0);
      if (i < nwf) {
        const float2 v = ur[((size_t)i * ntk + kq) * nr + r];
        bq[i] = make_double2((double)v.x * ph.x - (double)v.y * ph.y,
                             (double)v.x * ph.y + (double)v.y * ph.x);
      }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const int jw = FULL ? z * kRsumJB + a : a;
      if (jw >= nwf) continue;
      const float2 v = ur[((size_t)jw * ntk + k) * nr + r];
      // w_k u_jw,k(r)
      const double2 aw = make_double2((double)v.x * wv.x - (double)v.y * wv.y,
                                      (double)v.x * wv.y + (double)v.y * wv.x);
      if constexpr (FULL) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {  // conj(b_iw) (w_k a_jw)
          if (i >= nwf) break;  // uniform
          acc[a][i].x += bq[i].x * aw.x + bq[i].y * aw.y;
          acc[a][i].y += bq[i].x * aw.y - bq[i].y * aw.x;
        }
      } else {
        const double2 b = bq[a];  // a is a compile-time index here (unrolled)
        acc[a][0].x += b.x * aw.x + b.y * aw.y;
        acc[a][0].y += b.x * aw.y - b.y * aw.x;
      }
    }
  }
  double2* out = P + (size_t)q * npair * nr + r;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int jw = FULL ? z * kRsumJB + a : a;
    if (jw >= nwf) continue;
    if constexpr (FULL) {
#pragma unroll
      for (int i = 0; i < NI; ++i)
        if (i < nwf) out[(size_t)(i + nwf * jw) * nr] = acc[a][i];
    } else {
      out[(size_t)jw * nr] = acc[a][0];
    }
  }
}

// The same sums with the band grids of a tile of TR grid points (every
// Wannier function, every k) held in LDS: each grid value is then read from
// HBM once per tile and workgroup instead of once per q.  Arithmetic and
// summation order per (q, pair, r) are those of k_pair_rsum, except that the
// umklapp phase multiplies the jw factor (the same product, rounded
// differently: about 1e-16 relative).
// grid (ceil(nr/TR), ngroups, nqsplit), 512 threads; thread (rl = t % TR,
// ql = t / TR) takes the point r0 + rl and the q = qbeg + ql, qbeg + ql + QL, ...
// of its workgroup's range of qper q.  Dynamic LDS: the phase tables, w_k
// (ntk) and the grids [w][k][TR] (single precision, as in HBM).
constexpr int kRsumTileThreads = 512;
// jw per group in the full mode: the accumulators of JB jw x NI iw
template <int NI>
constexpr int rsum_tile_jb() { return NI <= 8 ? 5 : 2; }
__device__ __forceinline__ int rsum_wrap(int m, int n) {
  if (m >= n || m < -n) m %= n;  // only for umklapp vectors beyond +-1
  return m < 0 ? m + n : m;
}
constexpr int kRsumPF = 8;  // k of packed indices loaded ahead
template <bool FULL, int TR, int NI>
__global__ void __launch_bounds__(kRsumTileThreads) k_pair_rsum_tile(
    const float2* __restrict__ ur, int ntk, int nwf, int nx, int ny, int nz,
    const unsigned* __restrict__ kqs, const double2* __restrict__ wk, int nq, int qper,
    double2* __restrict__ P) {
  extern __shared__ double2 rs_lds[];
  double2* tab = rs_lds;               // 3 * kTwMax: e^{-i 2pi m/n} per axis
  double2* swk = tab + 3 * kTwMax;     // ntk
  float2* su = (float2*)(swk + ntk);   // [w][k][TR]
  constexpr int QL = kRsumTileThreads / TR;
  const double tpi = 2.0 * 3.141592653589793238462643383279502884;
  const int t = threadIdx.x;
  for (int i = t; i < nx + ny + nz; i += kRsumTileThreads) {
    const int ax = i < nx ? 0 : (i < nx + ny ? 1 : 2);
    const int m = ax == 0 ? i : (ax == 1 ? i - nx : i - nx - ny);
    const int n = ax == 0 ? nx : (ax == 1 ? ny : nz);
    double sn, cs;
    sincos(tpi * ((double)m / (double)n), &sn, &cs);
    tab[(ax == 0 ? 0 : (ax == 1 ? nx : nx + ny)) + m] = make_double2(cs, -sn);
  }
  for (int k = t; k < ntk; k += kRsumTileThreads) swk[k] = wk ? wk[k] : make_double2(1.0, 0.0);
  const int nr = nx * ny * nz;
  const int r0 = blockIdx.x * TR;
  for (int e = t; e < nwf * ntk * TR; e += kRsumTileThreads) {
    const int row = e / TR, rr = r0 + e % TR;  // row = w * ntk + k
    su[e] = rr < nr ? ur[(size_t)row * nr + rr] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const int rl = t % TR, ql = t / TR, z = blockIdx.y;
  const int r = r0 + rl;
  if (r >= nr) return;  // after the only barrier
  const int x = r % nx, y = (r / nx) % ny, zz = r / (nx * ny);
  const int npair = FULL ? nwf * nwf : nwf;
  constexpr int JB = rsum_tile_jb<NI>();
  constexpr int NA = FULL ? JB : NI;
  const int qend = min(nq, (int)(blockIdx.z + 1) * qper);
  for (int q = blockIdx.z * qper + ql; q < qend; q += QL) {
    double2 acc[NA][FULL ? NI : 1];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
      for (int b = 0; b < (FULL ? NI : 1); ++b) acc[a][b] = make_double2(0.0, 0.0);
    // packed (k+q, umklapp) of every k, kRsumPF k ahead of their use
    const unsigned* kq_row = kqs + (size_t)q * ntk;
    unsigned nxt[kRsumPF];
#pragma unroll
    for (int j = 0; j < kRsumPF; ++j) nxt[j] = j < ntk ? kq_row[j] : 0u;
    for (int k0 = 0; k0 < ntk; k0 += kRsumPF) {
      unsigned cur[kRsumPF];
#pragma unroll
      for (int j = 0; j < kRsumPF; ++j) {
        cur[j] = nxt[j];
        nxt[j] = k0 + kRsumPF + j < ntk ? kq_row[k0 + kRsumPF + j] : 0u;
      }
#pragma unroll
      for (int j = 0; j < kRsumPF; ++j) {
        const int k = k0 + j;
        if (k >= ntk) break;  // uniform
        const int kq = (int)(cur[j] & 0xfffffu);
        const int d1 = (int)((cur[j] >> 20) & 15u) - 8, d2 = (int)((cur[j] >> 24) & 15u) - 8,
                  d3 = (int)(cur[j] >> 28) - 8;
        double2 ph = make_double2(1.0, 0.0);
        if (d1 != 0 || d2 != 0 || d3 != 0) {
          const double2 px = tab[rsum_wrap(d1 * x, nx)];
          const double2 py = tab[nx + rsum_wrap(d2 * y, ny)];
          const double2 pz = tab[nx + ny + rsum_wrap(d3 * zz, nz)];
          const double2 tt = make_double2(px.x * py.x - px.y * py.y, px.x * py.y + px.y * py.x);
          ph = make_double2(tt.x * pz.x - tt.y * pz.y, tt.x * pz.y + tt.y * pz.x);
        }
        const double2 wv = swk[k];
        if constexpr (FULL) {
          // conj(u_iw,k+q ph) (w_k u_jw,k) = conj(u_iw,k+q) (conj(ph) w_k u_jw,k): the
          // phase goes with the jw factor, once per jw instead of once per iw
          const double2 cw = make_double2(ph.x * wv.x + ph.y * wv.y, ph.x * wv.y - ph.y * wv.x);
          double2 bq[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i) {
            bq[i] = make_double2(0.0, 0.0);
            if (i < nwf) {
              const float2 v = su[(i * ntk + kq) * TR + rl];
              bq[i] = make_double2((double)v.x, (double)v.y);
            }
          }
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            const int jw = z * JB + a;
            if (jw >= nwf) continue;
            const float2 v = su[(jw * ntk + k) * TR + rl];
            const double2 aw = make_double2((double)v.x * cw.x - (double)v.y * cw.y,
                                            (double)v.x * cw.y + (double)v.y * cw.x);
#pragma unroll
            for (int i = 0; i < NI; ++i) {
              if (i >= nwf) break;  // uniform
              acc[a][i].x += bq[i].x * aw.x + bq[i].y * aw.y;
              acc[a][i].y += bq[i].x * aw.y - bq[i].y * aw.x;
            }
          }
        } else {
          // conj(u_w,k+q ph) (w_k u_w,k) = conj(u_w,k+q) (conj(ph) w_k u_w,k)
          const double2 cw = make_double2(ph.x * wv.x + ph.y * wv.y, ph.x * wv.y - ph.y * wv.x);
#pragma unroll
          for (int a = 0; a < NA; ++a) {
            if (a >= nwf) continue;
            const float2 u = su[(a * ntk + kq) * TR + rl];
            const double2 b = make_double2((double)u.x, (double)u.y);
            const float2 v = su[(a * ntk + k) * TR + rl];
            const double2 aw = make_double2((double)v.x * cw.x - (double)v.y * cw.y,
                                            (double)v.x * cw.y + (double)v.y * cw.x);
            acc[a][0].x += b.x * aw.x + b.y * aw.y;
            acc[a][0].y += b.x * aw.y - b.y * aw.x;
          }
        }
      }
    }
    double2* out = P + (size_t)q * npair * nr + r;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const int jw = FULL ? z * JB + a : a;
      if (jw >= nwf) continue;
      if constexpr (FULL) {
#pragma unroll
        for (int i = 0; i < NI; ++i)
          if (i < nwf) out[(size_t)(i + nwf * jw) * nr] = acc[a][i];
      } else {
        out[(size_t)jw * nr] = acc[a][0];
      }
    }
  }
}

template <int N>
__device__ __forceinline__ void pd_zaxis_grid(const double2* __restrict__ src, int nx, int ny,
                                              unsigned zmask, const double2* tw, double2* g) {
  const int nl = nx * ny;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    double2 v[N];
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = src[l + nl * j];
    masked_line<N, true>(v, zmask, tw, g + l, nl);
  }
}

// rho_p(q+G) from the summed product grid P[q][p][r]: the pruned forward
// transform (e^{-i}, 1/N) read at -G for the psi-sphere points of q, / |q+G|,
// / NTK, G = 0 at q = 0 set to zero (calc_intW.F:720-767).  grid (npair, nq);
// rho[rho_off[q] + p*ngpsi + i].
template <int MAXN>
__global__ void __launch_bounds__(256) k_rho_from_grid(
    const double2* __restrict__ P, int npair, int nx, int ny, int nz, const int* __restrict__ plan,
    const long long* __restrict__ plan_off, const double* __restrict__ qg,
    const long long* __restrict__ qg_off, const int* __restrict__ g0s,
    const int* __restrict__ ngpsi, const long long* __restrict__ rho_off, int q_offset, int ntk,
    const double2* __restrict__ twtab, double2* __restrict__ rho) {
  extern __shared__ double2 lds[];
  const int p = blockIdx.x, ql = blockIdx.y, q = ql + q_offset;
  const int* pl = plan + plan_off[q];
  const int nzz = pl[0], nxl = pl[1];
  const unsigned zmask = (unsigned)pl[2];
  const int* ymask = pl + 4;
  const int* xline = ymask + nzz;
  const int* xmask = xline + nxl;
  const int* pos = xmask + nxl;
  double2* g = lds;  // [nzz][ny][nx]
  const int nr = nx * ny * nz;
  const double2* src = P + ((size_t)ql * npair + p) * nr;
  RP_PD_SWITCH(nz, MAXN, (pd_zaxis_grid<N>(src, nx, ny, zmask, twiddles_of(twtab, N, -1), g)))
  __syncthreads();
  RP_PD_SWITCH(ny, MAXN, (pd_yaxis<N>(g, nx, nzz, ymask, twiddles_of(twtab, N, -1))))
  __syncthreads();
  RP_PD_SWITCH(nx, MAXN, (pd_xaxis<N>(g, ny, nxl, xline, xmask, twiddles_of(twtab, N, -1))))
  __syncthreads();
  const int np = ngpsi[q], g0 = g0s[q];
  const double* len = qg + qg_off[q];
  const double scl = 1.0 / ((double)nr * (double)ntk);
  double2* out = rho + rho_off[q] + (size_t)p * np;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    const double2 v = g[pos[i]];
    out[i] = i == g0 ? make_double2(0.0, 0.0)
                     : make_double2(v.x * scl / len[i], v.y * scl / len[i]);
  }
}
#undef RP_PD_CASE
#undef RP_PD_SWITCH

// q = 0: the G = 0 entry per direction ix is -i<vir|p|occ>/de (calc_VMab,
// chiqw_sub.F:207-232, and chiqw.F:1787-1805), zero when |de| <= delt.
// Writes three copies A3[ix][K][lda] from A[K][lda] with that entry replaced.
// grid: K blocks, 64 threads.  coef as in k_band_grids.
__global__ void __launch_bounds__(64) k_q0_heads(
    const double2* __restrict__ coef, const int* __restrict__ kg, const int* __restrict__ ng0,
    int ntk, int ntg, const int* __restrict__ pocc, const int* __restrict__ pvir,
    const double* __restrict__ eband, const double* bvec, double delt, int g0, int ng,
    size_t lda, size_t rows, const float2* __restrict__ A, float2* __restrict__ A3) {
  const int K = blockIdx.x;
  const int p = K / ntk, k = K % ntk;
  const int jo = pocc[p], iv = pvir[p];
  const double2* ck = coef + ((size_t)jo * ntk + k) * ntg;
  const double2* ckq = coef + ((size_t)iv * ntk + k) * ntg;
  const int* kk = kg + (size_t)k * ntg * 3;
  double s[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  const int n = ng0[k];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double2 a = ck[i], b = ckq[i];
    // conj(b) * a
    double pr = b.x * a.x + b.y * a.y, pi = b.x * a.y - b.y * a.x;
    for (int d = 0; d < 3; ++d) {
      double gv = (double)kk[3 * i] * bvec[d] + (double)kk[3 * i + 1] * bvec[3 + d] +
                  (double)kk[3 * i + 2] * bvec[6 + d];
      s[d][0] += gv * pr;
      s[d][1] += gv * pi;
    }
  }
  __shared__ double red[6][64];
  for (int d = 0; d < 3; ++d) { red[2 * d][threadIdx.x] = s[d][0]; red[2 * d + 1][threadIdx.x] = s[d][1]; }
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  double de = eband[(size_t)iv * ntk + k] - eband[(size_t)jo * ntk + k];
  if (fabs(de) < delt) de = delt;
  const float2* src = A + (size_t)K * lda;
  for (int ix = 0; ix < 3; ++ix) {
    float2* dst = A3 + ((size_t)ix * rows + K) * lda;
    for (int i = threadIdx.x; i < (int)lda; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    if (threadIdx.x == 0 && g0 >= 0 && g0 < ng) {
      // vm = i*S ; vm = -i*(vm - 0)/de = S/de
      double vr = red[2 * ix][0] / de, vi = red[2 * ix + 1][0] / de;
      // m_tmp is complex(8) in the reference (no single rounding of this entry)
      dst[g0] = fabs(de) <= delt ? make_float2(0.f, 0.f) : make_float2((float)vr, (float)vi);
    }
  }
}

size_t twiddle_entries() { return 2 * (size_t)kTwHalf; }

void launch_twiddles(double2* tab, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_twiddles, dim3(8), dim3(256), 0, s, tab);
  RP_HIP(hipGetLastError());
}

#define RP_LD_CASE(NN)                                                                       \
  case NN:                                                                                   \
    hipLaunchKernelGGL(k_line_dft<NN>, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s, g, \
                       nl, axis, nx, ny, nz, twiddles_of(tab, NN, sign));                    \
    break;
void launch_line_dft3(double2* g, long long nb, int nx, int ny, int nz, int sign,
                      const double2* tab, hipStream_t s) {
  const long long nr = (long long)nx * ny * nz;
  const int axes[3] = {2, 1, 0};  // z, then y, then x (m_fft3d_20150826.f90:1598-1720)
  for (int axis : axes) {
    const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const long long nl = nb * (nr / n);
    switch (n) {
      RP_LD_CASE(1) RP_LD_CASE(2) RP_LD_CASE(3) RP_LD_CASE(4) RP_LD_CASE(5) RP_LD_CASE(6)
      RP_LD_CASE(8) RP_LD_CASE(9) RP_LD_CASE(10) RP_LD_CASE(12) RP_LD_CASE(15) RP_LD_CASE(16)
      RP_LD_CASE(18) RP_LD_CASE(20) RP_LD_CASE(24) RP_LD_CASE(25) RP_LD_CASE(27) RP_LD_CASE(30)
      RP_LD_CASE(32)
      default: throw Error("FFT axis length " + std::to_string(n) + " not supported");
    }
    RP_HIP(hipGetLastError());
  }
}
#undef RP_LD_CASE

// Grids that do not fit the LDS go through HBM in batches of at most this size.
constexpr size_t kHbmBatchBytes = size_t(1) << 30;

size_t band_grids_scratch_elems(int nx, int ny, int nz, size_t ngrids) {
  const size_t nr = (size_t)nx * ny * nz;
  if (nr * sizeof(double2) <= kLdsBytes) return 0;
  return std::min<size_t>(ngrids, std::max<size_t>(1, kHbmBatchBytes / (nr * sizeof(double2)))) * nr;
}

void launch_band_grids(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                       int nband, int nx, int ny, int nz, const double2* twtab, float2* ur,
                       double2* scratch, size_t scratch_elems, hipStream_t s) {
  const size_t nr = (size_t)nx * ny * nz;
  if (nr * sizeof(double2) > kLdsBytes) {
    // backward transform as line passes over HBM, in batches of grids
    RP_CHECK(scratch && scratch_elems >= nr, "band grids: HBM scratch missing");
    const long long total = (long long)nband * ntk;
    const long long per = (long long)(scratch_elems / nr);
    for (long long gb0 = 0; gb0 < total; gb0 += per) {
      const long long nb = std::min(per, total - gb0);
      launch_zero(scratch, (size_t)nb * nr * sizeof(double2), s);
      hipLaunchKernelGGL(k_scatter_coef, dim3((ntg + 255) / 256, (unsigned)nb), dim3(256), 0, s, coef,
                         kg, ng0, ntk, ntg, nx, ny, nz, gb0, scratch);
      RP_HIP(hipGetLastError());
      launch_line_dft3(scratch, nb, nx, ny, nz, +1, twtab, s);
      hipLaunchKernelGGL(k_round_grids, dim3(4096), dim3(256), 0, s, scratch, (size_t)nb * nr,
                         ur + (size_t)gb0 * nr);
      RP_HIP(hipGetLastError());
    }
    return;
  }
  size_t lds = sizeof(double2) * nr;
  const bool small = nx <= 16 && ny <= 16 && nz <= 16;
  const void* fn = small ? (const void*)k_band_grids<16> : (const void*)k_band_grids<32>;
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(ntk, nband);
  if (small)
    hipLaunchKernelGGL(k_band_grids<16>, grid, dim3(256), lds, s, coef, kg, ng0, ntk, ntg, nx, ny, nz, twtab, ur);
  else
    hipLaunchKernelGGL(k_band_grids<32>, grid, dim3(256), lds, s, coef, kg, ng0, ntk, ntg, nx, ny, nz, twtab, ur);
  RP_HIP(hipGetLastError());
}

void launch_pair_densities_pruned(const float2* ur, int ntk, const int* pocc, const int* pvir,
                                  const int* ikq, const int* shift, int nx, int ny, int nz,
                                  const int* plan, int nzz, int ng, size_t lda, int nK,
                                  const double2* twtab, float2* A, hipStream_t s) {
  const size_t lds = sizeof(double2) * ((size_t)nx * ny * nzz + nx + ny + nz);
  const bool small = nx <= 16 && ny <= 16 && nz <= 16;
  const void* fn =
      small ? (const void*)k_pair_densities_pruned<16> : (const void*)k_pair_densities_pruned<32>;
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int chunk = 1 << 30;
  for (int off = 0; off < nK; off += chunk) {
    const int n = nK - off < chunk ? nK - off : chunk;
    if (small)
      hipLaunchKernelGGL(k_pair_densities_pruned<16>, dim3(n), dim3(256), lds, s, ur, ntk, pocc,
                         pvir, ikq, shift, nx, ny, nz, plan, ng, lda, off, twtab, A);
    else
      hipLaunchKernelGGL(k_pair_densities_pruned<32>, dim3(n), dim3(256), lds, s, ur, ntk, pocc,
                         pvir, ikq, shift, nx, ny, nz, plan, ng, lda, off, twtab, A);
    RP_HIP(hipGetLastError());
  }
}

void launch_pair_rsum(const float2* ur, int ntk, int nwf, bool full, int nx, int ny, int nz,
                      const int* ikq, const int* shift, const unsigned* kqs, const double2* wk,
                      int nq, double2* P, hipStream_t s) {
  RP_CHECK(nwf >= 1 && nwf <= 16, "Wannier pair sums take at most 16 Wannier functions");
  RP_CHECK(nx <= kTwMax && ny <= kTwMax && nz <= kTwMax, "FFT axis too long for the phase tables");
  const int nr = nx * ny * nz;
  if (nq <= 0) return;
  const bool wide = nwf > 8;  // NI = 16 forms
  const int jb = wide ? rsum_tile_jb<16>() : rsum_tile_jb<8>();
  // The LDS-tiled form when the packed indices exist and the grids of one
  // point fit (NWF * NTK <= 20000).  Tile size and q split: least (rounds of
  // workgroups) x (q per thread), with the workgroups a compute unit holds by
  // LDS and by registers (two of the diagonal form with NI = 8, else one).
  const size_t fixed = sizeof(double2) * (3 * kTwMax + (size_t)ntk);
  const size_t per_point = sizeof(float2) * (size_t)nwf * ntk;
  int tr = 0, qper = nq;
  if (kqs) {
    int dev = 0, ncu = 256;
    RP_HIP(hipGetDevice(&dev));
    RP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int groups = full ? (nwf + jb - 1) / jb : 1;
    long long best = -1;
    for (int c : {16, 8, 4, 2, 1}) {
      const size_t lds = fixed + per_point * c;
      if (lds > kLdsBytes) continue;
      const int per_cu = std::min<int>((int)(kLdsBytes / lds), (full || wide) ? 1 : 2);
      const int tiles = (nr + c - 1) / c, ql = kRsumTileThreads / c;
      for (int split = 1; split <= 8; ++split) {
        const int qp = (nq + split - 1) / split;
        const int ns = (nq + qp - 1) / qp;
        const long long rounds = ((long long)tiles * groups * ns + (long long)ncu * per_cu - 1) /
                                 ((long long)ncu * per_cu);
        const long long cost = rounds * ((qp + ql - 1) / ql);
        if (best < 0 || cost < best) { best = cost; tr = c; qper = qp; }
      }
    }
  }
  if (tr) {
    const int groups = full ? (nwf + jb - 1) / jb : 1;
    const int tiles = (nr + tr - 1) / tr;
    const int nsplit = (nq + qper - 1) / qper;
    const size_t lds = fixed + per_point * tr;
    dim3 grid(tiles, groups, nsplit);
    auto go = [&](auto fn) {
      RP_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(fn, grid, dim3(kRsumTileThreads), lds, s, ur, ntk, nwf, nx, ny, nz, kqs,
                         wk, nq, qper, P);
    };
    auto pick = [&](auto full_tag, auto ni_tag) {
      constexpr bool F = decltype(full_tag)::value;
      constexpr int NI = decltype(ni_tag)::value;
      if (tr == 16) go(k_pair_rsum_tile<F, 16, NI>);
      else if (tr == 8) go(k_pair_rsum_tile<F, 8, NI>);
      else if (tr == 4) go(k_pair_rsum_tile<F, 4, NI>);
      else if (tr == 2) go(k_pair_rsum_tile<F, 2, NI>);
      else go(k_pair_rsum_tile<F, 1, NI>);
    };
    using N8 = std::integral_constant<int, 8>;
    using N16 = std::integral_constant<int, 16>;
    if (full) {
      if (wide) pick(std::true_type{}, N16{});
      else pick(std::true_type{}, N8{});
    } else {
      if (wide) pick(std::false_type{}, N16{});
      else pick(std::false_type{}, N8{});
    }
    RP_HIP(hipGetLastError());
    return;
  }
  const int groups = full ? (nwf + kRsumJB - 1) / kRsumJB : 1;
  dim3 grid((nr + 255) / 256, nq, groups);
  auto direct = [&](auto fn) {
    hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, ur, ntk, nwf, nx, ny, nz, ikq, shift, wk, P);
  };
  if (full) {
    if (wide) direct(k_pair_rsum<true, 16>);
    else direct(k_pair_rsum<true, 8>);
  } else {
    if (wide) direct(k_pair_rsum<false, 16>);
    else direct(k_pair_rsum<false, 8>);
  }
  RP_HIP(hipGetLastError());
}

void launch_rho_from_grid(const double2* P, int npair, int nx, int ny, int nz, const int* plan,
                          const long long* plan_off, int max_nzz, const double* qg,
                          const long long* qg_off, const int* g0s, const int* ngpsi,
                          const long long* rho_off, int q_offset, int nq, int ntk,
                          const double2* twtab, double2* rho, hipStream_t s) {
  const size_t lds = sizeof(double2) * ((size_t)nx * ny * max_nzz);
  const bool small = nx <= 16 && ny <= 16 && nz <= 16;
  const void* fn = small ? (const void*)k_rho_from_grid<16> : (const void*)k_rho_from_grid<32>;
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(npair, nq);
  if (small)
    hipLaunchKernelGGL(k_rho_from_grid<16>, grid, dim3(256), lds, s, P, npair, nx, ny, nz, plan,
                       plan_off, qg, qg_off, g0s, ngpsi, rho_off, q_offset, ntk, twtab, rho);
  else
    hipLaunchKernelGGL(k_rho_from_grid<32>, grid, dim3(256), lds, s, P, npair, nx, ny, nz, plan,
                       plan_off, qg, qg_off, g0s, ngpsi, rho_off, q_offset, ntk, twtab, rho);
  RP_HIP(hipGetLastError());
}

// rho from transformed grids P[q][p][r] (the HBM path of k_rho_from_grid): the
// points -G of the psi sphere (lg at lg_off, 3 ints per point), same scaling.
__global__ void __launch_bounds__(256) k_rho_gather(
    const double2* __restrict__ P, int npair, int nx, int ny, int nz, const int* __restrict__ lg,
    const double* __restrict__ qg, const long long* __restrict__ off, const int* __restrict__ g0s,
    const int* __restrict__ ngpsi, const long long* __restrict__ rho_off, int q_offset, int ntk,
    double2* __restrict__ rho) {
  const int p = blockIdx.x, ql = blockIdx.y, q = ql + q_offset;
  const int np = ngpsi[q], g0 = g0s[q];
  const int nr = nx * ny * nz;
  const double2* src = P + ((size_t)ql * npair + p) * nr;
  const int* l = lg + 3 * off[q];
  const double* len = qg + off[q];
  const double scl = 1.0 / ((double)nr * (double)ntk);
  double2* out = rho + rho_off[q] + (size_t)p * np;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    const int r = wrap(-l[3 * i], nx) + nx * (wrap(-l[3 * i + 1], ny) + ny * wrap(-l[3 * i + 2], nz));
    const double2 v = src[r];
    out[i] = i == g0 ? make_double2(0.0, 0.0)
                     : make_double2(v.x * scl / len[i], v.y * scl / len[i]);
  }
}

void launch_rho_from_grid_hbm(double2* P, int npair, int nx, int ny, int nz, const int* lg,
                              const double* qg, const long long* off, const int* g0s,
                              const int* ngpsi, const long long* rho_off, int q_offset, int nq,
                              int ntk, const double2* twtab, double2* rho, hipStream_t s) {
  launch_line_dft3(P, (long long)nq * npair, nx, ny, nz, -1, twtab, s);
  hipLaunchKernelGGL(k_rho_gather, dim3(npair, nq), dim3(256), 0, s, P, npair, nx, ny, nz, lg, qg,
                     off, g0s, ngpsi, rho_off, q_offset, ntk, rho);
  RP_HIP(hipGetLastError());
}

size_t pruned_transform_lds_bytes(int nx, int ny, int nz, int nzz) {
  return sizeof(double2) * ((size_t)nx * ny * nzz + nx + ny + nz);
}

void launch_q0_heads(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                     const int* pocc, const int* pvir, const double* eband, const double* bvec,
                     double delt, int g0, int ng, size_t lda, size_t rows, int nK, const float2* A,
                     float2* A3, hipStream_t s) {
  hipLaunchKernelGGL(k_q0_heads, dim3(nK), dim3(64), 0, s, coef, kg, ng0, ntk, ntg, pocc, pvir,
                     eband, bvec, delt, g0, ng, lda, rows, A, A3);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
