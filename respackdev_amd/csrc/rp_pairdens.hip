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
// Both transforms are LDS-resident: the FP64 grid (<= 10000 points) lives in
// one workgroup's LDS and each axis is done in place as independent per-thread
// line DFTs (the axis lengths are the reference's 2,3,5-smooth sizes, <= 32),
// so HBM sees one read of the inputs and one write of the outputs per
// transform.  The forward transform of stage 2 is pruned to the outputs the
// sphere needs (below).
#include <vector>

#include "rp_common.h"
#include "rp_fft.h"
#include "rp_kernels.h"

namespace rp {

// Largest grid the LDS-resident transform holds (the FP64 grid within the
// 160 KB of LDS a gfx950 workgroup may use): 20*20*25 points.
constexpr int kMaxGrid = 10000;

// Twiddle tables in HBM for every line length N = 1..32, one row per output
// index: T_N[k][j] = w^{(jk) mod N} with w = exp(-2 pi i/N) (forward table) or
// its conjugate (backward table), at tab[dir*kTwHalf + off(N) + k*N + j].  A line
// DFT reads row k with offsets fixed at compile time from a wave-uniform base,
// so the twiddles arrive through scalar loads into SGPRs: no LDS traffic, no
// vector instructions and no per-term index arithmetic.
constexpr int kTwMax = 32;
constexpr int kTwHalf = (kTwMax - 1) * kTwMax * (2 * kTwMax - 1) / 6 + kTwMax * kTwMax;
__host__ __device__ constexpr int tw_off(int n) { return (n - 1) * n * (2 * n - 1) / 6; }
__device__ __forceinline__ const double2* twiddles_of(const double2* tab, int n, int sign) {
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
int max_grid_points() { return kMaxGrid; }

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

void launch_band_grids(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                       int nband, int nx, int ny, int nz, const double2* twtab, float2* ur,
                       hipStream_t s) {
  size_t lds = sizeof(double2) * ((size_t)nx * ny * nz);
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


void launch_q0_heads(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                     const int* pocc, const int* pvir, const double* eband, const double* bvec,
                     double delt, int g0, int ng, size_t lda, size_t rows, int nK, const float2* A,
                     float2* A3, hipStream_t s) {
  hipLaunchKernelGGL(k_q0_heads, dim3(nK), dim3(64), 0, s, coef, kg, ng0, ntk, ntg, pocc, pvir,
                     eband, bvec, delt, g0, ng, lda, rows, A, A3);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
