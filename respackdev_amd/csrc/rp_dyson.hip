// Dyson step of calc_eps_rpa / calc_eps_rpa_q_0 (chiqw_sub.F:386-568) in FP64
// on gfx950: eps_GG' = delta - 2 chi_GG' 4pi/(|q+G||q+G'|) (head, wings and
// the Drude term at q = 0), then eps^-1 for every frequency of the batch.
//
// The inversion is Gauss-Jordan with partial pivoting by |Re|+|Im| (the pivot
// rule of LAPACK's izamax in zgetrf, which the reference calls through
// invZGE.F90:1-23), blocked by panels of up to 16 columns: one workgroup per
// matrix factors a panel (in registers for n <= 768, else in LDS), a batched
// kernel applies the panel's rank-nb update to all other columns (for n <= 768
// in the same launch as the factoring of the next panel), and a final kernel
// undoes the row interchanges as column swaps.
#include <algorithm>

#include <utility>

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

// acc + a b and acc - a b for complex FP64 as four fused multiply-adds each
// (a product's two terms, each folded into the accumulator).
__device__ __forceinline__ double2 cfma(double2 acc, double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y)));
}
__device__ __forceinline__ double2 cfms(double2 acc, double2 a, double2 b) {
  return make_double2(fma(-a.x, b.x, fma(a.y, b.y, acc.x)), fma(-a.x, b.y, fma(-a.y, b.x, acc.y)));
}

// Largest panel width; panel rows moved by the pivoting of one panel: at most
// the nb panel rows plus nb others.
constexpr int kPanelMax = 16;
constexpr int kMoveMax = 2 * kPanelMax;
// Per matrix the panel kernel leaves {count, dst[kMoveMax], src[kMoveMax]}.
constexpr int kMoveInts = 1 + 2 * kMoveMax;

// Blocked Gauss-Jordan.  The nb pivot steps of panel [c0, c0+nb) are
// elementary transforms E_k that differ from the identity only in column k;
// their product differs from the identity only in the panel columns, and
// running the in-place Gauss-Jordan steps on the panel columns alone leaves
// exactly those columns W in the panel.  Every other column j then becomes
//   a'(:,j) = P a(:,j) with the panel rows zeroed + W (P a)(panel rows, j),
// P the panel's row interchanges: one rank-nb update.

// One workgroup per matrix: the panel (n x nb, FP64 complex) is factored in
// LDS; row interchanges are applied to the panel and recorded as a move list
// for the update kernel.  LDS: nb*n double2 + n ints.
__global__ void __launch_bounds__(1024) k_gj_panel(double2* __restrict__ a, int n, int c0, int nb,
                                                   int* __restrict__ ipiv, int* __restrict__ moves,
                                                   int* __restrict__ status) {
  extern __shared__ double2 lds[];
  double2* P = lds;                      // P[l*n + i] = a(i, c0+l)
  int* orig = (int*)(P + (size_t)nb * n);  // orig[i]: original row now in row i
  __shared__ double wv[16];
  __shared__ int wi[16];
  __shared__ double2 prow[kPanelMax];
  __shared__ int nmove;
  __shared__ int singular;
  const int b = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  double2* m = a + (size_t)b * n * n;
  for (int l = 0; l < nb; ++l)
    for (int i = tid; i < n; i += nt) P[(size_t)l * n + i] = m[i + (size_t)n * (c0 + l)];
  for (int i = tid; i < n; i += nt) orig[i] = i;
  if (tid == 0) { nmove = 0; singular = 0; }
  __syncthreads();
  for (int l = 0; l < nb; ++l) {
    const int c = c0 + l;
    const double2* col = P + (size_t)l * n;
    // pivot: largest |Re|+|Im| in rows c..n-1, first index on ties (izamax)
    double best = -1.0;
    int bi = n;
    for (int i = c + tid; i < n; i += nt) {
      const double v = cabs1(col[i]);
      if (v > best) { best = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double v2 = __shfl_xor(best, off);
      const int i2 = __shfl_xor(bi, off);
      if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
    }
    if ((tid & 63) == 0) { wv[tid >> 6] = best; wi[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < nt / 64; ++w)
        if (wv[w] > best || (wv[w] == best && wi[w] < bi)) { best = wv[w]; bi = wi[w]; }
      if (bi >= n) bi = c;  // no candidate compared (NaN column): keep row c, as izamax does
      wi[0] = bi;
      ipiv[(size_t)b * n + c] = bi;
      const int t = orig[c];
      orig[c] = orig[bi];
      orig[bi] = t;
    }
    __syncthreads();
    const int p = wi[0];
    if (p != c && tid < nb) {
      double2* cl = P + (size_t)tid * n;
      const double2 t = cl[c];
      cl[c] = cl[p];
      cl[p] = t;
    }
    __syncthreads();
    const double2 piv = col[c];
    if (piv.x == 0.0 && piv.y == 0.0) {
      if (tid == 0) { atomicOr(status, 4); singular = 1; }
      break;  // uniform: every thread reads the same pivot
    }
    // pivot row c scaled (its column l becomes 1/piv)
    if (tid < nb) {
      const zd pinv = mkz(1.0, 0.0) / mkz(piv.x, piv.y);
      const double2 v = P[(size_t)tid * n + c];
      const zd u = tid == l ? pinv : mkz(v.x, v.y) * pinv;
      prow[tid] = make_double2(u.x, u.y);
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
      if (i == c) {
        for (int q = 0; q < nb; ++q) P[(size_t)q * n + i] = prow[q];
      } else {
        const double2 t = col[i];
        for (int q = 0; q < nb; ++q) {
          double2* e = P + (size_t)q * n + i;
          const double2 v = q == l ? make_double2(0.0, 0.0) : *e;
          *e = cfms(v, t, prow[q]);
        }
      }
    }
    __syncthreads();
  }
  for (int l = 0; l < nb; ++l)
    for (int i = tid; i < n; i += nt) m[i + (size_t)n * (c0 + l)] = P[(size_t)l * n + i];
  // move list: the nb panel rows first (in order), then other rows whose
  // content changed
  int* mv = moves + (size_t)b * kMoveInts;
  if (tid < nb) {
    mv[1 + tid] = c0 + tid;
    mv[1 + kMoveMax + tid] = orig[c0 + tid];
  }
  for (int i = tid; i < n; i += nt) {
    if ((i < c0 || i >= c0 + nb) && orig[i] != i) {
      const int e = nb + atomicAdd(&nmove, 1);
      mv[1 + e] = i;
      mv[1 + kMoveMax + e] = orig[i];
    }
  }
  __syncthreads();
  if (tid == 0) mv[0] = singular ? -1 : nb + nmove;
}

// The same panel step for full panels (nb = 16) of n <= 768 rows, with the
// panel in registers: thread i starts with row i of the 16 panel columns.  The
// 16 column steps are expanded at compile time (gj_reg_step<L>), so every
// register index is a constant.  Rows are never exchanged between threads:
// each thread carries the matrix row it holds (lr), the pivot search takes the
// largest |Re|+|Im| over rows lr >= c with the smallest row on ties (izamax),
// and the interchange of rows c and p only swaps the two threads' row numbers.
// The pivot row goes to the others through LDS: two barriers per column.  The
// elimination multiplies by 1/piv once per row, r - (r_L/piv) rowp, where
// k_gj_panel scales the pivot row first, r - r_L (rowp/piv): the same
// arithmetic up to the order of one product.  The rows go back to the matrix at
// their row numbers (gj_reg_store).
struct GjRegShared {
  double wv[16];
  int wi[16];
  double2 rowp[kPanelMax];
  double2 pinv;  // 1/piv, formed by the pivot row's thread
  int orig[768];
  int origb[768];  // the rows' sources since the last reset (one half of a 32-column panel)
  int nmove;
  int singular;
};

// A double moved across lanes by DPP (two 32-bit moves; every lane has a source).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
  const unsigned lo2 = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xf, 0xf, false);
  const unsigned hi2 = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi2 << 32) | lo2);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// A zero pivot sets the matrix's status bit and sh.singular and the steps go
// on (on values that are then discarded): every thread runs the same steps and
// meets the same barriers, and no branch keeps an older copy of the registers.
template <int L>
__device__ __forceinline__ void gj_reg_step(double2 (&r)[kPanelMax], int& lr, GjRegShared& sh,
                                            int i, int n, int c0, bool mine, int b,
                                            int* __restrict__ ipiv, int* __restrict__ status) {
  constexpr int NB = kPanelMax;
  const int lane = i & 63, wave = i >> 6, nwave = (int)(blockDim.x >> 6);
  const int c = c0 + L;
  // pivot: largest |Re|+|Im| in rows c..n-1, first row on ties (izamax)
  const double own = (mine && lr >= c) ? fabs(r[L].x) + fabs(r[L].y) : -1.0;
  // wave maximum: butterfly within rows of 16 lanes (DPP), then the four rows
  double best = own;
  best = fmax(best, dpp_d<0xB1>(best));   // quad_perm [1,0,3,2]
  best = fmax(best, dpp_d<0x4E>(best));   // quad_perm [2,3,0,1]
  best = fmax(best, dpp_d<0x124>(best));  // row_ror:4
  best = fmax(best, dpp_d<0x128>(best));  // row_ror:8
  best = fmax(fmax(readlane_d(best, 0), readlane_d(best, 16)),
              fmax(readlane_d(best, 32), readlane_d(best, 48)));
  // the smallest row holding it (n: no candidate here)
  int key = (best >= 0.0 && own == best) ? lr : n;
  key = min(key, __builtin_amdgcn_update_dpp(key, key, 0xB1, 0xf, 0xf, false));
  key = min(key, __builtin_amdgcn_update_dpp(key, key, 0x4E, 0xf, 0xf, false));
  key = min(key, __builtin_amdgcn_update_dpp(key, key, 0x124, 0xf, 0xf, false));
  key = min(key, __builtin_amdgcn_update_dpp(key, key, 0x128, 0xf, 0xf, false));
  key = min(min(__builtin_amdgcn_readlane(key, 0), __builtin_amdgcn_readlane(key, 16)),
            min(__builtin_amdgcn_readlane(key, 32), __builtin_amdgcn_readlane(key, 48)));
  if (lane == 0) { sh.wv[wave] = best; sh.wi[wave] = key; }
  __syncthreads();
  best = sh.wv[0];
  int bi = sh.wi[0];
  for (int w = 1; w < nwave; ++w)
    if (sh.wv[w] > best || (sh.wv[w] == best && sh.wi[w] < bi)) { best = sh.wv[w]; bi = sh.wi[w]; }
  const int p = bi < n ? bi : c;  // no candidate compared (NaN column): keep row c
  if (i == 0) {
    ipiv[(size_t)b * n + c] = p;
    const int t = sh.orig[c];
    sh.orig[c] = sh.orig[p];
    sh.orig[p] = t;
    const int tb = sh.origb[c];
    sh.origb[c] = sh.origb[p];
    sh.origb[p] = tb;
  }
  const bool is_p = mine && lr == p;
  if (is_p) {
#pragma unroll
    for (int q = 0; q < NB; ++q) sh.rowp[q] = r[q];
    const zd pz = mkz(1.0, 0.0) / mkz(r[L].x, r[L].y);
    sh.pinv = make_double2(pz.x, pz.y);
    if (r[L].x == 0.0 && r[L].y == 0.0) {
      atomicOr(status, 4);
      sh.singular = 1;
    }
  }
  __syncthreads();
  const double2 pinv = sh.pinv;
  if (is_p) {
    // the pivot row becomes row c, scaled (its column L becomes 1/piv)
#pragma unroll
    for (int q = 0; q < NB; ++q)
      r[q] = q == L ? pinv : cfma(make_double2(0.0, 0.0), r[q], pinv);
    lr = c;
  } else if (mine) {
    if (lr == c) lr = p;  // row c's content is row p's now
    const double2 s = cfma(make_double2(0.0, 0.0), r[L], pinv);
#pragma unroll
    for (int q = 0; q < NB; ++q)
      r[q] = q == L ? make_double2(-s.x, -s.y) : cfms(r[q], s, sh.rowp[q]);
  }
}

template <int... L>
__device__ __forceinline__ void gj_reg_steps(std::integer_sequence<int, L...>, double2 (&r)[kPanelMax],
                                             int& lr, GjRegShared& sh, int i, int n, int c0,
                                             bool mine, int b, int* __restrict__ ipiv,
                                             int* __restrict__ status) {
  (gj_reg_step<L>(r, lr, sh, i, n, c0, mine, b, ipiv, status), ...);
}

// Columns [c0, c0+16) of the rows the threads hold, back to the matrix at
// their row numbers.
__device__ __forceinline__ void gj_reg_store(const double2 (&r)[kPanelMax], int lr, bool mine,
                                             double2* __restrict__ m, int n, int c0) {
  if (mine)
#pragma unroll
    for (int q = 0; q < kPanelMax; ++q) m[lr + (size_t)n * (c0 + q)] = r[q];
}

// Factors the panel [c0, c0+16) of matrix m held in r (row i of the panel in
// thread i), writes it back and leaves the move list in mv.
__device__ __forceinline__ void gj_reg_factor(double2 (&r)[kPanelMax], GjRegShared& sh, int i,
                                              int n, int c0, bool mine, int b,
                                              double2* __restrict__ m, int* __restrict__ ipiv,
                                              int* __restrict__ mv, int* __restrict__ status) {
  constexpr int NB = kPanelMax;
  if (mine) sh.orig[i] = i;
  if (i == 0) { sh.nmove = 0; sh.singular = 0; }
  __syncthreads();
  int lr = i;
  gj_reg_steps(std::make_integer_sequence<int, NB>{}, r, lr, sh, i, n, c0, mine, b, ipiv, status);
  __syncthreads();
  const bool singular = sh.singular != 0;
  if (!singular) gj_reg_store(r, lr, mine, m, n, c0);
  // move list: the nb panel rows first (in order), then other rows whose
  // content changed
  if (i < NB) {
    mv[1 + i] = c0 + i;
    mv[1 + kMoveMax + i] = sh.orig[c0 + i];
  }
  if (mine && (i < c0 || i >= c0 + NB) && sh.orig[i] != i) {
    const int e = NB + atomicAdd(&sh.nmove, 1);
    mv[1 + e] = i;
    mv[1 + kMoveMax + e] = sh.orig[i];
  }
  __syncthreads();
  if (i == 0) mv[0] = singular ? -1 : NB + sh.nmove;
}

__global__ void __launch_bounds__(768) k_gj_panel_reg(double2* __restrict__ a, int n, int c0,
                                                      int* __restrict__ ipiv,
                                                      int* __restrict__ moves,
                                                      int* __restrict__ status) {
  constexpr int NB = kPanelMax;
  __shared__ GjRegShared sh;
  const int b = blockIdx.x, i = threadIdx.x;
  double2* m = a + (size_t)b * n * n;
  const bool mine = i < n;
  double2 r[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) r[q] = mine ? m[i + (size_t)n * (c0 + q)] : make_double2(0.0, 0.0);
  gj_reg_factor(r, sh, i, n, c0, mine, b, m, ipiv, moves + (size_t)b * kMoveInts, status);
}

// One step of the register path (n <= 768, full panels) with look-ahead: the
// rank-16 update of panel k (columns [ck, ck+16), factored by the previous
// launch, its move list in moves_k) and, when next is set, the factoring of
// panel k+1 in the same launch.  Blocks [0, nbatch) (next set) are panel
// blocks: they apply panel k's update to the 16 columns of panel k+1 in
// registers and factor them at once (gj_reg_factor), writing the move list to
// moves_next.  Every other block updates one group of up to 16 other columns
// of one matrix: the columns outside [ck, ck + 16 (1 + next)), in groups of
// 16, ngroups per matrix.  The two kinds touch disjoint columns and read
// panel k's columns only, so nothing in the launch waits for anything else:
// the panel factoring (one block per matrix, latency-bound) runs beside the
// bandwidth-bound update instead of after it.  Thread i handles row i; the
// arithmetic is k_gj_update's, operation for operation.  The rows the move
// list names (at most 32) are staged in LDS first; after that each thread
// reads and writes its own row only, so the waves run without further
// barriers and one wave's loads overlap another's arithmetic.  With nbatch a
// multiple of 8 the update blocks of matrix b run on XCD b mod 8 (blocks are
// dealt to the XCDs round-robin), so each XCD's L2 holds its matrices' panels.
struct GjStepShared {
  GjRegShared p;
  // old values of the moved rows in the block's columns: vals[e][t] = a(src_e, cj + t);
  // the first 16 entries are panel k's pivot rows
  double2 vals[kMoveMax][kPanelMax];
  int dst[kMoveMax], src[kMoveMax];
};

__global__ void __launch_bounds__(768) k_gj_step(double2* __restrict__ a, int n, int ck, int next,
                                                 int nbatch, int ngroups, int* __restrict__ ipiv,
                                                 const int* __restrict__ moves_k,
                                                 int* __restrict__ moves_next,
                                                 int* __restrict__ status) {
  constexpr int NB = kPanelMax;
  __shared__ GjStepShared sh;
  const int i = threadIdx.x;
  const bool mine = i < n;
  const bool panel = next && (int)blockIdx.x < nbatch;
  int b, cj, ntj;
  if (panel) {
    b = blockIdx.x;
    cj = ck + NB;
    ntj = NB;
  } else {
    const int u = (int)blockIdx.x - (next ? nbatch : 0);
    int t;
    if (nbatch % 8 == 0) {
      const int x = u % 8, s = u / 8;
      b = x + 8 * (s / ngroups);
      t = s % ngroups;
    } else {
      b = u / ngroups;
      t = u % ngroups;
    }
    const int skip = NB * (1 + next);  // columns [ck, ck + skip) are not in the groups
    const int t0 = t * NB;             // groups never straddle ck (both multiples of 16)
    cj = t0 < ck ? t0 : t0 + skip;
    ntj = min(NB, n - skip - t0);
  }
  double2* m = a + (size_t)b * n * n;
  const int* mv = moves_k + (size_t)b * kMoveInts;
  const int nm = mv[0];
  if (nm < 0) {  // a zero pivot earlier (the host reports it): nothing more for this matrix
    if (panel && i == 0) moves_next[(size_t)b * kMoveInts] = -1;
    return;  // uniform: every thread read the same word
  }
  if (i < nm) {
    sh.dst[i] = mv[1 + i];
    sh.src[i] = mv[1 + kMoveMax + i];
  }
  __syncthreads();
  // the only rows a thread reads besides its own: those the move list names
  for (int x = i; x < nm * NB; x += blockDim.x) {
    const int e = x / NB, t = x % NB;
    sh.vals[e][t] = t < ntj ? m[sh.src[e] + (size_t)n * (cj + t)] : make_double2(0.0, 0.0);
  }
  __syncthreads();  // after this each thread reads and writes its own row only
  const bool inpanel = i >= ck && i < ck + NB;
  int slot = -1;
  if (!inpanel)
    for (int e = NB; e < nm; ++e)
      if (sh.dst[e] == i) slot = e;
  double2 r[NB];
#pragma unroll
  for (int t = 0; t < NB; ++t)
    r[t] = (!mine || inpanel || t >= ntj) ? make_double2(0.0, 0.0)
           : slot >= 0                    ? sh.vals[slot][t]
                                          : m[i + (size_t)n * (cj + t)];
  if (mine) {
    for (int q = 0; q < NB; ++q) {
      const double2 wq = m[i + (size_t)n * (ck + q)];
#pragma unroll
      for (int t = 0; t < NB; ++t) r[t] = cfma(r[t], wq, sh.vals[q][t]);
    }
  }
  if (!panel) {
    if (mine)
#pragma unroll
      for (int t = 0; t < NB; ++t)
        if (t < ntj) m[i + (size_t)n * (cj + t)] = r[t];
    return;  // uniform per block
  }
  gj_reg_factor(r, sh.p, i, n, cj, mine, b, m, ipiv, moves_next + (size_t)b * kMoveInts, status);
}

// ---------------------------------------------------------------------------
// Register path with 32-column steps (n <= 768).  Each launch applies the
// rank-32 transform of one 32-column panel (two 16-column halves factored by
// the previous launch) to every other column: half the passes over the
// matrices of 16-column steps.  One block per matrix brings the next panel up
// to date and factors it in the same launch (look-ahead), as two halves in
// registers: half A, then half B after A's transform, then A after B's.
// The transform of the two halves together differs from the identity in the
// 32 panel columns only, so every other column j becomes
//   a'(:,j) = P a(:,j) with the panel rows zeroed + W (P a)(panel rows, j),
// W the panel's 32 final columns and P the panel's row interchanges (B's act
// on rows below A's panel rows only, so A's transform commutes past them).
constexpr int kSuper = 32;
constexpr int kMoveMax32 = 2 * kSuper;          // panel rows and rows they displaced
constexpr int kMoveInts32 = 1 + 2 * kMoveMax32;  // {count, dst[64], src[64]}

constexpr int kUpdCols = 32;  // columns per update workgroup of a 32-column step

struct Gj32Shared {
  GjRegShared p;
  double2 vals[kMoveMax32][kUpdCols];  // old values of the moved rows in up to 32 columns
  int dst[kMoveMax32], src[kMoveMax32];
  int nm;
};

// Loads into sh.vals the old values of the nm moved rows (sources sh.src) in
// columns [cj, cj + ntj) of a CW-column group and ends with a barrier.
template <int CW>
__device__ __forceinline__ void gj32_stage(Gj32Shared& sh, const double2* __restrict__ m, int n,
                                           int cj, int ntj, int nm) {
  for (int x = threadIdx.x; x < nm * CW; x += blockDim.x) {
    const int e = x / CW, t = x % CW;
    sh.vals[e][t] = t < ntj ? m[sh.src[e] + (size_t)n * (cj + t)] : make_double2(0.0, 0.0);
  }
  __syncthreads();
}

// Row i of columns [cj, cj + ntj) (at most CW) after the rank-W transform of
// the panel [ck, ck + W): the staged move list (sh.dst/sh.vals, its first W
// entries the panel rows in order) and the panel columns W(i, q) read from m.
template <int W, int CW>
__device__ __forceinline__ void gj32_apply(double2 (&r)[CW], const Gj32Shared& sh,
                                           const double2* __restrict__ m, int n, int ck, int cj,
                                           int ntj, int nm, int i, bool mine) {
  const bool inpanel = i >= ck && i < ck + W;
  int slot = -1;
  if (!inpanel)
    for (int e = W; e < nm; ++e)
      if (sh.dst[e] == i) slot = e;
#pragma unroll
  for (int t = 0; t < CW; ++t)
    r[t] = (!mine || inpanel || t >= ntj) ? make_double2(0.0, 0.0)
           : slot >= 0                    ? sh.vals[slot][t]
                                          : m[i + (size_t)n * (cj + t)];
  if (!mine) return;
  for (int q = 0; q < W; ++q) {
    const double2 wq = m[i + (size_t)n * (ck + q)];
#pragma unroll
    for (int t = 0; t < CW; ++t) r[t] = cfma(r[t], wq, sh.vals[q][t]);
  }
}

// The move list of rows whose source differs from themselves in map (orig or
// origb of sh.p), the nb panel rows [c0, c0 + nb) first, into sh.dst/src/nm.
__device__ __forceinline__ void gj32_list(Gj32Shared& sh, const int* map, int n, int c0, int nb,
                                          int i, bool mine) {
  if (i == 0) sh.nm = nb;
  if (i < nb) {
    sh.dst[i] = c0 + i;
    sh.src[i] = map[c0 + i];
  }
  __syncthreads();
  if (mine && (i < c0 || i >= c0 + nb) && map[i] != i) {
    const int e = atomicAdd(&sh.nm, 1);
    sh.dst[e] = i;
    sh.src[e] = map[i];
  }
  __syncthreads();
}

// Factors the 32-column panel [c, c + 32) of matrix b: half A's columns are
// in r (up to date), half B's in m (up to date).  Leaves the panel's final
// columns in m and its move list in mv.
__device__ __forceinline__ void gj32_chain(double2 (&r)[kPanelMax], Gj32Shared& sh, double2* m,
                                           int n, int c, int b, int* __restrict__ ipiv,
                                           int* __restrict__ mv, int* __restrict__ status) {
  const int i = threadIdx.x;
  const bool mine = i < n;
  if (mine) sh.p.orig[i] = sh.p.origb[i] = i;
  if (i == 0) sh.p.singular = 0;
  __syncthreads();
  // half A
  int lr = i;
  gj_reg_steps(std::make_integer_sequence<int, kPanelMax>{}, r, lr, sh.p, i, n, c, mine, b, ipiv,
               status);
  __syncthreads();
  if (sh.p.singular) {
    if (i == 0) mv[0] = -1;
    return;  // uniform
  }
  gj_reg_store(r, lr, mine, m, n, c);
  // half B after A's transform (A's moves: origb since the reset)
  gj32_list(sh, sh.p.origb, n, c, kPanelMax, i, mine);
  gj32_stage<kPanelMax>(sh, m, n, c + kPanelMax, kPanelMax, sh.nm);
  gj32_apply<kPanelMax, kPanelMax>(r, sh, m, n, c, c + kPanelMax, kPanelMax, sh.nm, i, mine);
  if (mine) sh.p.origb[i] = i;
  __syncthreads();
  lr = i;
  gj_reg_steps(std::make_integer_sequence<int, kPanelMax>{}, r, lr, sh.p, i, n, c + kPanelMax, mine,
               b, ipiv, status);
  __syncthreads();
  if (sh.p.singular) {
    if (i == 0) mv[0] = -1;
    return;  // uniform
  }
  gj_reg_store(r, lr, mine, m, n, c + kPanelMax);
  // half A after B's transform (B's moves: origb since its reset)
  gj32_list(sh, sh.p.origb, n, c + kPanelMax, kPanelMax, i, mine);
  gj32_stage<kPanelMax>(sh, m, n, c, kPanelMax, sh.nm);
  gj32_apply<kPanelMax, kPanelMax>(r, sh, m, n, c + kPanelMax, c, kPanelMax, sh.nm, i, mine);
  __syncthreads();  // every staged read of A is done before A is overwritten
  if (mine)
#pragma unroll
    for (int q = 0; q < kPanelMax; ++q) m[i + (size_t)n * (c + q)] = r[q];
  // the panel's move list: its 32 rows, then the rows they displaced
  gj32_list(sh, sh.p.orig, n, c, kSuper, i, mine);
  for (int e = i; e < sh.nm; e += blockDim.x) {
    mv[1 + e] = sh.dst[e];
    mv[1 + kMoveMax32 + e] = sh.src[e];
  }
  if (i == 0) mv[0] = sh.nm;
}

// The first panel [0, 32): one block per matrix.
__global__ void __launch_bounds__(768) k_gj_first32(double2* __restrict__ a, int n,
                                                    int* __restrict__ ipiv, int* __restrict__ moves,
                                                    int* __restrict__ status) {
  __shared__ Gj32Shared sh;
  const int b = blockIdx.x, i = threadIdx.x;
  double2* m = a + (size_t)b * n * n;
  double2 r[kPanelMax];
#pragma unroll
  for (int q = 0; q < kPanelMax; ++q) r[q] = i < n ? m[i + (size_t)n * q] : make_double2(0.0, 0.0);
  gj32_chain(r, sh, m, n, 0, b, ipiv, moves + (size_t)b * kMoveInts32, status);
}

// One 32-column step: the rank-32 transform of panel k (columns [ck, ck+32),
// move list moves_k) on every other column and, when next is set, the
// factoring of panel k+1.  Blocks [0, nbatch) (next set) are panel blocks:
// they update the 32 columns of panel k+1 (half B to HBM, half A kept in
// registers) and factor it (gj32_chain).  Every other block updates 32 other
// columns of one matrix (fewer in the last group): those outside
// [ck, ck + 32 (1 + next)), ngroups per matrix, so a block reads the panel's
// 32 columns once for 32 of its own.  The two kinds touch disjoint columns and
// read panel k's columns only.  With nbatch a multiple of 8 the update blocks
// of matrix b run on XCD b mod 8.
__global__ void __launch_bounds__(768) k_gj_step32(double2* __restrict__ a, int n, int ck, int next,
                                                   int nbatch, int ngroups, int* __restrict__ ipiv,
                                                   const int* __restrict__ moves_k,
                                                   int* __restrict__ moves_next,
                                                   int* __restrict__ status) {
  __shared__ Gj32Shared sh;
  const int i = threadIdx.x;
  const bool mine = i < n;
  const bool panel = next && (int)blockIdx.x < nbatch;
  int b, cj = 0, ntj = kPanelMax;
  if (panel) {
    b = blockIdx.x;
  } else {
    const int u = (int)blockIdx.x - (next ? nbatch : 0);
    int t;
    if (nbatch % 8 == 0) {
      const int x = u % 8, s = u / 8;
      b = x + 8 * (s / ngroups);
      t = s % ngroups;
    } else {
      b = u / ngroups;
      t = u % ngroups;
    }
    const int skip = kSuper * (1 + next);  // columns [ck, ck + skip) are not in the groups
    const int t0 = t * kUpdCols;           // groups never straddle ck (multiples of 32)
    cj = t0 < ck ? t0 : t0 + skip;
    ntj = min(kUpdCols, n - skip - t0);
  }
  double2* m = a + (size_t)b * n * n;
  const int* mv = moves_k + (size_t)b * kMoveInts32;
  const int nm = mv[0];
  if (nm < 0) {  // a zero pivot earlier (the host reports it): nothing more for this matrix
    if (panel && i == 0) moves_next[(size_t)b * kMoveInts32] = -1;
    return;  // uniform: every thread read the same word
  }
  if (i < nm) {
    sh.dst[i] = mv[1 + i];
    sh.src[i] = mv[1 + kMoveMax32 + i];
  }
  __syncthreads();
  if (!panel) {
    double2 ru[kUpdCols];
    gj32_stage<kUpdCols>(sh, m, n, cj, ntj, nm);
    gj32_apply<kSuper, kUpdCols>(ru, sh, m, n, ck, cj, ntj, nm, i, mine);
    if (mine)
#pragma unroll
      for (int t = 0; t < kUpdCols; ++t)
        if (t < ntj) m[i + (size_t)n * (cj + t)] = ru[t];
    return;  // uniform per block
  }
  double2 r[kPanelMax];
  // panel k+1: half B brought up to date into HBM, then half A into registers
  const int c = ck + kSuper;
  gj32_stage<kPanelMax>(sh, m, n, c + kPanelMax, kPanelMax, nm);
  gj32_apply<kSuper, kPanelMax>(r, sh, m, n, ck, c + kPanelMax, kPanelMax, nm, i, mine);
  __syncthreads();  // staged reads of B done before B is overwritten
  if (mine)
#pragma unroll
    for (int t = 0; t < kPanelMax; ++t) m[i + (size_t)n * (c + kPanelMax + t)] = r[t];
  gj32_stage<kPanelMax>(sh, m, n, c, kPanelMax, nm);
  gj32_apply<kSuper, kPanelMax>(r, sh, m, n, ck, c, kPanelMax, nm, i, mine);
  __syncthreads();  // every thread's B row is written before the chain stages B rows
  gj32_chain(r, sh, m, n, c, b, ipiv, moves_next + (size_t)b * kMoveInts32, status);
}

// Rank-nb update of every column outside the panel.  grid: (ceil((n-nb)/TJ),
// nbatch), 256 threads; workgroup = TJ whole columns of one matrix, so the
// reads of moved rows (staged in LDS before the barrier) never race with
// another workgroup's writes.
constexpr int kTJ = 16;
__global__ void __launch_bounds__(256) k_gj_update(double2* __restrict__ a, int n, int c0, int nb,
                                                   const int* __restrict__ moves) {
  __shared__ double2 vals[kMoveMax][kTJ];
  __shared__ int dst[kMoveMax];
  const int b = blockIdx.y, tid = threadIdx.x;
  double2* m = a + (size_t)b * n * n;
  const int* mv = moves + (size_t)b * kMoveInts;
  const int nm = mv[0];
  if (nm < 0) return;  // singular pivot: the host reports it
  const int nrest = n - nb;
  const int jt0 = blockIdx.x * kTJ;
  const int ntj = min(kTJ, nrest - jt0);
  if (tid < nm) dst[tid] = mv[1 + tid];
  for (int e = tid; e < nm * kTJ; e += blockDim.x) {
    const int mi = e / kTJ, t = e % kTJ;
    if (t < ntj) {
      const int jj = jt0 + t, j = jj < c0 ? jj : jj + nb;
      vals[mi][t] = m[mv[1 + kMoveMax + mi] + (size_t)n * j];
    }
  }
  __syncthreads();
  const double2* W = m + (size_t)n * c0;
  for (int i = tid; i < n; i += blockDim.x) {
    double2 w[kPanelMax];
#pragma unroll
    for (int q = 0; q < kPanelMax; ++q)
      if (q < nb) w[q] = W[i + (size_t)n * q];
    const bool inpanel = i >= c0 && i < c0 + nb;
    int slot = -1;
    if (!inpanel)
      for (int e = nb; e < nm; ++e)
        if (dst[e] == i) slot = e;
    for (int t = 0; t < ntj; ++t) {
      const int jj = jt0 + t, j = jj < c0 ? jj : jj + nb;
      double2* out = m + i + (size_t)n * j;
      double2 acc = inpanel ? make_double2(0.0, 0.0) : (slot >= 0 ? vals[slot][t] : *out);
#pragma unroll
      for (int q = 0; q < kPanelMax; ++q)
        if (q < nb) acc = cfma(acc, w[q], vals[q][t]);
      *out = acc;
    }
  }
}

// Undo the row interchanges: swap columns c and ipiv[c] for c = n-1 .. 0.
__global__ void __launch_bounds__(256) k_gj_colswap(double2* __restrict__ a, int n,
                                                    const int* __restrict__ ipiv) {
  const int b = blockIdx.x;
  double2* m = a + (size_t)b * n * n;
  for (int c = n - 1; c >= 0; --c) {
    const int p = ipiv[(size_t)b * n + c];
    if (p != c && p >= 0 && p < n)  // a singular matrix leaves later entries unset
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double2 t = m[i + (size_t)n * c];
        m[i + (size_t)n * c] = m[i + (size_t)n * p];
        m[i + (size_t)n * p] = t;
      }
    __syncthreads();
  }
}

// kappa[b] = max_j sum_i |a_b(i, j)| (the 1-norm of each matrix of the batch).
// grid: (ceil(n / 4), nbatch) blocks of 256: one wave per column, its lanes
// reading consecutive rows; the column sums meet in kappa[b] through a 64-bit
// atomic maximum on the bit patterns (non-negative doubles order as their
// bits).  kappa must start at zero.
__global__ void __launch_bounds__(256) k_norm1(const double2* __restrict__ a, int n,
                                               unsigned long long* __restrict__ kappa) {
  const int b = blockIdx.y, j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;  // uniform per wave
  const double2* col = a + ((size_t)b * n + j) * n;
  double s = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double2 z = col[i];
    s += sqrt(z.x * z.x + z.y * z.y);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) atomicMax(kappa + b, __double_as_longlong(s));
}

void launch_norm1(const double2* a, int n, int nbatch, double* kappa, hipStream_t s) {
  launch_zero(kappa, sizeof(double) * (size_t)nbatch, s);
  hipLaunchKernelGGL(k_norm1, dim3((n + 3) / 4, nbatch), dim3(256), 0, s, a, n,
                     (unsigned long long*)kappa);
  RP_HIP(hipGetLastError());
}

void launch_eps_build(const double2* chi, const double* qn, int n, int nbatch, int nen, int q0,
                      int g0, const double* wd, const double2* em, double delt, double2* eps,
                      hipStream_t s) {
  dim3 grid((n + 255) / 256, nbatch, n);
  hipLaunchKernelGGL(k_eps_build, grid, dim3(256), 0, s, chi, qn, n, nen, q0, g0, wd, em, delt, eps);
  RP_HIP(hipGetLastError());
}

// Two move lists per matrix: the fused step reads panel k's while its panel
// blocks write panel k+1's.
size_t gauss_jordan_work_ints(int nbatch) { return 2 * (size_t)nbatch * kMoveInts32; }

void launch_gauss_jordan(double2* a, int n, int nbatch, int* ipiv, int* work, int* status,
                         hipStream_t s) {
  // move list of panel number k: work + (k & 1) * nbatch * kMoveInts
  auto moves = [&](int k) { return work + (size_t)(k & 1) * nbatch * kMoveInts; };
  int c0 = 0, panel = 0;
  if (n <= 768 && n >= kSuper) {
    // a thread per row (at least 256: the steps stage 64 x 16 moved-row values)
    const int nt = std::max(256, (n + 63) / 64 * 64);
    auto moves32 = [&](int k) { return work + (size_t)(k & 1) * nbatch * kMoveInts32; };
    const int nfull = n / kSuper;
    hipLaunchKernelGGL(k_gj_first32, dim3(nbatch), dim3(nt), 0, s, a, n, ipiv, moves32(0), status);
    for (int k = 0; k < nfull; ++k) {
      const int next = k + 1 < nfull ? 1 : 0;
      const int nrest = n - kSuper * (1 + next);
      const int ngroups = (nrest + kUpdCols - 1) / kUpdCols;
      const int nblk = (next ? nbatch : 0) + nbatch * ngroups;
      if (nblk > 0)
        hipLaunchKernelGGL(k_gj_step32, dim3(nblk), dim3(nt), 0, s, a, n, k * kSuper, next, nbatch,
                           ngroups, ipiv, (const int*)moves32(k), moves32(k + 1), status);
    }
    c0 = nfull * kSuper;
    panel = nfull;  // the LDS path's move lists continue the alternation
  } else if (n <= 768 && n >= kPanelMax) {
    // a thread per row (at least 256: the step stages 16 x 16 pivot-row values)
    const int nt = std::max(256, (n + 63) / 64 * 64);
    // full panels in registers, each launch updating panel k and factoring k+1
    const int nfull = n / kPanelMax;
    hipLaunchKernelGGL(k_gj_panel_reg, dim3(nbatch), dim3(nt), 0, s, a, n, 0, ipiv, moves(0), status);
    for (int k = 0; k < nfull; ++k) {
      const int next = k + 1 < nfull ? 1 : 0;
      const int nrest = n - kPanelMax * (1 + next);
      const int ngroups = (nrest + kPanelMax - 1) / kPanelMax;
      const int nblk = (next ? nbatch : 0) + nbatch * ngroups;
      if (nblk > 0)
        hipLaunchKernelGGL(k_gj_step, dim3(nblk), dim3(nt), 0, s, a, n, k * kPanelMax, next,
                           nbatch, ngroups, ipiv, (const int*)moves(k), moves(k + 1), status);
    }
    c0 = panel = nfull;
    c0 *= kPanelMax;
  }
  // the remaining columns (every column when n > 768): panels in LDS, as many
  // columns as fit next to the row map, at most 16, then a rank-nb update
  const size_t budget = 160 * 1024 - 1024 - sizeof(int) * (size_t)n;
  const int nb = (int)std::min<size_t>(kPanelMax, budget / (sizeof(double2) * (size_t)n));
  RP_CHECK(nb >= 1, "Dyson: matrix too large for the LDS panel");
  int maxlds = 0;
  for (; c0 < n; ++panel) {
    const int w = std::min(nb, n - c0);
    const size_t lds = sizeof(double2) * (size_t)w * n + sizeof(int) * (size_t)n;
    if ((int)lds > maxlds && lds > 65536) {
      RP_HIP(hipFuncSetAttribute((const void*)k_gj_panel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      maxlds = (int)lds;
    }
    hipLaunchKernelGGL(k_gj_panel, dim3(nbatch), dim3(1024), lds, s, a, n, c0, w, ipiv,
                       moves(panel), status);
    if (n - w > 0)
      hipLaunchKernelGGL(k_gj_update, dim3((n - w + kTJ - 1) / kTJ, nbatch), dim3(256), 0, s, a, n,
                         c0, w, (const int*)moves(panel));
    c0 += w;
  }
  hipLaunchKernelGGL(k_gj_colswap, dim3(nbatch), dim3(256), 0, s, a, n, (const int*)ipiv);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
