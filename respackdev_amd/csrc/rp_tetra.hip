// Generalized-tetrahedron weights x_ij(k, w) of ttrhdrn_chi0
// (m_tetrahedron_20170325.F:99-503) for every band pair, in FP64 on gfx950.
//
// One workgroup per (pair, slice of nf frequencies).  Its k-grid accumulator
// x[nk3][nf] (complex FP64) lives in LDS.  The workgroup first classifies its
// 2 x 6 NTK (tetrahedron, dsgn) items (chk1) into two LDS lists, non-degenerate
// items first, so that the waves of the second pass run one closed form
// without divergence.  For a non-degenerate tetrahedron the frequency-
// independent factors (reciprocal vertex differences) are formed once and each
// frequency needs only the four vertex logarithms: Log(a/b) = Log a - Log b,
// brought back to the principal branch.  Degenerate items (rare) use the
// reference's closed forms as written.  Corner contributions go to LDS with
// FP64 atomics.  Tetrahedra cut by E_F (metals: partially occupied bands)
// are left to k_tetra_cut, which runs after this kernel on the same output.
//
// Each closed form is the exact integral W_v = int lambda_v / (e + g(lambda))
// over the unit simplex for its class's nodes (near-equal vertex values
// replaced by their mean).  Where the nodes are close together against
// |e + c| (c their mean) the closed forms cancel: their terms are ~24 r^-3
// times W, r = max_j |g_j - c| / |e + c|, so at high frequency a 1-ulp error
// of a logarithm becomes 1e-5 of W on small Fermi-surface pieces and more on
// tiny ones.  For r < kSeriesR every kernel here evaluates the same integral
// by its convergent series instead (series_w), which loses nothing.
//
// The output is the contraction weight d[w][K] = x(k,w) (1 - P_occ P_vir)/NTK
// with K = pair*NTK + k (the factors of chiqw.F:1404-1406).
#include <algorithm>

#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

// ---------------------------------------------------------- closed forms
// Fujiwara-Yamamoto-Ishii closed forms as evaluated in
// m_tetrahedron_20170325.F:304-451, one frequency e at a time.
__device__ inline zd sq(zd a) { return a * a; }
__device__ inline zd cu(zd a) { return a * a * a; }

__device__ __forceinline__ zd f_sub1g30g21(zd g0, zd g1, zd e) {
  zd d = g0 - g1;
  zd x = (0.5 * (g0 + g1) + e) / sq(d);
  return x - logz((g0 + e) / (g1 + e)) * (g0 + e) * (g1 + e) / cu(d);
}
__device__ __forceinline__ zd f_sub1g30g20(zd g0, zd g1, zd e) {
  zd d = g0 - g1;
  zd x = (g0 - 3.0 * g1 - 2.0 * e) * 0.25 / sq(d);
  return x + sq(g1 + e) * logz((g0 + e) / (g1 + e)) * 0.5 / cu(d);
}
__device__ __forceinline__ zd f_sub1g30(zd g0, zd g1, zd g2, zd e) {
  zd x = 0.5 / (g0 - g2);
  zd p = (g0 - g2) * (g0 - g1);
  x = x - 0.5 / sq(p) * (g0 + e) * (-2.0 * (g1 + e) * (g2 + e) + (g0 + e) * (g1 + g2 + 2.0 * e)) *
              logz((g0 + e) / (g1 + e));
  x = x + 0.5 / sq(g0 - g2) * sq(g2 + e) * logz((g1 + e) / (g2 + e)) / (g1 - g2);
  return x + 0.5 / ((g0 - g2) * (g0 - g1)) * (g1 + e);
}
__device__ __forceinline__ zd f_sub1normal(zd g0, zd g1, zd g2, zd g3, zd e) {
  zd l1 = sq(g1 + e) * logz((g1 + e) / (g0 + e));
  zd x = (-l1) / ((g0 - g2) * (g0 - g3) * (g0 - g1)) * 0.5;
  x = x - (l1 - sq(g2 + e) * logz((g2 + e) / (g0 + e))) / ((g0 - g2) * (g1 - g2) * (g2 - g3)) * 0.5;
  return x + (l1 - sq(g3 + e) * logz((g3 + e) / (g0 + e))) / ((g1 - g3) * (g0 - g3) * (g2 - g3)) * 0.5;
}
__device__ __forceinline__ zd f_sub2simple(zd g0, zd e) { return 1.0 / ((g0 + e) * 24.0); }
__device__ __forceinline__ zd f_sub2g30g20(zd g0, zd g1, zd e) {
  zd d = g0 - g1, d2 = d * d;
  zd x = (d * (g0 + 2.0 * g1 + 3.0 * e) - 6.0 * (g0 + e) * (g1 + e)) / (12.0 * (d2 * d));
  return x + (g0 + e) * sq(g1 + e) * logz((g0 + e) / (g1 + e)) / (d2 * d2) * 0.5;
}
__device__ __forceinline__ zd f_sub2g30g10(zd g0, zd g2, zd e) {
  zd a = g0 + e, c = g2 + e, d = g0 - g2, d2 = d * d;
  zd x = (2.0 * sq(a) - 7.0 * a * c + 11.0 * sq(c)) / (36.0 * (d2 * d));
  return x - cu(c) * logz(a / c) / (6.0 * (d2 * d2));
}
__device__ __forceinline__ zd f_sub2g10(zd g0, zd g2, zd g3, zd e) {
  zd a = g0 + e, c2 = g2 + e, c3 = g3 + e;
  zd x = a * (sq(a) + 5.0 * c2 * c3 - 3.0 * a * (g2 + g3 + 2.0 * e)) /
         (12.0 * sq(g0 - g2) * sq(g0 - g3));
  x = x - cu(c2) * logz(c2 / a) / (6.0 * cu(g0 - g2) * (g2 - g3));
  return x + cu(c3) * logz(c3 / a) / (6.0 * cu(g0 - g3) * (g2 - g3));
}
__device__ __forceinline__ zd f_sub2g30(zd g0, zd g1, zd g2, zd e) {
  zd a = g0 + e, b = g1 + e, c = g2 + e;
  zd d01 = g0 - g1, d02 = g0 - g2, d12 = g1 - g2;
  zd x = (sq(a) * d12 + sq(b) * d02) / (6.0 * d02 * sq(d01) * d12);
  x = x - sq(a) * (2.0 * b * d02 + c * d01) * logz(a / c) / (6.0 * sq(d02) * cu(d01));
  return x + sq(b) * (2.0 * a * d12 - c * d01) * logz(b / c) / (6.0 * cu(d01) * sq(d12));
}
__device__ __forceinline__ zd f_sub2normal(zd g0, zd g1, zd g2, zd g3, zd e) {
  zd a = g0 + e, b = g1 + e, c = g2 + e, d = g3 + e;
  zd x = -sq(b) / (6.0 * (g0 - g1) * (g1 - g2) * (g1 - g3));
  x = x - cu(c) * logz(c / b) / (6.0 * (g0 - g2) * sq(g1 - g2) * (g2 - g3));
  x = x + cu(d) * logz(d / b) / (6.0 * (g0 - g3) * sq(g1 - g3) * (g2 - g3));
  return x + cu(a) * logz(a / b) / (6.0 * sq(g0 - g1) * (g0 - g2) * (g0 - g3));
}

// chk1 (m_tetrahedron_20170325.F:457-503): vertices compared in the
// reference's rotated order p = mod(i,4)+1.  Returns the packed class:
// bits 0-3 kind, 4-11 the degenerate vertices a3 (2 bits each, in discovery
// order), 12-19 the others a4.
enum : int { kNormal = 0, kPair = 1, kTriple = 2, kTwoPairs = 3, kAll = 4, kBad = 5 };

__device__ __forceinline__ int classify(double dmna, double dmnr, const zd (&g)[4]) {
  unsigned ia2 = 0, a3 = 0, a4 = 0;
  int i3 = 0, i4 = 0;
#pragma unroll
  for (int i1 = 1; i1 <= 3; ++i1)
#pragma unroll
    for (int i2 = i1 + 1; i2 <= 4; ++i2) {
      const int p = i1 & 3, q = i2 & 3;
      const double d1 = absz(g[p] - g[q]);
      const double d2 = absz(g[p] + g[q]) * 0.5;
      if (d1 < dmna || d1 / d2 < dmnr) {
        ++i3;
        if (!((ia2 >> p) & 1u)) { a3 |= (unsigned)p << (2 * i4); ++i4; ia2 |= 1u << p; }
        if (!((ia2 >> q) & 1u)) { a3 |= (unsigned)q << (2 * i4); ++i4; ia2 |= 1u << q; }
      }
    }
  int i5 = 0;
#pragma unroll
  for (int i1 = 1; i1 <= 4; ++i1) {
    const int p = i1 & 3;
    if (!((ia2 >> p) & 1u)) { a4 |= (unsigned)p << (2 * i5); ++i5; }
  }
  int kind;
  switch (i4) {
    case 0: kind = kNormal; break;
    case 2: kind = kPair; break;
    case 3: kind = kTriple; break;
    case 4: kind = i3 == 2 ? kTwoPairs : (i3 >= 3 && i3 <= 6 ? kAll : kBad); break;
    default: kind = kBad; break;
  }
  return kind | (int)(a3 << 4) | (int)(a4 << 12);
}
__device__ __forceinline__ int vtx(int code, int set, int i) {  // set 0: a3, 1: a4
  return (code >> (4 + 8 * set + 2 * i)) & 3;
}
// v_i for a data-dependent i in 0..3, as selects on scalars (a select of
// aggregates would be lowered to an indexed private-memory array).
__device__ __forceinline__ double sel4d(int i, double v0, double v1, double v2, double v3) {
  double r = v0;
  r = i == 1 ? v1 : r;
  r = i == 2 ? v2 : r;
  r = i == 3 ? v3 : r;
  return r;
}
__device__ __forceinline__ int sel4(int i, int v0, int v1, int v2, int v3) {
  int r = v0;
  r = i == 1 ? v1 : r;
  r = i == 2 ? v2 : r;
  r = i == 3 ? v3 : r;
  return r;
}
__device__ __forceinline__ zd sel4(int i, zd v0, zd v1, zd v2, zd v3) {
  return mkz(sel4d(i, v0.x, v1.x, v2.x, v3.x), sel4d(i, v0.y, v1.y, v2.y, v3.y));
}

// The vertex weights of nodes g0..g3 (one shared imaginary part, the
// broadening) at frequency e by the series of the integral around the mean
// node c: with s = e + c and d_j = Re g_j - Re c,
//   W_v = (1/s) sum_n (-1/s)^n n!/(n+4)! P_{v,n},
//   P_{v,n} = [z^n] prod_j 1/(1 - d_j z) * 1/(1 - d_v z)
// (the Dirichlet moments of the simplex).  Term n is at most r^n / (24 |s|),
// r = max_j |d_j| / |s|: the sum runs until r^n < 2^-54 and loses nothing to
// cancellation.  Returns false, outputs untouched, for r >= kSeriesR, where
// the closed forms lose at most ~24 r^-3 = 200 ulps.
constexpr double kSeriesR = 0.5;
constexpr int kSeriesMax = 64;  // terms at r just below kSeriesR: 55
// the last index n of the sum for ratio r (r^n < 2^-54)
__device__ __forceinline__ int series_terms(double r) {
  return r > 0.0 ? min(kSeriesMax - 1, (int)ceil(-37.42994775023705 / log(r))) : 0;
}
__constant__ double kMomentInv[kSeriesMax] = {
#define RP_MI(n) 1.0 / ((double)((n) + 1) * ((n) + 2) * ((n) + 3) * ((n) + 4))
    RP_MI(0),  RP_MI(1),  RP_MI(2),  RP_MI(3),  RP_MI(4),  RP_MI(5),  RP_MI(6),  RP_MI(7),
    RP_MI(8),  RP_MI(9),  RP_MI(10), RP_MI(11), RP_MI(12), RP_MI(13), RP_MI(14), RP_MI(15),
    RP_MI(16), RP_MI(17), RP_MI(18), RP_MI(19), RP_MI(20), RP_MI(21), RP_MI(22), RP_MI(23),
    RP_MI(24), RP_MI(25), RP_MI(26), RP_MI(27), RP_MI(28), RP_MI(29), RP_MI(30), RP_MI(31),
    RP_MI(32), RP_MI(33), RP_MI(34), RP_MI(35), RP_MI(36), RP_MI(37), RP_MI(38), RP_MI(39),
    RP_MI(40), RP_MI(41), RP_MI(42), RP_MI(43), RP_MI(44), RP_MI(45), RP_MI(46), RP_MI(47),
    RP_MI(48), RP_MI(49), RP_MI(50), RP_MI(51), RP_MI(52), RP_MI(53), RP_MI(54), RP_MI(55),
    RP_MI(56), RP_MI(57), RP_MI(58), RP_MI(59), RP_MI(60), RP_MI(61), RP_MI(62), RP_MI(63),
#undef RP_MI
};
// The sweep over non-degenerate tetrahedra takes the series only below
// kFastR (at most 37 terms; the closed forms lose at most ~24 r^-3 = 560 ulps
// above it) and counts terms from r^2 in single precision: r^n < 2^-54 for
// n > 108 / -log2(r^2), with a quarter term of margin for the rounding.
constexpr double kFastR = 0.35;
constexpr double kFastInvR2 = 1.0 / (kFastR * kFastR);
__device__ __forceinline__ int fast_terms(double r2) {
  const float lg = __log2f((float)r2);
  return r2 > 0.0 ? min(kSeriesMax - 1, (int)ceilf(-108.0f / lg + 0.25f)) : 0;
}
// Both items of a tetrahedron by one series: the nodes of dsgn = -1 are the
// negated nodes of dsgn = +1, so d_j and the mean change sign and P_{v,n}
// takes (-1)^n; one set of recurrences serves both, and the two items' terms
// (the second taken with its factor dsgn = -1) sum to
//   T_v = W_v(+) - W_v(-) = sum c_n P_{v,n} ((1/s+)(-1/s+)^n - (1/s-)^(n+1))
//       = -sum c_n P_{v,n} (u^(n+1) + v^(n+1)),  u = -1/s+, v = 1/s-,
// with s+- = e +- c and |s+-|^2 = ap, am (u, v formed as conj(s)/|s|^2).
// r^2 = dm^2 / min(ap, am) sets the terms.
__device__ __forceinline__ void joint_series(double d0, double d1, double d2, double d3, double dm,
                                             zd sp, zd sm, double ap, double am, zd& T0, zd& T1,
                                             zd& T2, zd& T3) {
  const int nmax = fast_terms(dm * dm / fmin(ap, am));
  const double q = 1.0 / (ap * am), ip = am * q, im = ap * q;
  const zd u = mkz(-sp.x * ip, sp.y * ip), v = mkz(sm.x * im, -sm.y * im);
  zd U = u, V = v;
  T0 = T1 = T2 = T3 = mkz(0.0, 0.0);
  double h1 = 1.0, h2 = 1.0, h3 = 1.0, h = 1.0;
  double p0 = 1.0, p1 = 1.0, p2 = 1.0, p3 = 1.0;
  for (int n = 0; n <= nmax; ++n) {
    const double cf = -kMomentInv[n];
    const zd W = (U + V) * cf;
    T0 = T0 + W * p0;
    T1 = T1 + W * p1;
    T2 = T2 + W * p2;
    T3 = T3 + W * p3;
    U = U * u;
    V = V * v;
    h1 *= d0;
    h2 = h1 + d1 * h2;
    h3 = h2 + d2 * h3;
    h = h3 + d3 * h;
    p0 = h + d0 * p0;
    p1 = h + d1 * p1;
    p2 = h + d2 * p2;
    p3 = h + d3 * p3;
  }
}

__device__ __forceinline__ bool series_w(zd g0, zd g1, zd g2, zd g3, zd e, zd& w0, zd& w1,
                                         zd& w2, zd& w3, double rmax = kSeriesR) {
  const zd c = (g0 + g1 + g2 + g3) * 0.25;
  const zd s = e + c;
  const double d0 = g0.x - c.x, d1 = g1.x - c.x, d2 = g2.x - c.x, d3 = g3.x - c.x;
  const double dm = fmax(fmax(fabs(d0), fabs(d1)), fmax(fabs(d2), fabs(d3)));
  const double as = absz(s);
  if (!(dm < rmax * as)) return false;
  const int nmax = series_terms(dm / as);
  const zd t = -1.0 / s;
  zd tn = 1.0 / s;
  double h1 = 1.0, h2 = 1.0, h3 = 1.0, h = 1.0;
  double p0 = 1.0, p1 = 1.0, p2 = 1.0, p3 = 1.0;
  zd a0 = mkz(0.0, 0.0), a1 = a0, a2 = a0, a3 = a0;
  for (int n = 0; n <= nmax; ++n) {
    const double cf = kMomentInv[n];
    a0 = a0 + tn * (cf * p0);
    a1 = a1 + tn * (cf * p1);
    a2 = a2 + tn * (cf * p2);
    a3 = a3 + tn * (cf * p3);
    tn = tn * t;
    h1 *= d0;
    h2 = h1 + d1 * h2;
    h3 = h2 + d2 * h3;
    h = h3 + d3 * h;
    p0 = h + d0 * p0;
    p1 = h + d1 * p1;
    p2 = h + d2 * p2;
    p3 = h + d3 * p3;
  }
  w0 = a0;
  w1 = a1;
  w2 = a2;
  w3 = a3;
  return true;
}

// Log(a/b) from Log a and Log b, on the principal branch (-pi, pi].
__device__ __forceinline__ zd logdiff(zd la, zd lb) {
  zd d = la - lb;
  if (d.y > M_PI) d.y -= 2.0 * M_PI;
  else if (d.y <= -M_PI) d.y += 2.0 * M_PI;
  return d;
}

__device__ inline void lds_add(double2* p, zd v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
}

// Does ttct need to split?  (Some pair of vertices strictly on both sides.)
__device__ inline bool straddles(double f, double v0, double v1, double v2, double v3) {
  // all six vertex pairs (the order does not matter for "any")
  const double d0 = v0 - f, d1 = v1 - f, d2 = v2 - f, d3 = v3 - f;
  return (d0 * -d1 > 0.0) || (d0 * -d2 > 0.0) || (d0 * -d3 > 0.0) || (d1 * -d2 > 0.0) ||
         (d1 * -d3 > 0.0) || (d2 * -d3 > 0.0);
}

// Item it < 2 ntet: tetrahedron it % ntet, dsgn = +1 for the first half.
// Vertex values g_v = -(e2 - e1 - i delt) dsgn (chiqw.F:1374-1383 with f_v = dsgn).
struct Item {
  int c[4];
  zd g[4];
  double dsgn;
};
__device__ __forceinline__ Item load_item(int it, int ntet, const int* __restrict__ imt,
                                          const double* e1, const double* e2, double delt) {
  Item m;
  const int t = it % ntet;
  m.dsgn = it < ntet ? 1.0 : -1.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    m.c[v] = imt[4 * t + v];
    m.g[v] = mkz(-(e2[m.c[v]] - e1[m.c[v]]), delt) * m.dsgn;
  }
  return m;
}

// grid: (npair, ceil(nen / nf)) blocks.  eband[band][k] energies (Ha, full
// mesh, FP64); the tetrahedron sees them rounded to single precision (CMPLX,
// chiqw.F:1353,1356).  Tetrahedra cut by E_F are skipped when cuts_elsewhere
// (k_tetra_cut adds them), else they raise status bit 1; a class the
// reference rejects (errexec) raises bit 2.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_tetra_weights(
    int ntk, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const double* __restrict__ eband,
    const double2* __restrict__ pband, double ef, double delt, double dmna, double dmnr, int nen,
    int nf, const double2* __restrict__ em, size_t ldd, int cuts_elsewhere, int qcap,
    int* __restrict__ status, double2* __restrict__ dout) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  const int ntet = 6 * nk3;
  double2* x = lds;                               // [nk3][nf]
  double2* lv = x + (size_t)nk3 * nf;             // [nk3][nf][2] log(s A_k + e_f), s = +1, -1
  double* e1 = (double*)(lv + (size_t)nk3 * nf * 2);  // [nk3]
  double* e2 = e1 + nk3;                          // [nk3]
  int2* list = (int2*)(e2 + nk3);                 // [ntet] {tetrahedron, class}
  int* queue = (int*)(list + ntet);               // [qcap] (item, frequency) for the closed forms
  __shared__ int count[3];
  const int p = blockIdx.x;
  const int f0 = blockIdx.y * nf, nfb = min(nf, nen - f0);
  const int jo = pocc[p], iv = pvir[p];
  for (int i = threadIdx.x; i < nk3 * nf; i += blockDim.x) x[i] = make_double2(0, 0);
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    const int k = index_kpt[gi];
    e1[gi] = (double)(float)eband[(size_t)jo * ntk + k];
    e2[gi] = (double)(float)eband[(size_t)iv * ntk + ikq[k]];
  }
  if (threadIdx.x < 3) count[threadIdx.x] = 0;
  __syncthreads();
  // ---- pass 1: cut test (ttct), keep test, class, per tetrahedron: its two
  // items (dsgn = +1 and -1) have vertex values g and -g, hence the same cut
  // and keep tests and the same class (chk1 compares |g_p - g_q| and
  // |g_p + g_q|); the list holds the dsgn = +1 item
  for (int it = threadIdx.x; it < ntet; it += blockDim.x) {
    const int t = it;
    const int c0 = imt[4 * t], c1 = imt[4 * t + 1], c2 = imt[4 * t + 2], c3 = imt[4 * t + 3];
    const double a0 = e1[c0], a1 = e1[c1], a2 = e1[c2], a3 = e1[c3];
    const double b0 = e2[c0], b1 = e2[c1], b2 = e2[c2], b3 = e2[c3];
    if (straddles(ef, b0, b1, b2, b3) || straddles(ef, a0, a1, a2, a3)) {
      if (!cuts_elsewhere) atomicOr(status, 1);
      continue;
    }
    // ttct keeps the tetrahedron iff some e2 vertex is above and some e1 vertex below E_F
    const bool k2 = (b0 - ef > 0.0) || (b1 - ef > 0.0) || (b2 - ef > 0.0) || (b3 - ef > 0.0);
    const bool k1 = (-(a0 - ef) > 0.0) || (-(a1 - ef) > 0.0) || (-(a2 - ef) > 0.0) || (-(a3 - ef) > 0.0);
    if (!(k1 && k2)) continue;
    const Item m = load_item(it, ntet, imt, e1, e2, delt);
    const int code = classify(dmna, dmnr, m.g);
    const int kind = code & 15;
    if (kind == kBad) {
      atomicOr(status, 2);
      continue;
    }
    const int slot = kind == kNormal ? atomicAdd(&count[0], 1) : ntet - 1 - atomicAdd(&count[1], 1);
    list[slot] = make_int2(it, code);
  }
  // the vertex logarithms of the non-degenerate closed forms: a vertex value
  // depends on the k point alone, and each k is a corner of 24 tetrahedra
  for (int i = threadIdx.x; i < nk3 * nfb * 2; i += blockDim.x) {
    const int k = i / (2 * nfb), f = (i >> 1) % nfb;
    const double sg = (i & 1) ? -1.0 : 1.0;
    const double2 ev = em[f0 + f];
    const zd v = mkz(-(e2[k] - e1[k]), delt) * sg + mkz(ev.x, ev.y);
    const zd L = logz(v);
    lv[((size_t)k * nf + f) * 2 + (i & 1)] = make_double2(L.x, L.y);
  }
  __syncthreads();
  const int nnormal = count[0], nother = count[1];
  // ---- pass 2: non-degenerate tetrahedra (calc, case(0); ttrhdrn_sub2normal x3 and
  // ttrhdrn_sub1normal, m_tetrahedron_20170325.F:215-229,431-451)
  // Closed forms of one (item j, frequency f): per sign the series where its
  // ratio is below kFastR, else the closed forms from the vertex logarithms.
  auto closed_unit = [&](int j, int f) {
    const int2 ent = list[j];
    const Item m = load_item(ent.x, ntet, imt, e1, e2, delt);
    const int code = ent.y;
    const int i0 = vtx(code, 1, 0), i1 = vtx(code, 1, 1), i2 = vtx(code, 1, 2), i3 = vtx(code, 1, 3);
    const zd A0 = sel4(i0, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A1 = sel4(i1, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A2 = sel4(i2, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A3 = sel4(i3, m.g[0], m.g[1], m.g[2], m.g[3]);
    const int k0 = sel4(i0, m.c[0], m.c[1], m.c[2], m.c[3]);
    const int k1 = sel4(i1, m.c[0], m.c[1], m.c[2], m.c[3]);
    const int k2 = sel4(i2, m.c[0], m.c[1], m.c[2], m.c[3]);
    const int k3 = sel4(i3, m.c[0], m.c[1], m.c[2], m.c[3]);
    const double2* l0 = lv + (size_t)k0 * nf * 2;
    const double2* l1 = lv + (size_t)k1 * nf * 2;
    const double2* l2 = lv + (size_t)k2 * nf * 2;
    const double2* l3 = lv + (size_t)k3 * nf * 2;
    // r_ij = 1/(A_i - A_j); P0..P3 the triple products around each vertex
    const zd r01 = 1.0 / (A0 - A1), r02 = 1.0 / (A0 - A2), r03 = 1.0 / (A0 - A3);
    const zd r12 = 1.0 / (A1 - A2), r13 = 1.0 / (A1 - A3), r23 = 1.0 / (A2 - A3);
    const zd P0 = r01 * r02 * r03, P1 = r01 * r12 * r13, P2 = r02 * r12 * r23,
             P3 = r03 * r13 * r23;
    const double s6 = 1.0 / 6.0;
    // sub2normal(A1,A0,A2,A3): 1/(6(A1-A0)(A0-A2)(A0-A3)), 1/(6(A1-A2)(A0-A2)^2(A2-A3)), ...
    const zd K11 = P0 * -s6, K12 = r02 * P2 * s6, K13 = r03 * P3 * s6, K14 = r01 * P1 * s6;
    // sub2normal(A0,A1,A2,A3)
    const zd K21 = P1 * s6, K22 = r12 * P2 * s6, K23 = r13 * P3 * s6, K24 = r01 * P0 * s6;
    // sub2normal(A0,A2,A1,A3)
    const zd K31 = P2 * -s6, K32 = r12 * P1 * s6, K33 = r23 * P3 * s6, K34 = r02 * P0 * s6;
    // sub1normal(A0,A3,A1,A2)
    const zd J1 = P0 * 0.5, J2 = P1 * -0.5, J3 = P2 * -0.5;
    // Both items of the tetrahedron: dsgn = s takes the vertex values s*A, so
    // every r_ij takes the factor s: the factors of three of them (K11, K21,
    // K31, J1..J3) change sign, those of four do not (exactly, in IEEE
    // arithmetic).  Their corner terms are summed before the LDS atomics.
    const double2 ev = em[f0 + f];
    const zd e = mkz(ev.x, ev.y);
    zd T0 = mkz(0.0, 0.0), T1 = mkz(0.0, 0.0), T2 = mkz(0.0, 0.0), T3 = mkz(0.0, 0.0);
#pragma unroll
    for (int sg = 0; sg < 2; ++sg) {
      const double s = sg ? -1.0 : 1.0;
      {
        zd w0, w1, w2, w3;
        if (series_w(A0 * s, A1 * s, A2 * s, A3 * s, e, w0, w1, w2, w3, kFastR)) {
          T0 = T0 + w0 * s;
          T1 = T1 + w1 * s;
          T2 = T2 + w2 * s;
          T3 = T3 + w3 * s;
          continue;
        }
      }
      const zd v0 = A0 * s + e, v1 = A1 * s + e, v2 = A2 * s + e, v3 = A3 * s + e;
      // log(v_i), v_i = s A_i + e, from the vertex table (the same operations)
      const double2 w0 = l0[2 * f + sg], w1 = l1[2 * f + sg], w2 = l2[2 * f + sg],
                    w3 = l3[2 * f + sg];
      const zd L0 = mkz(w0.x, w0.y), L1 = mkz(w1.x, w1.y), L2 = mkz(w2.x, w2.y),
               L3 = mkz(w3.x, w3.y);
      const zd D10 = logdiff(L1, L0), D20 = logdiff(L2, L0), D30 = logdiff(L3, L0);
      const zd D21 = logdiff(L2, L1), D31 = logdiff(L3, L1), D32 = logdiff(L3, L2);
      const zd q0 = v0 * v0, q1 = v1 * v1, q2 = v2 * v2, q3 = v3 * v3;
      const zd u0 = q0 * v0, u1 = q1 * v1, u2 = q2 * v2, u3 = q3 * v3;
      zd c1 = -q0 * (K11 * s);
      c1 = c1 - u2 * D20 * K12;
      c1 = c1 + u3 * D30 * K13;
      c1 = c1 + u1 * D10 * K14;
      zd c2 = -q1 * (K21 * s);
      c2 = c2 - u2 * D21 * K22;
      c2 = c2 + u3 * D31 * K23;
      c2 = c2 - u0 * D10 * K24;
      zd c3 = -q2 * (K31 * s);
      c3 = c3 + u1 * D21 * K32;
      c3 = c3 + u3 * D32 * K33;
      c3 = c3 - u0 * D20 * K34;
      const zd l1 = q3 * D30;
      zd c4 = -l1 * (J1 * s);
      c4 = c4 - (l1 - q1 * D10) * (J2 * s);
      c4 = c4 + (l1 - q2 * D20) * (J3 * s);
      T0 = T0 + c1 * s;
      T1 = T1 + c2 * s;
      T2 = T2 + c3 * s;
      T3 = T3 + (c4 - c3 - c2 - c1) * s;
    }
    lds_add(x + (size_t)k0 * nf + f, T0);
    lds_add(x + (size_t)k1 * nf + f, T1);
    lds_add(x + (size_t)k2 * nf + f, T2);
    lds_add(x + (size_t)k3 * nf + f, T3);
  };
  // The sweep: work units (item, chunk of frequencies); with few tetrahedra
  // (small k meshes) each item's frequencies are split so that about twice
  // the block's threads have work.  At a frequency where both signs' ratios
  // are below kFastR (the usual case) the two series are summed here; the
  // other (item, frequency) pairs are queued for closed_unit after the sweep
  // (inline when the queue is full), so that the lanes of a wave do not wait
  // on the closed forms of a few.
  const int fsplit = max(1, min(nfb, (2 * (int)blockDim.x + max(nnormal, 1) - 1) / max(nnormal, 1)));
  const int fchunk = (nfb + fsplit - 1) / fsplit;
  for (int jf = threadIdx.x; jf < nnormal * fsplit; jf += blockDim.x) {
    const int j = jf / fsplit, fa = (jf % fsplit) * fchunk, fb = min(nfb, fa + fchunk);
    const int2 ent = list[j];
    const Item m = load_item(ent.x, ntet, imt, e1, e2, delt);
    const int code = ent.y;
    const int i0 = vtx(code, 1, 0), i1 = vtx(code, 1, 1), i2 = vtx(code, 1, 2), i3 = vtx(code, 1, 3);
    const zd A0 = sel4(i0, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A1 = sel4(i1, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A2 = sel4(i2, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd A3 = sel4(i3, m.g[0], m.g[1], m.g[2], m.g[3]);
    double2* x0 = x + (size_t)sel4(i0, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* x1 = x + (size_t)sel4(i1, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* x2 = x + (size_t)sel4(i2, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* x3 = x + (size_t)sel4(i3, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    // the series' nodes for dsgn = +1 (series_w's operations): mean sc, offsets sd
    const zd sc = (A0 + A1 + A2 + A3) * 0.25;
    const double sd0 = A0.x - sc.x, sd1 = A1.x - sc.x, sd2 = A2.x - sc.x, sd3 = A3.x - sc.x;
    const double sdm = fmax(fmax(fabs(sd0), fabs(sd1)), fmax(fabs(sd2), fabs(sd3)));
    const double lim = sdm * sdm * kFastInvR2;  // r < kFastR  <=>  |s|^2 > lim
#pragma unroll 1
    for (int f = fa; f < fb; ++f) {
      const double2 ev = em[f0 + f];
      // both signs by the series (joint_series) where both ratios are below kFastR
      const zd sp = mkz(ev.x + sc.x, ev.y + sc.y), sm = mkz(ev.x - sc.x, ev.y - sc.y);
      const double ap = sp.x * sp.x + sp.y * sp.y, am = sm.x * sm.x + sm.y * sm.y;
      if (!(lim < ap && lim < am)) {
        const int slot = atomicAdd(&count[2], 1);
        if (slot < qcap) queue[slot] = j * nf + f;
        else closed_unit(j, f);
        continue;
      }
      zd T0, T1, T2, T3;
      joint_series(sd0, sd1, sd2, sd3, sdm, sp, sm, ap, am, T0, T1, T2, T3);
      lds_add(x0 + f, T0);
      lds_add(x1 + f, T1);
      lds_add(x2 + f, T2);
      lds_add(x3 + f, T3);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < min(count[2], qcap); i += blockDim.x) closed_unit(queue[i] / nf, queue[i] % nf);
  // ---- pass 3: degenerate tetrahedra (calc, m_tetrahedron_20170325.F:231-300)
  // Work units (entry, chunk of frequencies) as in the sweep, both items of
  // the entry (dsgn = +1 and -1) per unit.
  const int fsplit3 = max(1, min(nfb, (2 * (int)blockDim.x + max(nother, 1) - 1) / max(nother, 1)));
  const int fchunk3 = (nfb + fsplit3 - 1) / fsplit3;
  for (int u3 = threadIdx.x; u3 < nother * fsplit3; u3 += blockDim.x) {
    const int j = u3 / fsplit3, fa = (u3 % fsplit3) * fchunk3, fb = min(nfb, fa + fchunk3);
    const int2 ent = list[ntet - 1 - j];
    const Item m = load_item(ent.x, ntet, imt, e1, e2, delt);  // dsgn = +1
    const int code = ent.y, kind = code & 15;
    const int b0 = vtx(code, 0, 0), b1 = vtx(code, 0, 1), b2 = vtx(code, 0, 2), b3 = vtx(code, 0, 3);
    const int o0 = vtx(code, 1, 0), o1 = vtx(code, 1, 1);
    const zd G30 = sel4(b0, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd G31 = sel4(b1, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd G32 = sel4(b2, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd G33 = sel4(b3, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd G40 = sel4(o0, m.g[0], m.g[1], m.g[2], m.g[3]);
    const zd G41 = sel4(o1, m.g[0], m.g[1], m.g[2], m.g[3]);
    double2* y30 = x + (size_t)sel4(b0, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* y31 = x + (size_t)sel4(b1, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* y32 = x + (size_t)sel4(b2, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* y33 = x + (size_t)sel4(b3, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* y40 = x + (size_t)sel4(o0, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    double2* y41 = x + (size_t)sel4(o1, m.c[0], m.c[1], m.c[2], m.c[3]) * nf;
    // the class's nodes (near-equal values replaced by their mean) and the
    // corners they belong to, for the series; kAll's closed form is exact
    zd n2 = G40, n3 = G41;
    double2* z2 = y40;
    double2* z3 = y41;
    zd m1 = (G30 + G31) / 2.0;
    if (kind == kTriple) {
      m1 = (G30 + G31 + G32) / 3.0;
      n2 = m1;
      n3 = G40;
      z2 = y32;
      z3 = y40;
    } else if (kind == kTwoPairs) {
      n2 = n3 = (G32 + G33) / 2.0;
      z2 = y32;
      z3 = y33;
    }
    // series_w's nodes, mean and offsets for dsgn = +1; the series where the
    // ratio is below kSeriesR (both signs at once where both are)
    const zd sc = (m1 + m1 + n2 + n3) * 0.25;
    const double sd0 = m1.x - sc.x, sd2 = n2.x - sc.x, sd3 = n3.x - sc.x;
    const double sdm = fmax(fabs(sd0), fmax(fabs(sd2), fabs(sd3)));
    const double lim = kind == kAll ? INFINITY : sdm * sdm * (1.0 / (kSeriesR * kSeriesR));
#pragma unroll 1
    for (int f = fa; f < fb; ++f) {
      const double2 ev = em[f0 + f];
      const zd e = mkz(ev.x, ev.y);
      const zd sp = mkz(ev.x + sc.x, ev.y + sc.y), sm = mkz(ev.x - sc.x, ev.y - sc.y);
      const double ap = sp.x * sp.x + sp.y * sp.y, am = sm.x * sm.x + sm.y * sm.y;
      if (lim < ap && lim < am) {
        zd T0, T1, T2, T3;
        joint_series(sd0, sd0, sd2, sd3, sdm, sp, sm, ap, am, T0, T1, T2, T3);
        lds_add(y30 + f, T0);
        lds_add(y31 + f, T1);
        lds_add(z2 + f, T2);
        lds_add(z3 + f, T3);
        continue;
      }
#pragma unroll 1
      for (int sg = 0; sg < 2; ++sg) {
        const double ds = sg ? -1.0 : 1.0;  // the item's dsgn: its nodes are ds times those above
        if (kind != kAll) {
          zd w0, w1, w2, w3;
          if (series_w(m1 * ds, m1 * ds, n2 * ds, n3 * ds, e, w0, w1, w2, w3)) {
            lds_add(y30 + f, w0 * ds);
            lds_add(y31 + f, w1 * ds);
            lds_add(z2 + f, w2 * ds);
            lds_add(z3 + f, w3 * ds);
            continue;
          }
        }
        const zd H30 = G30 * ds, H31 = G31 * ds, H32 = G32 * ds, H33 = G33 * ds;
        const zd H40 = G40 * ds, H41 = G41 * ds;
        if (kind == kPair) {
          const zd q1 = (H30 + H31) / 2.0;
          const zd c1 = f_sub2g10(q1, H40, H41, e);
          const zd c2 = f_sub2g30(q1, H40, H41, e);
          const zd c3 = f_sub1g30(q1, H41, H40, e);
          lds_add(y30 + f, c1 * ds);
          lds_add(y31 + f, c1 * ds);
          lds_add(y40 + f, c2 * ds);
          lds_add(y41 + f, (c3 - c2 - c1 - c1) * ds);
        } else if (kind == kTriple) {
          const zd q1 = (H30 + H31 + H32) / 3.0;
          const zd c1 = f_sub2g30g10(q1, H40, e);
          const zd c2 = f_sub1g30g20(q1, H40, e);
          lds_add(y30 + f, c1 * ds);
          lds_add(y31 + f, c1 * ds);
          lds_add(y32 + f, c1 * ds);
          lds_add(y40 + f, (c2 - c1 - c1 - c1) * ds);
        } else if (kind == kTwoPairs) {
          const zd q1 = (H30 + H31) / 2.0, q2 = (H32 + H33) / 2.0;
          const zd c1 = f_sub2g30g20(q1, q2, e);
          const zd c2 = f_sub1g30g21(q1, q2, e);
          lds_add(y30 + f, c1 * ds);
          lds_add(y31 + f, c1 * ds);
          lds_add(y32 + f, (c2 - c1 - c1) * ds * 0.5);
          lds_add(y33 + f, (c2 - c1 - c1) * ds * 0.5);
        } else {  // kAll
          const zd q1 = (H30 + H31 + H32 + H33) / 4.0;
          const zd c1 = f_sub2simple(q1, e);
          lds_add(y30 + f, c1 * ds);
          lds_add(y31 + f, c1 * ds);
          lds_add(y32 + f, c1 * ds);
          lds_add(y33 + f, c1 * ds);
        }
      }
    }
  }
  __syncthreads();
  const double inv_ntk = 1.0 / (double)ntk;
  for (int i = threadIdx.x; i < nk3 * nfb; i += blockDim.x) {
    const int gi = i / nfb, f = i % nfb;
    const int k = index_kpt[gi];
    const double2 po = pband[(size_t)jo * ntk + k], pv = pband[(size_t)iv * ntk + ikq[k]];
    // 1 - P_occ(j,k) P_vir(i,k+q)
    const zd fac = mkz(1.0, 0.0) - mkz(po.x, po.y) * mkz(pv.x, pv.y);
    const double2 xv = x[(size_t)gi * nf + f];
    const zd v = mkz(xv.x, xv.y) * inv_ntk * fac;
    dout[(size_t)(f0 + f) * ldd + (size_t)p * ntk + k] = make_double2(v.x, v.y);
  }
}

// ------------------------------------------------- tetrahedra cut by E_F
// Vertex terms of calc (m_tetrahedron_20170325.F:209-300) for one
// tetrahedron with vertex values g0..g3 of class `code` at frequency e, the
// factor fk left out: xo(v) += fk * t_v.  The closed forms are those of the
// reference, each evaluated directly (the uncut kernel shares logarithms).
__device__ __forceinline__ void add_at(int i, zd v, zd& t0, zd& t1, zd& t2, zd& t3) {
  t0.x += i == 0 ? v.x : 0.0;
  t0.y += i == 0 ? v.y : 0.0;
  t1.x += i == 1 ? v.x : 0.0;
  t1.y += i == 1 ? v.y : 0.0;
  t2.x += i == 2 ? v.x : 0.0;
  t2.y += i == 2 ? v.y : 0.0;
  t3.x += i == 3 ? v.x : 0.0;
  t3.y += i == 3 ? v.y : 0.0;
}

__device__ __forceinline__ void calc_terms(int code, zd g0, zd g1, zd g2, zd g3, zd e, zd& t0,
                                           zd& t1, zd& t2, zd& t3) {
  t0 = t1 = t2 = t3 = mkz(0.0, 0.0);
  const int kind = code & 15;
  {
    // the class's nodes by vertex (near-equal values replaced by their mean,
    // as the closed forms below take them), then the series where it converges
    const int b0 = vtx(code, 0, 0), b1 = vtx(code, 0, 1), b2 = vtx(code, 0, 2), b3 = vtx(code, 0, 3);
    const zd G30 = sel4(b0, g0, g1, g2, g3), G31 = sel4(b1, g0, g1, g2, g3);
    const zd G32 = sel4(b2, g0, g1, g2, g3), G33 = sel4(b3, g0, g1, g2, g3);
    zd q1 = mkz(0.0, 0.0), q2 = q1;
    unsigned in1 = 0, in2 = 0;  // vertex masks of the classes with means q1, q2
    if (kind == kPair) {
      q1 = (G30 + G31) / 2.0;
      in1 = (1u << b0) | (1u << b1);
    } else if (kind == kTriple) {
      q1 = (G30 + G31 + G32) / 3.0;
      in1 = (1u << b0) | (1u << b1) | (1u << b2);
    } else if (kind == kTwoPairs) {
      q1 = (G30 + G31) / 2.0;
      q2 = (G32 + G33) / 2.0;
      in1 = (1u << b0) | (1u << b1);
      in2 = (1u << b2) | (1u << b3);
    } else if (kind == kAll) {
      q1 = (G30 + G31 + G32 + G33) / 4.0;
      in1 = 15u;
    }
    const zd n0 = (in1 & 1u) ? q1 : (in2 & 1u) ? q2 : g0;
    const zd n1 = (in1 & 2u) ? q1 : (in2 & 2u) ? q2 : g1;
    const zd n2 = (in1 & 4u) ? q1 : (in2 & 4u) ? q2 : g2;
    const zd n3 = (in1 & 8u) ? q1 : (in2 & 8u) ? q2 : g3;
    if (series_w(n0, n1, n2, n3, e, t0, t1, t2, t3)) return;
  }
  if (kind == kNormal) {  // case(0)
    const int i0 = vtx(code, 1, 0), i1 = vtx(code, 1, 1), i2 = vtx(code, 1, 2), i3 = vtx(code, 1, 3);
    const zd A0 = sel4(i0, g0, g1, g2, g3), A1 = sel4(i1, g0, g1, g2, g3);
    const zd A2 = sel4(i2, g0, g1, g2, g3), A3 = sel4(i3, g0, g1, g2, g3);
    const zd c1 = f_sub2normal(A1, A0, A2, A3, e);
    const zd c2 = f_sub2normal(A0, A1, A2, A3, e);
    const zd c3 = f_sub2normal(A0, A2, A1, A3, e);
    const zd c4 = f_sub1normal(A0, A3, A1, A2, e);
    add_at(i0, c1, t0, t1, t2, t3);
    add_at(i1, c2, t0, t1, t2, t3);
    add_at(i2, c3, t0, t1, t2, t3);
    add_at(i3, c4 - c3 - c2 - c1, t0, t1, t2, t3);
    return;
  }
  const int b0 = vtx(code, 0, 0), b1 = vtx(code, 0, 1), b2 = vtx(code, 0, 2), b3 = vtx(code, 0, 3);
  const int o0 = vtx(code, 1, 0), o1 = vtx(code, 1, 1);
  const zd G30 = sel4(b0, g0, g1, g2, g3), G31 = sel4(b1, g0, g1, g2, g3);
  const zd G32 = sel4(b2, g0, g1, g2, g3), G33 = sel4(b3, g0, g1, g2, g3);
  const zd G40 = sel4(o0, g0, g1, g2, g3), G41 = sel4(o1, g0, g1, g2, g3);
  if (kind == kPair) {  // case(2)
    const zd q1 = (G30 + G31) / 2.0;
    const zd c1 = f_sub2g10(q1, G40, G41, e);
    const zd c2 = f_sub2g30(q1, G40, G41, e);
    const zd c3 = f_sub1g30(q1, G41, G40, e);
    add_at(b0, c1, t0, t1, t2, t3);
    add_at(b1, c1, t0, t1, t2, t3);
    add_at(o0, c2, t0, t1, t2, t3);
    add_at(o1, c3 - c2 - c1 - c1, t0, t1, t2, t3);
  } else if (kind == kTriple) {  // case(3)
    const zd q1 = (G30 + G31 + G32) / 3.0;
    const zd c1 = f_sub2g30g10(q1, G40, e);
    const zd c2 = f_sub1g30g20(q1, G40, e);
    add_at(b0, c1, t0, t1, t2, t3);
    add_at(b1, c1, t0, t1, t2, t3);
    add_at(b2, c1, t0, t1, t2, t3);
    add_at(o0, c2 - c1 - c1 - c1, t0, t1, t2, t3);
  } else if (kind == kTwoPairs) {  // case(4), i3 = 2
    const zd q1 = (G30 + G31) / 2.0, q2 = (G32 + G33) / 2.0;
    const zd c1 = f_sub2g30g20(q1, q2, e);
    const zd c2 = f_sub1g30g21(q1, q2, e);
    add_at(b0, c1, t0, t1, t2, t3);
    add_at(b1, c1, t0, t1, t2, t3);
    add_at(b2, (c2 - c1 - c1) * 0.5, t0, t1, t2, t3);
    add_at(b3, (c2 - c1 - c1) * 0.5, t0, t1, t2, t3);
  } else {  // case(4), i3 = 3..6
    const zd q1 = (G30 + G31 + G32 + G33) / 4.0;
    const zd c1 = f_sub2simple(q1, e);
    add_at(b0, c1, t0, t1, t2, t3);
    add_at(b1, c1, t0, t1, t2, t3);
    add_at(b2, c1, t0, t1, t2, t3);
    add_at(b3, c1, t0, t1, t2, t3);
  }
}

// A node of ttct's split tree (m_tetrahedron_20170325.F:143-207): row v holds
// the weights of the four original corners in vertex v (the reference's
// dwgt(:, v, t)), vol the volume fraction, marks the vertices already moved
// onto the cutting plane (bvtx).  Only constant indices after unrolling, and
// selects for data-dependent ones, so the node stays in registers.
struct CutNode {
  double r[4][4];
  double vol;
  unsigned marks;
};

__device__ __forceinline__ void node_values(const CutNode& n, const double (&dv)[4],
                                            double (&val)[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v)
    val[v] = n.r[v][0] * dv[0] + n.r[v][1] * dv[1] + n.r[v][2] * dv[2] + n.r[v][3] * dv[3];
}

// First pair i1 < i2 of unmoved vertices strictly on opposite sides of the
// plane, in the reference's loop order.
__device__ __forceinline__ bool find_cut(const CutNode& n, const double (&val)[4], double dval,
                                         int& i1, int& i2) {
  bool found = false;
  i1 = i2 = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a + 1; b < 4; ++b) {
      const bool ok = !found && !((n.marks >> a) & 1u) && !((n.marks >> b) & 1u) &&
                      (val[a] - dval) * (dval - val[b]) > 0.0;
      if (ok) {
        found = true;
        i1 = a;
        i2 = b;
      }
    }
  return found;
}

// Split along edge (i1, i2) where it meets the plane.  child 1 is the
// reference's appended tetrahedron (vertex i1 moved, volume * d1), child 0
// the one split in place (vertex i2 moved, volume * d2).
__device__ __forceinline__ void split_node(CutNode& n, const double (&val)[4], double dval, int i1,
                                           int i2, unsigned child) {
  const double v1 = sel4d(i1, val[0], val[1], val[2], val[3]);
  const double v2 = sel4d(i2, val[0], val[1], val[2], val[3]);
  const double d1 = (v2 - dval) / (v2 - v1), d2 = (dval - v1) / (v2 - v1);
  const int tgt = child ? i1 : i2;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double w = sel4d(i1, n.r[0][j], n.r[1][j], n.r[2][j], n.r[3][j]) * d1 +
                     sel4d(i2, n.r[0][j], n.r[1][j], n.r[2][j], n.r[3][j]) * d2;
#pragma unroll
    for (int v = 0; v < 4; ++v) n.r[v][j] = v == tgt ? w : n.r[v][j];
  }
  n.vol *= child ? d1 : d2;
  n.marks |= 1u << tgt;
}

// ttct's keep test for a node that is not split: some unmoved vertex on the
// imod side of the plane.
__device__ __forceinline__ bool node_kept(const CutNode& n, const double (&val)[4], double dval,
                                          double imod) {
  bool k = false;
#pragma unroll
  for (int v = 0; v < 4; ++v) k = k || (!((n.marks >> v) & 1u) && imod * (val[v] - dval) > 0.0);
  return k;
}

// The tetrahedra of a pair that straddle E_F (in the empty band e2 or the
// occupied band e1), as ttrhdrn_chi0 treats them (m_tetrahedron_20170325.F:
// 114-139): ttct cuts by e2 = E_F keeping the e2 > E_F side, then every piece
// by e1 = E_F keeping e1 < E_F, and calc runs on each piece with f = vol*dsgn
// and g_v = -(sum_j (e2_j - e1_j) dwgt(j,v) - i delt) dsgn, its vertex terms
// spread back to the corners with dwgt.
//
// A plane moves one more vertex onto itself with every cut and needs two
// unmoved vertices to cut again, so a tree has at most 3 cut levels per plane:
// at most 2^6 pieces.  Each thread owns one tetrahedron and walks the leaves
// of its tree depth first, rebuilding each from the root along its path of
// child bits, so the pieces never need storage; both signs dsgn share the
// pieces and the degeneracy class (chk1 depends on |g_p - g_q| and |g_p + g_q|).
//
// grid as k_tetra_weights, which must have run first: the pieces are added
// to the weights it wrote, with the same 1/NTK and cRPA factor.  Status bit 2
// as there; bit 8 if a tree had more leaves than the bound (cannot happen).
__global__ void __launch_bounds__(256) k_tetra_cut(
    int ntk, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const double* __restrict__ eband,
    const double2* __restrict__ pband, double ef, double delt, double dmna, double dmnr, int nen,
    int nf, const double2* __restrict__ em, size_t ldd, int* __restrict__ status,
    double2* __restrict__ dout) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  const int ntet = 6 * nk3;
  double2* x = lds;                              // [nk3][nf]
  double* e1 = (double*)(x + (size_t)nk3 * nf);  // [nk3]
  double* e2 = e1 + nk3;                         // [nk3]
  int* list = (int*)(e2 + nk3);                  // [ntet]
  __shared__ int count;
  const int p = blockIdx.x;
  const int f0 = blockIdx.y * nf, nfb = min(nf, nen - f0);
  const int jo = pocc[p], iv = pvir[p];
  for (int i = threadIdx.x; i < nk3 * nf; i += blockDim.x) x[i] = make_double2(0, 0);
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    const int k = index_kpt[gi];
    e1[gi] = (double)(float)eband[(size_t)jo * ntk + k];
    e2[gi] = (double)(float)eband[(size_t)iv * ntk + ikq[k]];
  }
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < ntet; t += blockDim.x) {
    const int c0 = imt[4 * t], c1 = imt[4 * t + 1], c2 = imt[4 * t + 2], c3 = imt[4 * t + 3];
    if (straddles(ef, e2[c0], e2[c1], e2[c2], e2[c3]) ||
        straddles(ef, e1[c0], e1[c1], e1[c2], e1[c3]))
      list[atomicAdd(&count, 1)] = t;
  }
  __syncthreads();
  const int ncut = count;
  if (ncut == 0) return;  // uniform over the block
  for (int j = threadIdx.x; j < ncut; j += blockDim.x) {
    const int t = list[j];
    const int c0 = imt[4 * t], c1 = imt[4 * t + 1], c2 = imt[4 * t + 2], c3 = imt[4 * t + 3];
    const double de2[4] = {e2[c0], e2[c1], e2[c2], e2[c3]};
    const double de1[4] = {e1[c0], e1[c1], e1[c2], e1[c3]};
    double2* y0 = x + (size_t)c0 * nf;
    double2* y1 = x + (size_t)c1 * nf;
    double2* y2 = x + (size_t)c2 * nf;
    double2* y3 = x + (size_t)c3 * nf;
    unsigned path = 0;
    bool more = true;
#pragma unroll 1
    for (int leaf = 0; leaf < 64 && more; ++leaf) {
      CutNode n;
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int c = 0; c < 4; ++c) n.r[v][c] = v == c ? 1.0 : 0.0;
      n.vol = 1.0;
      n.marks = 0;
      int depth = 0;
      bool keep = true;
#pragma unroll
      for (int plane = 0; plane < 2; ++plane) {
        if (keep) {
          const double(&dv)[4] = plane == 0 ? de2 : de1;
          const double imod = plane == 0 ? 1.0 : -1.0;
          n.marks = 0;
          bool leafy = false;
          double val[4];
#pragma unroll 1
          for (int step = 0; step < 4 && !leafy; ++step) {
            node_values(n, dv, val);
            int i1, i2;
            if (find_cut(n, val, ef, i1, i2)) {
              split_node(n, val, ef, i1, i2, (path >> depth) & 1u);
              ++depth;
            } else {
              leafy = true;
              keep = node_kept(n, val, ef, imod);
            }
          }
          if (!leafy) {  // more than three cuts in one plane: impossible
            atomicOr(status, 8);
            keep = false;
          }
        }
      }
      if (keep) {
        double s[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
          s[v] = (de2[0] - de1[0]) * n.r[v][0] + (de2[1] - de1[1]) * n.r[v][1] +
                 (de2[2] - de1[2]) * n.r[v][2] + (de2[3] - de1[3]) * n.r[v][3];
        const zd g[4] = {mkz(-s[0], delt), mkz(-s[1], delt), mkz(-s[2], delt), mkz(-s[3], delt)};
        const int code = classify(dmna, dmnr, g);
        if ((code & 15) == kBad) {
          atomicOr(status, 2);
        } else {
#pragma unroll 1
          for (int f = 0; f < nfb; ++f) {
            const zd e = mkz(em[f0 + f].x, em[f0 + f].y);
            zd T0 = mkz(0, 0), T1 = mkz(0, 0), T2 = mkz(0, 0), T3 = mkz(0, 0);
#pragma unroll 1
            for (int sg = 0; sg < 2; ++sg) {  // dsgn = +1, then -1: g -> -g, f -> -f
              const double ds = sg == 0 ? 1.0 : -1.0;
              zd t0, t1, t2, t3;
              calc_terms(code, g[0] * ds, g[1] * ds, g[2] * ds, g[3] * ds, e, t0, t1, t2, t3);
              T0 = T0 + t0 * ds;
              T1 = T1 + t1 * ds;
              T2 = T2 + t2 * ds;
              T3 = T3 + t3 * ds;
            }
            T0 = T0 * n.vol;
            T1 = T1 * n.vol;
            T2 = T2 * n.vol;
            T3 = T3 * n.vol;
            // corner c: sum_v T_v dwgt(c, v)
            lds_add(y0 + f, T0 * n.r[0][0] + T1 * n.r[1][0] + T2 * n.r[2][0] + T3 * n.r[3][0]);
            lds_add(y1 + f, T0 * n.r[0][1] + T1 * n.r[1][1] + T2 * n.r[2][1] + T3 * n.r[3][1]);
            lds_add(y2 + f, T0 * n.r[0][2] + T1 * n.r[1][2] + T2 * n.r[2][2] + T3 * n.r[3][2]);
            lds_add(y3 + f, T0 * n.r[0][3] + T1 * n.r[1][3] + T2 * n.r[2][3] + T3 * n.r[3][3]);
          }
        }
      }
      // next leaf in depth-first order: flip the deepest 0 bit of the path
      int b = depth - 1;
      while (b >= 0 && ((path >> b) & 1u)) --b;
      if (b < 0) more = false;
      else path = (path & ((1u << b) - 1u)) | (1u << b);
    }
    if (more) atomicOr(status, 8);
  }
  __syncthreads();
  const double inv_ntk = 1.0 / (double)ntk;
  for (int i = threadIdx.x; i < nk3 * nfb; i += blockDim.x) {
    const int gi = i / nfb, f = i % nfb;
    const int k = index_kpt[gi];
    const double2 po = pband[(size_t)jo * ntk + k], pv = pband[(size_t)iv * ntk + ikq[k]];
    const zd fac = mkz(1.0, 0.0) - mkz(po.x, po.y) * mkz(pv.x, pv.y);
    const double2 xv = x[(size_t)gi * nf + f];
    const zd v = mkz(xv.x, xv.y) * inv_ntk * fac;
    double2* o = dout + (size_t)(f0 + f) * ldd + (size_t)p * ntk + k;
    *o = make_double2(o->x + v.x, o->y + v.y);
  }
}

// Drude sums at q = 0 (chiqw.F:1631-1733), one intraband band per block:
// p(k) = -i calc_VMaa(n, n, k) = sum_G (k+G) conj(C) C (Cartesian), the
// ttrhdrn_simple weights x(k) for f = 1 and g = CMPLX(E_n(k) - E_F, -delt)
// (single precision) at em = 0, then part[block][d] = sum_k Im x(k) p_d(k)^2.
// coef/kg/ng0 as k_q0_heads; sk0 the full-mesh k (fractional).
__global__ void __launch_bounds__(256) k_drude(
    const double2* __restrict__ coef, const int* __restrict__ kg, const int* __restrict__ ng0,
    int ntk, int ntg, const int* __restrict__ bands, const double* __restrict__ sk0,
    const double* __restrict__ bvec, const double* __restrict__ eband, double ef, double delt,
    double dmna, double dmnr, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, int* __restrict__ status, double2* __restrict__ part) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  double2* pk = lds;           // [ntk][3]
  double2* xo = pk + 3 * ntk;  // [nk3]
  __shared__ double red[6][4];
  const int n = bands[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  for (int k = wave; k < ntk; k += nwave) {
    const double2* c = coef + ((size_t)n * ntk + k) * ntg;
    const int* kk = kg + (size_t)k * ntg * 3;
    const double k0 = sk0[3 * k], k1 = sk0[3 * k + 1], k2 = sk0[3 * k + 2];
    double s[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int i = lane; i < ng0[k]; i += 64) {
      const double2 a = c[i];
      const double pr = a.x * a.x + a.y * a.y, pi = a.x * a.y - a.y * a.x;  // conj(C) C
      const double u0 = k0 + (double)kk[3 * i], u1 = k1 + (double)kk[3 * i + 1],
                   u2 = k2 + (double)kk[3 * i + 2];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double gv = u0 * bvec[d] + u1 * bvec[3 + d] + u2 * bvec[6 + d];
        s[d][0] += gv * pr;
        s[d][1] += gv * pi;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        s[d][0] += __shfl_xor(s[d][0], off);
        s[d][1] += __shfl_xor(s[d][1], off);
      }
    // vm = i S and p = -i vm = S
    if (lane == 0) {
      pk[3 * k] = make_double2(s[0][0], s[0][1]);
      pk[3 * k + 1] = make_double2(s[1][0], s[1][1]);
      pk[3 * k + 2] = make_double2(s[2][0], s[2][1]);
    }
  }
  for (int i = threadIdx.x; i < nk3; i += blockDim.x) xo[i] = make_double2(0, 0);
  __syncthreads();
  const zd e0 = mkz(0.0, 0.0);
  for (int t = threadIdx.x; t < 6 * nk3; t += blockDim.x) {
    const int c0 = imt[4 * t], c1 = imt[4 * t + 1], c2 = imt[4 * t + 2], c3 = imt[4 * t + 3];
    const double mdl = (double)(float)(-delt);
    const zd g[4] = {mkz((double)(float)(eband[(size_t)n * ntk + index_kpt[c0]] - ef), mdl),
                     mkz((double)(float)(eband[(size_t)n * ntk + index_kpt[c1]] - ef), mdl),
                     mkz((double)(float)(eband[(size_t)n * ntk + index_kpt[c2]] - ef), mdl),
                     mkz((double)(float)(eband[(size_t)n * ntk + index_kpt[c3]] - ef), mdl)};
    const int code = classify(dmna, dmnr, g);
    if ((code & 15) == kBad) {
      atomicOr(status, 2);
      continue;
    }
    zd t0, t1, t2, t3;
    calc_terms(code, g[0], g[1], g[2], g[3], e0, t0, t1, t2, t3);
    lds_add(xo + c0, t0);
    lds_add(xo + c1, t1);
    lds_add(xo + c2, t2);
    lds_add(xo + c3, t3);
  }
  __syncthreads();
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    const int k = index_kpt[gi];
    const double im = xo[gi].y;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double2 q = pk[3 * k + d];
      acc[2 * d] += im * (q.x * q.x - q.y * q.y);
      acc[2 * d + 1] += im * (2.0 * q.x * q.y);
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
    if (lane == 0) red[c][wave] = acc[c];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double re = 0, im = 0;
    for (int w = 0; w < nwave; ++w) {
      re += red[2 * d][w];
      im += red[2 * d + 1][w];
    }
    part[3 * blockIdx.x + d] = make_double2(re, im);
  }
}

// wpl = 8 sum_bands part / (VOLUME NTK), wd = sqrt(Re wpl) (chiqw.F:1991-1993).
__global__ void k_drude_finish(const double2* __restrict__ part, int nb, double volume, int ntk,
                               double* __restrict__ wd) {
  const int d = threadIdx.x;
  if (d >= 3) return;
  double re = 0.0;
  for (int b = 0; b < nb; ++b) re += part[3 * b + d].x;
  wd[d] = sqrt(8.0 * re / volume / (double)ntk);
}

// Closed-form queue of the sweep: a quarter of the (tetrahedron, frequency)
// pairs of a workgroup (the rest, rarely needed, are evaluated inline).
int tetra_queue_len(int nk3, int nf) { return (6 * nk3 * nf) / 4 + 64; }

size_t tetra_lds_bytes(int nk3, int nf) {
  return 3 * sizeof(double2) * (size_t)nk3 * nf + 2 * sizeof(double) * (size_t)nk3 +
         sizeof(int2) * 6 * (size_t)nk3 + sizeof(int) * (size_t)tetra_queue_len(nk3, nf);
}

// Frequencies per workgroup: as many as keep two workgroups per CU (the
// register budget allows two waves per SIMD), evened out over the slices.
int tetra_freqs_per_block(int nk3, int nen) {
  const size_t fixed = tetra_lds_bytes(nk3, 0), per = tetra_lds_bytes(nk3, 1) - fixed;
  const size_t target = 80 * 1024;
  int nf = target > fixed ? (int)((target - fixed) / per) : 1;
  nf = std::max(1, std::min(nf, nen));
  const int nslice = (nen + nf - 1) / nf;
  return (nen + nslice - 1) / nslice;
}

void launch_tetra_weights(int ntk, int n1, int n2, int n3, const int* imt, const int* index_kpt,
                          const int* pocc, const int* pvir, const int* ikq, const double* eband,
                          const double2* pband, double ef, double delt, double dmna, double dmnr,
                          int nen, const double2* em, size_t ldd, int npair, int cuts, int* status,
                          double2* dout, hipStream_t s) {
  const int nk3 = n1 * n2 * n3;
  const int nf = tetra_freqs_per_block(nk3, nen);
  const size_t lds = tetra_lds_bytes(nk3, nf);
  RP_CHECK(lds <= 160 * 1024, "k mesh too large for the LDS accumulator of the tetrahedron kernel");
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute((const void*)k_tetra_weights,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const dim3 grid(npair, (nen + nf - 1) / nf);
  hipLaunchKernelGGL(k_tetra_weights, grid, dim3(256), lds, s, ntk, n1, n2, n3, imt, index_kpt,
                     pocc, pvir, ikq, eband, pband, ef, delt, dmna, dmnr, nen, nf, em, ldd,
                     cuts ? 1 : 0, tetra_queue_len(nk3, nf), status, dout);
  RP_HIP(hipGetLastError());
  if (!cuts) return;
  // x[nk3][nf], e1, e2, the cut list (<= lds of the uncut kernel)
  const size_t lcut = sizeof(double2) * (size_t)nk3 * nf + 2 * sizeof(double) * (size_t)nk3 +
                      sizeof(int) * 6 * (size_t)nk3;
  if (lcut > 65536)
    RP_HIP(hipFuncSetAttribute((const void*)k_tetra_cut,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lcut));
  hipLaunchKernelGGL(k_tetra_cut, grid, dim3(256), lcut, s, ntk, n1, n2, n3, imt, index_kpt, pocc,
                     pvir, ikq, eband, pband, ef, delt, dmna, dmnr, nen, nf, em, ldd, status, dout);
  RP_HIP(hipGetLastError());
}

void launch_drude(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                  const int* bands, int nbands, const double* sk0, const double* bvec,
                  const double* eband, double ef, double delt, double dmna, double dmnr, int n1,
                  int n2, int n3, const int* imt, const int* index_kpt, double volume,
                  int* status, double2* part, double* wd, hipStream_t s) {
  const int nk3 = n1 * n2 * n3;
  const size_t lds = sizeof(double2) * (3 * (size_t)ntk + nk3);
  RP_CHECK(lds <= 64 * 1024, "k mesh too large for the Drude kernel");
  hipLaunchKernelGGL(k_drude, dim3(nbands), dim3(256), lds, s, coef, kg, ng0, ntk, ntg, bands, sk0,
                     bvec, eband, ef, delt, dmna, dmnr, n1, n2, n3, imt, index_kpt, status, part);
  RP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drude_finish, dim3(1), dim3(64), 0, s, part, nbands, volume, ntk, wd);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
