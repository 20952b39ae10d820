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
// FP64 atomics.  Only tetrahedra not cut by E_F are handled (every system
// without partially occupied bands, where the reference's ttct never splits);
// the host rejects systems with partially occupied bands before launch.
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
// chiqw.F:1353,1356).  Tetrahedra cut by E_F raise status bit 1, a class the
// reference rejects (errexec) bit 2.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_tetra_weights(
    int ntk, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const double* __restrict__ eband,
    const double2* __restrict__ pband, double ef, double delt, double dmna, double dmnr, int nen,
    int nf, const double2* __restrict__ em, size_t ldd, int* __restrict__ status,
    double2* __restrict__ dout) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  const int ntet = 6 * nk3, nitem = 2 * ntet;
  double2* x = lds;                               // [nk3][nf]
  double* e1 = (double*)(x + (size_t)nk3 * nf);   // [nk3]
  double* e2 = e1 + nk3;                          // [nk3]
  int2* list = (int2*)(e2 + nk3);                 // [nitem] {item, class}
  __shared__ int count[2];
  const int p = blockIdx.x;
  const int f0 = blockIdx.y * nf, nfb = min(nf, nen - f0);
  const int jo = pocc[p], iv = pvir[p];
  for (int i = threadIdx.x; i < nk3 * nf; i += blockDim.x) x[i] = make_double2(0, 0);
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    const int k = index_kpt[gi];
    e1[gi] = (double)(float)eband[(size_t)jo * ntk + k];
    e2[gi] = (double)(float)eband[(size_t)iv * ntk + ikq[k]];
  }
  if (threadIdx.x < 2) count[threadIdx.x] = 0;
  __syncthreads();
  // ---- pass 1: cut test (ttct), keep test, class
  for (int it = threadIdx.x; it < nitem; it += blockDim.x) {
    const int t = it % ntet;
    const int c0 = imt[4 * t], c1 = imt[4 * t + 1], c2 = imt[4 * t + 2], c3 = imt[4 * t + 3];
    const double a0 = e1[c0], a1 = e1[c1], a2 = e1[c2], a3 = e1[c3];
    const double b0 = e2[c0], b1 = e2[c1], b2 = e2[c2], b3 = e2[c3];
    if (straddles(ef, b0, b1, b2, b3) || straddles(ef, a0, a1, a2, a3)) {
      atomicOr(status, 1);
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
    const int slot = kind == kNormal ? atomicAdd(&count[0], 1) : nitem - 1 - atomicAdd(&count[1], 1);
    list[slot] = make_int2(it, code);
  }
  __syncthreads();
  const int nnormal = count[0], nother = count[1];
  // ---- pass 2: non-degenerate tetrahedra (calc, case(0); ttrhdrn_sub2normal x3 and
  // ttrhdrn_sub1normal, m_tetrahedron_20170325.F:215-229,431-451)
  for (int j = threadIdx.x; j < nnormal; j += blockDim.x) {
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
    const double ds = m.dsgn;
#pragma unroll 1
    for (int f = 0; f < nfb; ++f) {
      const double2 ev = em[f0 + f];
      const zd e = mkz(ev.x, ev.y);
      const zd v0 = A0 + e, v1 = A1 + e, v2 = A2 + e, v3 = A3 + e;
      const zd L0 = logz(v0), L1 = logz(v1), L2 = logz(v2), L3 = logz(v3);
      const zd D10 = logdiff(L1, L0), D20 = logdiff(L2, L0), D30 = logdiff(L3, L0);
      const zd D21 = logdiff(L2, L1), D31 = logdiff(L3, L1), D32 = logdiff(L3, L2);
      const zd q0 = v0 * v0, q1 = v1 * v1, q2 = v2 * v2, q3 = v3 * v3;
      const zd u0 = q0 * v0, u1 = q1 * v1, u2 = q2 * v2, u3 = q3 * v3;
      zd c1 = -q0 * K11;
      c1 = c1 - u2 * D20 * K12;
      c1 = c1 + u3 * D30 * K13;
      c1 = c1 + u1 * D10 * K14;
      zd c2 = -q1 * K21;
      c2 = c2 - u2 * D21 * K22;
      c2 = c2 + u3 * D31 * K23;
      c2 = c2 - u0 * D10 * K24;
      zd c3 = -q2 * K31;
      c3 = c3 + u1 * D21 * K32;
      c3 = c3 + u3 * D32 * K33;
      c3 = c3 - u0 * D20 * K34;
      const zd l1 = q3 * D30;
      zd c4 = -l1 * J1;
      c4 = c4 - (l1 - q1 * D10) * J2;
      c4 = c4 + (l1 - q2 * D20) * J3;
      lds_add(x0 + f, c1 * ds);
      lds_add(x1 + f, c2 * ds);
      lds_add(x2 + f, c3 * ds);
      lds_add(x3 + f, (c4 - c3 - c2 - c1) * ds);
    }
  }
  // ---- pass 3: degenerate tetrahedra (calc, m_tetrahedron_20170325.F:231-300)
  for (int j = threadIdx.x; j < nother; j += blockDim.x) {
    const int2 ent = list[nitem - 1 - j];
    const Item m = load_item(ent.x, ntet, imt, e1, e2, delt);
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
    const double ds = m.dsgn;
    if (kind == kPair) {
      const zd q1 = (G30 + G31) / 2.0;
#pragma unroll 1
      for (int f = 0; f < nfb; ++f) {
        const zd e = mkz(em[f0 + f].x, em[f0 + f].y);
        const zd c1 = f_sub2g10(q1, G40, G41, e);
        const zd c2 = f_sub2g30(q1, G40, G41, e);
        const zd c3 = f_sub1g30(q1, G41, G40, e);
        lds_add(y30 + f, c1 * ds);
        lds_add(y31 + f, c1 * ds);
        lds_add(y40 + f, c2 * ds);
        lds_add(y41 + f, (c3 - c2 - c1 - c1) * ds);
      }
    } else if (kind == kTriple) {
      const zd q1 = (G30 + G31 + G32) / 3.0;
#pragma unroll 1
      for (int f = 0; f < nfb; ++f) {
        const zd e = mkz(em[f0 + f].x, em[f0 + f].y);
        const zd c1 = f_sub2g30g10(q1, G40, e);
        const zd c2 = f_sub1g30g20(q1, G40, e);
        lds_add(y30 + f, c1 * ds);
        lds_add(y31 + f, c1 * ds);
        lds_add(y32 + f, c1 * ds);
        lds_add(y40 + f, (c2 - c1 - c1 - c1) * ds);
      }
    } else if (kind == kTwoPairs) {
      const zd q1 = (G30 + G31) / 2.0, q2 = (G32 + G33) / 2.0;
#pragma unroll 1
      for (int f = 0; f < nfb; ++f) {
        const zd e = mkz(em[f0 + f].x, em[f0 + f].y);
        const zd c1 = f_sub2g30g20(q1, q2, e);
        const zd c2 = f_sub1g30g21(q1, q2, e);
        lds_add(y30 + f, c1 * ds);
        lds_add(y31 + f, c1 * ds);
        lds_add(y32 + f, (c2 - c1 - c1) * ds * 0.5);
        lds_add(y33 + f, (c2 - c1 - c1) * ds * 0.5);
      }
    } else {  // kAll
      const zd q1 = (G30 + G31 + G32 + G33) / 4.0;
#pragma unroll 1
      for (int f = 0; f < nfb; ++f) {
        const zd e = mkz(em[f0 + f].x, em[f0 + f].y);
        const zd c1 = f_sub2simple(q1, e);
        lds_add(y30 + f, c1 * ds);
        lds_add(y31 + f, c1 * ds);
        lds_add(y32 + f, c1 * ds);
        lds_add(y33 + f, c1 * ds);
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

size_t tetra_lds_bytes(int nk3, int nf) {
  return sizeof(double2) * (size_t)nk3 * nf + 2 * sizeof(double) * (size_t)nk3 +
         sizeof(int2) * 12 * (size_t)nk3;
}

// Frequencies per workgroup: as many as keep two workgroups per CU (the
// register budget allows two waves per SIMD), evened out over the slices.
int tetra_freqs_per_block(int nk3, int nen) {
  const size_t fixed = tetra_lds_bytes(nk3, 0), per = sizeof(double2) * (size_t)nk3;
  const size_t target = 80 * 1024;
  int nf = target > fixed ? (int)((target - fixed) / per) : 1;
  nf = std::max(1, std::min(nf, nen));
  const int nslice = (nen + nf - 1) / nf;
  return (nen + nslice - 1) / nslice;
}

void launch_tetra_weights(int ntk, int n1, int n2, int n3, const int* imt, const int* index_kpt,
                          const int* pocc, const int* pvir, const int* ikq, const double* eband,
                          const double2* pband, double ef, double delt, double dmna, double dmnr,
                          int nen, const double2* em, size_t ldd, int npair, int* status,
                          double2* dout, hipStream_t s) {
  const int nk3 = n1 * n2 * n3;
  const int nf = tetra_freqs_per_block(nk3, nen);
  const size_t lds = tetra_lds_bytes(nk3, nf);
  RP_CHECK(lds <= 160 * 1024, "k mesh too large for the LDS accumulator of the tetrahedron kernel");
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute((const void*)k_tetra_weights,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_tetra_weights, dim3(npair, (nen + nf - 1) / nf), dim3(256), lds, s, ntk,
                     n1, n2, n3, imt, index_kpt, pocc, pvir, ikq, eband, pband, ef, delt, dmna,
                     dmnr, nen, nf, em, ldd, status, dout);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
