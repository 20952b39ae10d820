// Generalized-tetrahedron weights x_ij(k, w) of ttrhdrn_chi0
// (m_tetrahedron_20170325.F:99-503) for every band pair, in FP64 on gfx950.
//
// One workgroup per pair.  Its k-grid accumulator x[nk3][nen] (complex FP64)
// lives in LDS; threads take (tetrahedron, dsgn) items, evaluate the closed
// forms for all frequencies and add the four corner contributions with LDS
// atomics.  Only tetrahedra not cut by E_F are handled (every system without
// partially occupied bands, where the reference's ttct never splits); the
// host rejects systems with partially occupied bands before launch.
//
// The output is the contraction weight d[w][K] = x(k,w) (1 - P_occ P_vir)/NTK
// with K = pair*NTK + k (the factors of chiqw.F:1404-1406).
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

// Vertex classes of chk1 packed 4 bits per entry.
struct Class {
  int n2, i3;          // vertices in degenerate pairs; number of degenerate pairs
  unsigned a3, a4;     // degenerate vertices (discovery order) / the others
};
__device__ __forceinline__ int nib(unsigned x, int i) { return (int)((x >> (4 * i)) & 15u); }

// chk1 (m_tetrahedron_20170325.F:457-503): vertices compared in the
// reference's rotated order p = mod(i,4)+1.  g: this thread's 4 vertex values
// (LDS, so the data-dependent indices below need no private memory).
__device__ __forceinline__ Class classify(double dmna, double dmnr, const zd* g) {
  Class c;
  unsigned ia2 = 0, a3 = 0, a4 = 0;
  int i3 = 0, i4 = 0;
  for (int i1 = 1; i1 <= 3; ++i1)
    for (int i2 = i1 + 1; i2 <= 4; ++i2) {
      const int p = i1 & 3, q = i2 & 3;
      const double d1 = absz(g[p] - g[q]);
      const double d2 = absz(g[p] + g[q]) * 0.5;
      if (d1 < dmna || d1 / d2 < dmnr) {
        ++i3;
        if (!((ia2 >> p) & 1u)) { a3 |= (unsigned)p << (4 * i4); ++i4; ia2 |= 1u << p; }
        if (!((ia2 >> q) & 1u)) { a3 |= (unsigned)q << (4 * i4); ++i4; ia2 |= 1u << q; }
      }
    }
  int i5 = 0;
  for (int i1 = 1; i1 <= 4; ++i1) {
    const int p = i1 & 3;
    if (!((ia2 >> p) & 1u)) { a4 |= (unsigned)p << (4 * i5); ++i5; }
  }
  c.n2 = i4;
  c.i3 = i3;
  c.a3 = a3;
  c.a4 = a4;
  return c;
}

__device__ inline void lds_add(double2* p, zd v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
}

// calc (m_tetrahedron_20170325.F:209-302) for one frequency e with the four
// vertex weights all equal to f; vertex i's contribution is added to
// xw[cn[i] * nen] (LDS accumulator column of this frequency).  Returns false for
// a class the reference rejects (errexec).
__device__ __forceinline__ bool calc_one(const Class& c, const zd* g, const int* cn, zd f, zd e,
                                         double2* xw, int nen) {
  const int a30 = nib(c.a3, 0), a31 = nib(c.a3, 1), a32 = nib(c.a3, 2), a33 = nib(c.a3, 3);
  const int a40 = nib(c.a4, 0), a41 = nib(c.a4, 1), a42 = nib(c.a4, 2), a43 = nib(c.a4, 3);
#define ADD(i, v) lds_add(&xw[(size_t)cn[(i)] * nen], (v))
  switch (c.n2) {
    case 0: {
      const zd A0 = g[a40], A1 = g[a41], A2 = g[a42], A3 = g[a43];
      zd c1 = f_sub2normal(A1, A0, A2, A3, e);
      zd c2 = f_sub2normal(A0, A1, A2, A3, e);
      zd c3 = f_sub2normal(A0, A2, A1, A3, e);
      zd c4 = f_sub1normal(A0, A3, A1, A2, e);
      ADD(a40, f * c1);
      ADD(a41, f * c2);
      ADD(a42, f * c3);
      ADD(a43, f * (c4 - c3 - c2 - c1));
      return true;
    }
    case 2: {
      zd q1 = (g[a30] + g[a31]) / 2.0;
      const zd B0 = g[a40], B1 = g[a41];
      zd c1 = f_sub2g10(q1, B0, B1, e);
      zd c2 = f_sub2g30(q1, B0, B1, e);
      zd c3 = f_sub1g30(q1, B1, B0, e);
      ADD(a30, f * c1);
      ADD(a31, f * c1);
      ADD(a40, f * c2);
      ADD(a41, f * (c3 - c2 - c1 - c1));
      return true;
    }
    case 3: {
      zd q1 = (g[a30] + g[a31] + g[a32]) / 3.0;
      const zd B0 = g[a40];
      zd c1 = f_sub2g30g10(q1, B0, e);
      zd c2 = f_sub1g30g20(q1, B0, e);
      ADD(a30, f * c1);
      ADD(a31, f * c1);
      ADD(a32, f * c1);
      ADD(a40, f * (c2 - c1 - c1 - c1));
      return true;
    }
    case 4:
      if (c.i3 == 2) {
        zd q1 = (g[a30] + g[a31]) / 2.0, q2 = (g[a32] + g[a33]) / 2.0;
        zd c1 = f_sub2g30g20(q1, q2, e);
        zd c2 = f_sub1g30g21(q1, q2, e);
        ADD(a30, f * c1);
        ADD(a31, f * c1);
        ADD(a32, f * (c2 - c1 - c1) * 0.5);
        ADD(a33, f * (c2 - c1 - c1) * 0.5);
        return true;
      } else if (c.i3 >= 3 && c.i3 <= 6) {
        zd q1 = (g[a30] + g[a31] + g[a32] + g[a33]) / 4.0;
        zd c1 = f_sub2simple(q1, e);
        ADD(a30, c1 * f);
        ADD(a31, c1 * f);
        ADD(a32, c1 * f);
        ADD(a33, c1 * f);
        return true;
      }
      return false;
    default:
      return false;
  }
#undef ADD
}

// Does ttct need to split?  (Some pair of vertices strictly on both sides.)
__device__ inline bool straddles(double f, double v0, double v1, double v2, double v3) {
  // all six vertex pairs (the order does not matter for "any")
  const double d0 = v0 - f, d1 = v1 - f, d2 = v2 - f, d3 = v3 - f;
  return (d0 * -d1 > 0.0) || (d0 * -d2 > 0.0) || (d0 * -d3 > 0.0) || (d1 * -d2 > 0.0) ||
         (d1 * -d3 > 0.0) || (d2 * -d3 > 0.0);
}

// grid: npair blocks.  eband[band][k] energies (Ha, full mesh, FP64); the
// tetrahedron sees them rounded to single precision (CMPLX, chiqw.F:1353,1356).
// Tetrahedra cut by E_F (partially occupied bands) are not handled here: they
// raise status bit 1 (the host rejects such systems before launch).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_tetra_weights(
    int ntk, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const double* __restrict__ eband,
    const double2* __restrict__ pband, double ef, double delt, double dmna, double dmnr, int nen,
    const double2* __restrict__ em, size_t ldd, int* __restrict__ status,
    double2* __restrict__ dout) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  double2* x = lds;                               // [nk3][nen]
  zd* gslot = (zd*)(x + (size_t)nk3 * nen);       // [256][4] vertex values
  int* cslot = (int*)(gslot + 4 * 256);           // [256][4] corner grid indices
  double* e1 = (double*)(cslot + 4 * 256);        // [nk3]
  double* e2 = e1 + nk3;
  zd* gt = gslot + 4 * threadIdx.x;
  int* ct = cslot + 4 * threadIdx.x;
  const int p = blockIdx.x;
  const int jo = pocc[p], iv = pvir[p];
  for (int i = threadIdx.x; i < nk3 * nen; i += blockDim.x) x[i] = make_double2(0, 0);
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    const int k = index_kpt[gi];
    e1[gi] = (double)(float)eband[(size_t)jo * ntk + k];
    e2[gi] = (double)(float)eband[(size_t)iv * ntk + ikq[k]];
  }
  __syncthreads();
  const int ntet = 6 * nk3;
  for (int it = threadIdx.x; it < 2 * ntet; it += blockDim.x) {
    const int t = it % ntet;
    const double dsgn = it < ntet ? 1.0 : -1.0;
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
    // g_v = -(e2 - e1 - i delt) * dsgn, f_v = dvol * dsgn with dvol = 1
    gt[0] = mkz(-(b0 - a0), delt) * dsgn;
    gt[1] = mkz(-(b1 - a1), delt) * dsgn;
    gt[2] = mkz(-(b2 - a2), delt) * dsgn;
    gt[3] = mkz(-(b3 - a3), delt) * dsgn;
    ct[0] = c0;
    ct[1] = c1;
    ct[2] = c2;
    ct[3] = c3;
    const zd f = mkz(dsgn, 0.0);
    const Class cl = classify(dmna, dmnr, gt);
    for (int ie = 0; ie < nen; ++ie) {
      const zd e = mkz(em[ie].x, em[ie].y);
      if (!calc_one(cl, gt, ct, f, e, x + ie, nen)) {
        atomicOr(status, 2);
        break;
      }
    }
  }
  __syncthreads();
  const double inv_ntk = 1.0 / (double)ntk;
  for (int i = threadIdx.x; i < nk3 * nen; i += blockDim.x) {
    const int gi = i / nen, ie = i % nen;
    const int k = index_kpt[gi];
    const double2 po = pband[(size_t)jo * ntk + k], pv = pband[(size_t)iv * ntk + ikq[k]];
    // 1 - P_occ(j,k) P_vir(i,k+q)
    const zd fac = mkz(1.0, 0.0) - mkz(po.x, po.y) * mkz(pv.x, pv.y);
    const zd v = mkz(x[i].x, x[i].y) * inv_ntk * fac;
    dout[(size_t)ie * ldd + (size_t)p * ntk + k] = make_double2(v.x, v.y);
  }
}

size_t tetra_lds_bytes(int nk3, int nen) {
  return sizeof(double2) * (size_t)nk3 * nen + (sizeof(zd) + sizeof(int)) * 4 * 256 +
         2 * sizeof(double) * (size_t)nk3;
}

void launch_tetra_weights(int ntk, int n1, int n2, int n3, const int* imt, const int* index_kpt,
                          const int* pocc, const int* pvir, const int* ikq, const double* eband,
                          const double2* pband, double ef, double delt, double dmna, double dmnr,
                          int nen, const double2* em, size_t ldd, int npair, int* status,
                          double2* dout, hipStream_t s) {
  size_t lds = tetra_lds_bytes(n1 * n2 * n3, nen);
  if (lds > 65536)
    RP_HIP(hipFuncSetAttribute((const void*)k_tetra_weights,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_tetra_weights, dim3(npair), dim3(256), lds, s, ntk, n1, n2, n3, imt,
                     index_kpt, pocc, pvir, ikq, eband, pband, ef, delt, dmna, dmnr, nen, em, ldd,
                     status, dout);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
