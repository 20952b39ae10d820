// Generalized-tetrahedron weights x_ij(k, w) of ttrhdrn_chi0
// (m_tetrahedron_20170325.F:99-503) for every band pair, in FP64 on gfx950.
//
// One workgroup per pair.  Its k-grid accumulator x[nk3][nen] (complex FP64)
// lives in LDS; threads take (tetrahedron, dsgn) items, evaluate the closed
// forms for all frequencies and add the four corner contributions with LDS
// atomics.  Tetrahedra not cut by E_F (every gapped system) take a register
// fast path; cut ones run the reference's edge-splitting (ttct) on a private
// stack of at most kMaxSub sub-tetrahedra (an error flag is raised beyond it,
// as the reference STOPs beyond 99).
//
// The output is the contraction weight d[w][K] = x(k,w) (1 - P_occ P_vir)/NTK
// with K = pair*NTK + k (the factors of chiqw.F:1404-1406).
#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

constexpr int kMaxSub = 24;

// ---------------------------------------------------------- closed forms
// Fujiwara-Yamamoto-Ishii closed forms as evaluated in
// m_tetrahedron_20170325.F:304-451, one frequency e at a time.
__device__ inline zd sq(zd a) { return a * a; }
__device__ inline zd cu(zd a) { return a * a * a; }

__device__ zd f_sub1g30g21(zd g0, zd g1, zd e) {
  zd d = g0 - g1;
  zd x = (0.5 * (g0 + g1) + e) / sq(d);
  return x - logz((g0 + e) / (g1 + e)) * (g0 + e) * (g1 + e) / cu(d);
}
__device__ zd f_sub1g30g20(zd g0, zd g1, zd e) {
  zd d = g0 - g1;
  zd x = (g0 - 3.0 * g1 - 2.0 * e) * 0.25 / sq(d);
  return x + sq(g1 + e) * logz((g0 + e) / (g1 + e)) * 0.5 / cu(d);
}
__device__ zd f_sub1g30(zd g0, zd g1, zd g2, zd e) {
  zd x = 0.5 / (g0 - g2);
  zd p = (g0 - g2) * (g0 - g1);
  x = x - 0.5 / sq(p) * (g0 + e) * (-2.0 * (g1 + e) * (g2 + e) + (g0 + e) * (g1 + g2 + 2.0 * e)) *
              logz((g0 + e) / (g1 + e));
  x = x + 0.5 / sq(g0 - g2) * sq(g2 + e) * logz((g1 + e) / (g2 + e)) / (g1 - g2);
  return x + 0.5 / ((g0 - g2) * (g0 - g1)) * (g1 + e);
}
__device__ zd f_sub1normal(zd g0, zd g1, zd g2, zd g3, zd e) {
  zd l1 = sq(g1 + e) * logz((g1 + e) / (g0 + e));
  zd x = (-l1) / ((g0 - g2) * (g0 - g3) * (g0 - g1)) * 0.5;
  x = x - (l1 - sq(g2 + e) * logz((g2 + e) / (g0 + e))) / ((g0 - g2) * (g1 - g2) * (g2 - g3)) * 0.5;
  return x + (l1 - sq(g3 + e) * logz((g3 + e) / (g0 + e))) / ((g1 - g3) * (g0 - g3) * (g2 - g3)) * 0.5;
}
__device__ zd f_sub2simple(zd g0, zd e) { return 1.0 / ((g0 + e) * 24.0); }
__device__ zd f_sub2g30g20(zd g0, zd g1, zd e) {
  zd d = g0 - g1, d2 = d * d;
  zd x = (d * (g0 + 2.0 * g1 + 3.0 * e) - 6.0 * (g0 + e) * (g1 + e)) / (12.0 * (d2 * d));
  return x + (g0 + e) * sq(g1 + e) * logz((g0 + e) / (g1 + e)) / (d2 * d2) * 0.5;
}
__device__ zd f_sub2g30g10(zd g0, zd g2, zd e) {
  zd a = g0 + e, c = g2 + e, d = g0 - g2, d2 = d * d;
  zd x = (2.0 * sq(a) - 7.0 * a * c + 11.0 * sq(c)) / (36.0 * (d2 * d));
  return x - cu(c) * logz(a / c) / (6.0 * (d2 * d2));
}
__device__ zd f_sub2g10(zd g0, zd g2, zd g3, zd e) {
  zd a = g0 + e, c2 = g2 + e, c3 = g3 + e;
  zd x = a * (sq(a) + 5.0 * c2 * c3 - 3.0 * a * (g2 + g3 + 2.0 * e)) /
         (12.0 * sq(g0 - g2) * sq(g0 - g3));
  x = x - cu(c2) * logz(c2 / a) / (6.0 * cu(g0 - g2) * (g2 - g3));
  return x + cu(c3) * logz(c3 / a) / (6.0 * cu(g0 - g3) * (g2 - g3));
}
__device__ zd f_sub2g30(zd g0, zd g1, zd g2, zd e) {
  zd a = g0 + e, b = g1 + e, c = g2 + e;
  zd d01 = g0 - g1, d02 = g0 - g2, d12 = g1 - g2;
  zd x = (sq(a) * d12 + sq(b) * d02) / (6.0 * d02 * sq(d01) * d12);
  x = x - sq(a) * (2.0 * b * d02 + c * d01) * logz(a / c) / (6.0 * sq(d02) * cu(d01));
  return x + sq(b) * (2.0 * a * d12 - c * d01) * logz(b / c) / (6.0 * cu(d01) * sq(d12));
}
__device__ zd f_sub2normal(zd g0, zd g1, zd g2, zd g3, zd e) {
  zd a = g0 + e, b = g1 + e, c = g2 + e, d = g3 + e;
  zd x = -sq(b) / (6.0 * (g0 - g1) * (g1 - g2) * (g1 - g3));
  x = x - cu(c) * logz(c / b) / (6.0 * (g0 - g2) * sq(g1 - g2) * (g2 - g3));
  x = x + cu(d) * logz(d / b) / (6.0 * (g0 - g3) * sq(g1 - g3) * (g2 - g3));
  return x + cu(a) * logz(a / b) / (6.0 * sq(g0 - g1) * (g0 - g2) * (g0 - g3));
}

struct Class {
  int n2, i3;       // number of vertices in degenerate pairs, number of pairs
  int a3[4], a4[4]; // degenerate vertices (discovery order), the others
};

// chk1 (m_tetrahedron_20170325.F:457-503): vertices compared in the
// reference's rotated order p = mod(i,4)+1.
__device__ Class classify(double dmna, double dmnr, const zd* g) {
  Class c;
  int ia2[4] = {0, 0, 0, 0};
  int i3 = 0, i4 = 0;
  for (int v = 0; v < 4; ++v) { c.a3[v] = 0; c.a4[v] = 0; }
  for (int i1 = 1; i1 <= 3; ++i1)
    for (int i2 = i1 + 1; i2 <= 4; ++i2) {
      int p = i1 & 3, q = i2 & 3;
      double d1 = absz(g[p] - g[q]);
      double d2 = absz(g[p] + g[q]) * 0.5;
      if (d1 < dmna || d1 / d2 < dmnr) {
        ++i3;
        if (!ia2[p]) { c.a3[i4++] = p; ia2[p] = 1; }
        if (!ia2[q]) { c.a3[i4++] = q; ia2[q] = 1; }
      }
    }
  int i5 = 0;
  for (int i1 = 1; i1 <= 4; ++i1) {
    int p = i1 & 3;
    if (!ia2[p]) c.a4[i5++] = p;
  }
  c.n2 = i4;
  c.i3 = i3;
  return c;
}

// calc (m_tetrahedron_20170325.F:209-302) for one frequency: out[v] is the
// contribution to vertex v (f included).  Returns false for an impossible class.
__device__ bool calc_one(const Class& c, const zd* g, const zd* f, zd e, zd* out) {
  out[0] = out[1] = out[2] = out[3] = mkz(0, 0);
  const int* a3 = c.a3;
  const int* a4 = c.a4;
  switch (c.n2) {
    case 0: {
      zd c1 = f_sub2normal(g[a4[1]], g[a4[0]], g[a4[2]], g[a4[3]], e);
      zd c2 = f_sub2normal(g[a4[0]], g[a4[1]], g[a4[2]], g[a4[3]], e);
      zd c3 = f_sub2normal(g[a4[0]], g[a4[2]], g[a4[1]], g[a4[3]], e);
      zd c4 = f_sub1normal(g[a4[0]], g[a4[3]], g[a4[1]], g[a4[2]], e);
      out[a4[0]] = f[a4[0]] * c1;
      out[a4[1]] = f[a4[1]] * c2;
      out[a4[2]] = f[a4[2]] * c3;
      out[a4[3]] = f[a4[3]] * (c4 - c3 - c2 - c1);
      return true;
    }
    case 2: {
      zd q1 = (g[a3[0]] + g[a3[1]]) / 2.0;
      zd c1 = f_sub2g10(q1, g[a4[0]], g[a4[1]], e);
      zd c2 = f_sub2g30(q1, g[a4[0]], g[a4[1]], e);
      zd c3 = f_sub1g30(q1, g[a4[1]], g[a4[0]], e);
      out[a3[0]] = f[a3[0]] * c1;
      out[a3[1]] = f[a3[1]] * c1;
      out[a4[0]] = f[a4[0]] * c2;
      out[a4[1]] = f[a4[1]] * (c3 - c2 - c1 - c1);
      return true;
    }
    case 3: {
      zd q1 = (g[a3[0]] + g[a3[1]] + g[a3[2]]) / 3.0;
      zd c1 = f_sub2g30g10(q1, g[a4[0]], e);
      zd c2 = f_sub1g30g20(q1, g[a4[0]], e);
      out[a3[0]] = f[a3[0]] * c1;
      out[a3[1]] = f[a3[1]] * c1;
      out[a3[2]] = f[a3[2]] * c1;
      out[a4[0]] = f[a4[0]] * (c2 - c1 - c1 - c1);
      return true;
    }
    case 4:
      if (c.i3 == 2) {
        zd q1 = (g[a3[0]] + g[a3[1]]) / 2.0, q2 = (g[a3[2]] + g[a3[3]]) / 2.0;
        zd c1 = f_sub2g30g20(q1, q2, e);
        zd c2 = f_sub1g30g21(q1, q2, e);
        out[a3[0]] = f[a3[0]] * c1;
        out[a3[1]] = f[a3[1]] * c1;
        out[a3[2]] = f[a3[2]] * (c2 - c1 - c1) * 0.5;
        out[a3[3]] = f[a3[3]] * (c2 - c1 - c1) * 0.5;
        return true;
      } else if (c.i3 >= 3 && c.i3 <= 6) {
        zd q1 = (g[a3[0]] + g[a3[1]] + g[a3[2]] + g[a3[3]]) / 4.0;
        zd c1 = f_sub2simple(q1, e);
        for (int v = 0; v < 4; ++v) out[a3[v]] = c1 * f[a3[v]];
        return true;
      }
      return false;
    default:
      return false;
  }
}

__device__ inline void lds_add(double2* p, zd v) {
  atomicAdd(&p->x, v.x);
  atomicAdd(&p->y, v.y);
}

// ttct (m_tetrahedron_20170325.F:143-207) on a private stack.
struct SubTets {
  double w[kMaxSub][4][4];  // w[t][v][j]: weight of corner j in vertex v
  double vol[kMaxSub];
  int n;
};

__device__ bool ttct(double dval, const double* dvtx, int imod, SubTets& s) {
  unsigned char bv[kMaxSub][4];
  unsigned char keep[kMaxSub];
  for (int t = 0; t < kMaxSub; ++t) { keep[t] = 0; for (int v = 0; v < 4; ++v) bv[t][v] = 0; }
  int itet = 0;
  while (itet < s.n) {
    const int t = itet++;
    double v0[4];
    for (int i = 0; i < 4; ++i)
      v0[i] = s.w[t][i][0] * dvtx[0] + s.w[t][i][1] * dvtx[1] + s.w[t][i][2] * dvtx[2] +
              s.w[t][i][3] * dvtx[3];
    int i1 = -1, i2 = -1;
    for (int a = 0; a < 3 && i1 < 0; ++a) {
      if (bv[t][a]) continue;
      for (int b = a + 1; b < 4; ++b) {
        if (bv[t][b]) continue;
        if ((v0[a] - dval) * (dval - v0[b]) > 0.0) { i1 = a; i2 = b; break; }
      }
    }
    if (i1 >= 0) {
      double d1 = (v0[i2] - dval) / (v0[i2] - v0[i1]);
      double d2 = (dval - v0[i1]) / (v0[i2] - v0[i1]);
      if (s.n + 1 > kMaxSub) return false;
      const int n = s.n++;
      for (int v = 0; v < 4; ++v) {
        bv[n][v] = bv[t][v];
        for (int j = 0; j < 4; ++j) s.w[n][v][j] = s.w[t][v][j];
      }
      bv[n][i1] = 1;
      for (int j = 0; j < 4; ++j) s.w[n][i1][j] = s.w[t][i1][j] * d1 + s.w[t][i2][j] * d2;
      s.vol[n] = s.vol[t] * d1;
      bv[t][i2] = 1;
      for (int j = 0; j < 4; ++j) s.w[t][i2][j] = s.w[t][i1][j] * d1 + s.w[t][i2][j] * d2;
      s.vol[t] = s.vol[t] * d2;
      --itet;
    } else {
      for (int i = 0; i < 4; ++i) {
        if (bv[t][i]) continue;
        keep[t] = keep[t] || ((double)imod * (v0[i] - dval) > 0.0);
      }
    }
  }
  int k = 0;
  for (int t = 0; t < s.n; ++t) {
    if (!keep[t]) continue;
    if (k != t) {
      for (int v = 0; v < 4; ++v)
        for (int j = 0; j < 4; ++j) s.w[k][v][j] = s.w[t][v][j];
      s.vol[k] = s.vol[t];
    }
    ++k;
  }
  s.n = k;
  return true;
}

// Does ttct need to split?  (Any unflagged pair strictly on both sides.)
__device__ inline bool straddles(double f, const double* v) {
  for (int a = 0; a < 3; ++a)
    for (int b = a + 1; b < 4; ++b)
      if ((v[a] - f) * (f - v[b]) > 0.0) return true;
  return false;
}

// grid: npair blocks.  eg[band][k] energies (Ha, full mesh, FP64); the
// tetrahedron sees them rounded to single precision (CMPLX, chiqw.F:1353,1356).
__global__ void __launch_bounds__(256) k_tetra_weights(
    int ntk, int n1, int n2, int n3, const int* __restrict__ imt,
    const int* __restrict__ index_kpt, const int* __restrict__ pocc,
    const int* __restrict__ pvir, const int* __restrict__ ikq, const double* __restrict__ eband,
    const double2* __restrict__ pband, double ef, double delt, double dmna, double dmnr, int nen,
    const double2* __restrict__ em, size_t ldd, int* __restrict__ status,
    double2* __restrict__ dout) {
  extern __shared__ double2 lds[];
  const int nk3 = n1 * n2 * n3;
  double2* x = lds;                              // [nk3][nen]
  double* e1 = (double*)(x + (size_t)nk3 * nen);  // [nk3]
  double* e2 = e1 + nk3;
  const int p = blockIdx.x;
  const int jo = pocc[p], iv = pvir[p];
  for (int i = threadIdx.x; i < nk3 * nen; i += blockDim.x) x[i] = make_double2(0, 0);
  for (int gi = threadIdx.x; gi < nk3; gi += blockDim.x) {
    int k = index_kpt[gi];
    e1[gi] = (double)(float)eband[(size_t)jo * ntk + k];
    e2[gi] = (double)(float)eband[(size_t)iv * ntk + ikq[k]];
  }
  __syncthreads();
  const int ntet = 6 * nk3;
  for (int it = threadIdx.x; it < 2 * ntet; it += blockDim.x) {
    const int t = it % ntet;
    const double dsgn = it < ntet ? 1.0 : -1.0;
    int c[4];
    double ee1[4], ee2[4];
    for (int v = 0; v < 4; ++v) {
      c[v] = imt[4 * t + v];
      ee1[v] = e1[c[v]];
      ee2[v] = e2[c[v]];
    }
    if (!straddles(ef, ee2) && !straddles(ef, ee1)) {
      // no cut: one sub-tetrahedron (identity weights) or none
      bool k2 = false, k1 = false;
      for (int v = 0; v < 4; ++v) { k2 = k2 || (ee2[v] - ef > 0.0); k1 = k1 || (-(ee1[v] - ef) > 0.0); }
      if (!(k1 && k2)) continue;
      zd g[4], f[4];
      for (int v = 0; v < 4; ++v) {
        double s = (ee2[v] - ee1[v]);
        g[v] = mkz(-s, delt) * dsgn;
        f[v] = mkz(dsgn, 0.0);
      }
      Class cl = classify(dmna, dmnr, g);
      for (int ie = 0; ie < nen; ++ie) {
        zd out[4];
        zd e = mkz(em[ie].x, em[ie].y);
        if (!calc_one(cl, g, f, e, out)) { atomicOr(status, 2); break; }
        for (int v = 0; v < 4; ++v) lds_add(&x[(size_t)c[v] * nen + ie], out[v]);
      }
      continue;
    }
    SubTets s;
    for (int v = 0; v < 4; ++v)
      for (int j = 0; j < 4; ++j) s.w[0][v][j] = v == j ? 1.0 : 0.0;
    s.vol[0] = 1.0;
    s.n = 1;
    if (!ttct(ef, ee2, +1, s) || !ttct(ef, ee1, -1, s)) { atomicOr(status, 1); continue; }
    for (int u = 0; u < s.n; ++u) {
      zd g[4], f[4];
      for (int v = 0; v < 4; ++v) {
        double sum = 0.0;
        for (int j = 0; j < 4; ++j) sum += (ee2[j] - ee1[j]) * s.w[u][v][j];
        g[v] = mkz(-sum, delt) * dsgn;
        f[v] = mkz(s.vol[u] * dsgn, 0.0);
      }
      Class cl = classify(dmna, dmnr, g);
      for (int ie = 0; ie < nen; ++ie) {
        zd out[4];
        zd e = mkz(em[ie].x, em[ie].y);
        if (!calc_one(cl, g, f, e, out)) { atomicOr(status, 2); break; }
        for (int v = 0; v < 4; ++v)
          for (int j = 0; j < 4; ++j) {
            double w = s.w[u][v][j];
            if (w != 0.0) lds_add(&x[(size_t)c[j] * nen + ie], out[v] * w);
          }
      }
    }
  }
  __syncthreads();
  const double inv_ntk = 1.0 / (double)ntk;
  for (int i = threadIdx.x; i < nk3 * nen; i += blockDim.x) {
    const int gi = i / nen, ie = i % nen;
    const int k = index_kpt[gi];
    double2 po = pband[(size_t)jo * ntk + k], pv = pband[(size_t)iv * ntk + ikq[k]];
    // 1 - P_occ(j,k) P_vir(i,k+q)
    zd fac = mkz(1.0, 0.0) - mkz(po.x, po.y) * mkz(pv.x, pv.y);
    zd v = mkz(x[i].x, x[i].y) * inv_ntk * fac;
    dout[(size_t)ie * ldd + (size_t)p * ntk + k] = make_double2(v.x, v.y);
  }
}

size_t tetra_lds_bytes(int nk3, int nen) {
  return sizeof(double2) * (size_t)nk3 * nen + 2 * sizeof(double) * (size_t)nk3;
}

void launch_tetra_weights(int ntk, int n1, int n2, int n3, const int* imt, const int* index_kpt,
                          const int* pocc, const int* pvir, const int* ikq, const double* eband,
                          const double2* pband, double ef, double delt, double dmna, double dmnr,
                          int nen, const double2* em, size_t ldd, int npair, int* status,
                          double2* dout, hipStream_t s) {
  size_t lds = tetra_lds_bytes(n1 * n2 * n3, nen);
  RP_HIP(hipFuncSetAttribute((const void*)k_tetra_weights,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_tetra_weights, dim3(npair), dim3(256), lds, s, ntk, n1, n2, n3, imt,
                     index_kpt, pocc, pvir, ikq, eband, pband, ef, delt, dmna, dmnr, nen, em, ldd,
                     status, dout);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
