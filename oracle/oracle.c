/*
 * oracle.c -- CPU restatement of RESPACK's cRPA polarisation path (src/chiqw).
 *
 * TEST INFRASTRUCTURE: the checker for librespack_gpu.so and the CPU baseline
 * timed beside it.  Never linked, loaded or called by the product.
 *
 * Every function restates one reference routine (cited per function, paths
 * relative to the reference's src/).  Branch structure and evaluation order
 * follow the reference where they decide results (sphere membership, k-folding
 * thresholds, tetrahedron degeneracy classes, single-precision roundings).
 *
 * Single-precision roundings of the reference, reproduced here on purpose:
 * Fortran's CMPLX(x,y) without KIND returns default (single precision) complex,
 * so the reference rounds to float
 *   - u_k(r) after each inverse FFT   (chiqw_sub.F:143, :162),
 *   - the pair density M(G)           (chiqw_sub.F:201),
 *   - band energies fed to the tetrahedron weights (chiqw.F:1353, :1356,
 *     :1836-1837), and the q = 0 Drude term (chiqw_sub.F:512).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RO_PI 3.14159265358979323846
static const double ro_au = 27.21151; /* chiqw/config.h:19 */

static inline ro_z fround(ro_z z) { /* CMPLX(x,y) without KIND */
  return (double)(float)creal(z) + I * (double)(float)cimag(z);
}

/* ------------------------------------------------------------------ grids */

int ro_algn235(int n) { /* est_nwx2.f90:1-42 */
  for (;; ++n) {
    int i = n;
    while (i != 1) {
      if (i % 2 == 0) i /= 2;
      else if (i % 3 == 0) i /= 3;
      else if (i % 5 == 0) i /= 5;
      else break;
    }
    if (i == 1) return n;
  }
}

void ro_gen_grid(int nt, double e, ro_z *em) { /* gen_grid.f90:1-39 */
  const double dlt = 0.05 / ro_au, ratio = 0.1, eps = 1.0e-6;
  if (nt == 1) { em[0] = 0.0; return; }
  int m = (int)(ratio * (double)nt);
  int n = nt - m;
  double alp = e / dlt, x = 1.5, dl = 1.0;
  while (dl > eps) {
    double f = pow(x, n - 1) - alp * x + alp - 1.0;
    double g = (double)(n - 1) * pow(x, n - 2) - alp;
    double x1 = x - f / g;
    dl = fabs(x1 - x);
    x = x1;
  }
  double a = dlt / (x - 1.0), b = log(x);
  for (int i = 1; i <= n; ++i) em[i - 1] = a * exp(b * (double)(i - 1)) - a;
  double c = e * exp(-log(2.0) * (double)n / (double)m);
  double d = log(2.0) / (double)m;
  for (int i = n + 1; i <= n + m; ++i) em[i - 1] = c * exp(d * (double)i);
}

int ro_make_kg0(int ntg, const double *b1, const double *b2, const double *b3,
                double gcut, double q1, double q2, double q3, int32_t *kg) {
  /* chiqw_sub.F:773-812: g1 outer, g3 inner, over [-100,100]^3 */
  int n = 0;
  for (int i1 = -100; i1 <= 100; ++i1)
    for (int i2 = -100; i2 <= 100; ++i2)
      for (int i3 = -100; i3 <= 100; ++i3) {
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = (q1 + (double)i1) * b1[d] + (q2 + (double)i2) * b2[d] +
                 (q3 + (double)i3) * b3[d];
        double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (s <= gcut) {
          if (n < ntg) { kg[3 * n] = i1; kg[3 * n + 1] = i2; kg[3 * n + 2] = i3; }
          ++n;
        }
      }
  return n;
}

void ro_kcheck(double *k, int32_t *rw) { /* chiqw_sub.F:668-716 */
  const double d = 1.0e-6;
  for (int i = 0; i < 3; ++i) {
    if (k[i] > 1.5 + d) { k[i] -= 2.0; rw[i] = -2; }
    if (k[i] > 0.5 + d) { k[i] -= 1.0; rw[i] = -1; }
    if (k[i] <= -1.5 + d) { k[i] += 2.0; rw[i] = 2; }
    if (k[i] <= -0.5 + d) { k[i] += 1.0; rw[i] = 1; }
  }
}

void ro_kcheck_trs(double *k, int32_t *rw) { /* chiqw_sub.F:718-771 */
  const double d = -1.0e-6;
  for (int i = 0; i < 3; ++i) {
    if (k[i] >= 1.5 + d) { k[i] -= 2.0; rw[i] = -2; }
    if (k[i] >= 0.5 + d) { k[i] -= 1.0; rw[i] = -1; }
    if (k[i] < -1.5 + d) { k[i] += 2.0; rw[i] = 2; }
    if (k[i] < -0.5 + d) { k[i] += 1.0; rw[i] = 1; }
  }
}

void ro_search_list(const double *sqi, double *sq) { /* chiqw_sub.F:76-95 */
  const double d = 1.0e-6;
  for (int i = 0; i < 3; ++i) sq[i] = sqi[i];
  for (int i = 0; i < 2; ++i) {
    if (sq[i] > 1.5 + d) sq[i] -= 2.0;
    if (sq[i] > 0.5 + d) sq[i] -= 1.0;
    if (sq[i] <= -1.5 + d) sq[i] += 2.0;
    if (sq[i] <= -0.5 + d) sq[i] += 1.0;
  }
  /* the third component tests -0.5 before -1.5 (chiqw_sub.F:91-92) */
  if (sq[2] > 1.5 + d) sq[2] -= 2.0;
  if (sq[2] > 0.5 + d) sq[2] -= 1.0;
  if (sq[2] <= -0.5 + d) sq[2] += 1.0;
  if (sq[2] <= -1.5 + d) sq[2] += 2.0;
}

int ro_search_kq(int ntk, const double *sk0, double q1, double q2, double q3,
                 int ik, int32_t *sg) { /* chiqw_sub.F:1-75 */
  const double d = 1.0e-6;
  double s[3] = {sk0[3 * ik] + q1, sk0[3 * ik + 1] + q2, sk0[3 * ik + 2] + q3};
  sg[0] = sg[1] = sg[2] = 0;
  for (int i = 0; i < 3; ++i) {
    if (s[i] > 1.5 + d) { s[i] -= 2.0; sg[i] = 2; }
    if (s[i] > 0.5 + d) { s[i] -= 1.0; sg[i] = 1; }
    if (s[i] <= -1.5 + d) { s[i] += 2.0; sg[i] = -2; }
    if (s[i] <= -0.5 + d) { s[i] += 1.0; sg[i] = -1; }
  }
  int ikq = -1;
  for (int jk = 0; jk < ntk; ++jk)
    if (fabs(sk0[3 * jk] - s[0]) < 1e-6 && fabs(sk0[3 * jk + 1] - s[1]) < 1e-6 &&
        fabs(sk0[3 * jk + 2] - s[2]) < 1e-6)
      ikq = jk;
  return ikq;
}

void ro_est_nkbi(int n, const double *sk, int32_t *nkb) { /* est_NTK.f90:1-39 */
  for (int d = 0; d < 3; ++d) {
    double x = 1.0;
    for (int i = 0; i < n; ++i) {
      double a = fabs(sk[3 * i + d]);
      if (a < 1.0e-7) continue;
      if (a < x) x = a;
    }
    nkb[d] = (int32_t)lround(1.0 / x);
  }
}

int ro_unfold_k(int nk_irr, int nsym, const double *ski, const int32_t *rg,
                double *sk0, int32_t *numirr, int32_t *numrot, int32_t *trs,
                int32_t *rw) { /* chiqw.F:437-480 */
  int jk = 0;
  for (int ik = 0; ik < nk_irr; ++ik)
    for (int iop = 0; iop < nsym; ++iop) {
      const int32_t *r = rg + 9 * iop;
      for (int pass = 0; pass < 2; ++pass) {
        double k[3];
        int32_t w[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i)
          k[i] = (double)r[i] * ski[3 * ik] + (double)r[i + 3] * ski[3 * ik + 1] +
                 (double)r[i + 6] * ski[3 * ik + 2];
        double sgn = 1.0;
        if (pass == 0) ro_kcheck(k, w);
        else { ro_kcheck_trs(k, w); sgn = -1.0; }
        int found = 0;
        for (int j = 0; j < jk; ++j)
          if (fabs(sk0[3 * j] - sgn * k[0]) < 1e-4 &&
              fabs(sk0[3 * j + 1] - sgn * k[1]) < 1e-4 &&
              fabs(sk0[3 * j + 2] - sgn * k[2]) < 1e-4) { found = 1; break; }
        if (found) continue;
        for (int i = 0; i < 3; ++i) { sk0[3 * jk + i] = sgn * k[i]; rw[3 * jk + i] = w[i]; }
        numirr[jk] = ik; numrot[jk] = iop; trs[jk] = pass == 0 ? 1 : -1;
        ++jk;
      }
    }
  return jk;
}

void ro_make_index_kpt(int ntk, int n1, int n2, int n3, const double *sk0,
                       int32_t *idx) { /* chiqw_sub.F:324-384 */
  for (int ik = 0; ik < ntk; ++ik) {
    double x = sk0[3 * ik] * n1, y = sk0[3 * ik + 1] * n2, z = sk0[3 * ik + 2] * n3;
    long ix, iy, iz;
    if (ntk % 2 != 0) {
      ix = lround(x + (n1 - 1.0) / 2.0) + 1; iy = lround(y + (n2 - 1.0) / 2.0) + 1;
      iz = lround(z + (n3 - 1.0) / 2.0) + 1;
    } else {
      ix = lround(x + n1 / 2.0); iy = lround(y + n2 / 2.0); iz = lround(z + n3 / 2.0);
    }
    idx[(ix - 1) + n1 * ((iy - 1) + n2 * (iz - 1))] = ik;
  }
}

void ro_judge_fermi(int nk, int ntb, const double *eig, double ef, int32_t *win,
                    int32_t *tindx, double *eave, double *bmax, double *bmin) {
  /* chiqw_sub.F:570-613 (Tindx is 1-based band number or 0, as there) */
  for (int ib = 0; ib < ntb; ++ib) {
    double top = 0, btm = 0;
    for (int ik = 0; ik < nk; ++ik) {
      double e = eig[ib + (size_t)ntb * ik];
      if (ik == 0) top = btm = e;
      if (e < btm) btm = e;
      if (e > top) top = e;
    }
    int fs = 0;
    tindx[ib] = 0;
    if (btm > ef) fs = 2;
    if (btm <= ef && top >= ef) { fs = 1; tindx[ib] = ib + 1; }
    if (top < ef) fs = 0;
    win[ib] = fs;
    eave[ib] = (top + btm) / 2.0;
    bmax[ib] = top;
    bmin[ib] = btm;
  }
}

void ro_make_c0(int itrs, int ng, const int32_t *kg, const int32_t *rw,
                const double *rginv, const int32_t *pg, int nnp, int l1, int l2,
                int l3, const int32_t *pack, const ro_z *occ, ro_z *c0) {
  /* chiqw_sub.F:264-322 */
  int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1;
  double s = itrs == 1 ? 1.0 : -1.0;
  for (int ig = 0; ig < ng; ++ig) {
    int i1 = (int)s * kg[3 * ig], j1 = (int)s * kg[3 * ig + 1], k1 = (int)s * kg[3 * ig + 2];
    int i2 = i1 + rw[0], j2 = j1 + rw[1], k2 = k1 + rw[2];
    int i3 = (int)rginv[0] * i2 + (int)rginv[3] * j2 + (int)rginv[6] * k2;
    int j3 = (int)rginv[1] * i2 + (int)rginv[4] * j2 + (int)rginv[7] * k2;
    int k3 = (int)rginv[2] * i2 + (int)rginv[5] * j2 + (int)rginv[8] * k2;
    int jg = pack[(i3 + l1) + n1 * ((j3 + l2) + n2 * (k3 + l3))] - 1;
    double ph = 2.0 * RO_PI * ((double)i1 * pg[0] + (double)j1 * pg[1] + (double)k1 * pg[2]);
    ro_z pf = cexp(-I * ph / (double)nnp);
    ro_z v = occ[jg] * pf;
    c0[ig] = itrs == 1 ? v : conj(v);
  }
}

/* ------------------------------------------------------------ tetrahedra */

void ro_ttrhdrn_mkidx(int n1, int n2, int n3, const double *b1, const double *b2,
                      const double *b3, int32_t *idx) {
  /* m_tetrahedron_20170325.F:523-584 (libtetrabz corner table) */
  int ng[3] = {n1, n2, n3};
  double bv[3][3], b3v[4][3], l[4];
  for (int d = 0; d < 3; ++d) { bv[0][d] = b1[d] / n1; bv[1][d] = b2[d] / n2; bv[2][d] = b3[d] / n3; }
  for (int d = 0; d < 3; ++d) {
    b3v[0][d] = -bv[0][d] + bv[1][d] + bv[2][d];
    b3v[1][d] = bv[0][d] - bv[1][d] + bv[2][d];
    b3v[2][d] = bv[0][d] + bv[1][d] - bv[2][d];
    b3v[3][d] = bv[0][d] + bv[1][d] + bv[2][d];
  }
  int itype = 0;
  for (int i = 0; i < 4; ++i) {
    l[i] = b3v[i][0] * b3v[i][0] + b3v[i][1] * b3v[i][1] + b3v[i][2] * b3v[i][2];
    if (l[i] < l[itype]) itype = i;
  }
  int v0[4] = {0, 0, 0, 0}, dv[4][4] = {{0}}; /* dv[col][row] */
  for (int i = 0; i < 4; ++i) dv[i][i] = 1;
  v0[itype] = 1;
  dv[itype][itype] = -1;
  int vv[6][4][3], it = 0;
  for (int i1 = 0; i1 < 3; ++i1)
    for (int i2 = 0; i2 < 3; ++i2) {
      if (i2 == i1) continue;
      for (int i3 = 0; i3 < 3; ++i3) {
        if (i3 == i1 || i3 == i2) continue;
        for (int d = 0; d < 3; ++d) {
          vv[it][0][d] = v0[d];
          vv[it][1][d] = vv[it][0][d] + dv[i1][d];
          vv[it][2][d] = vv[it][1][d] + dv[i2][d];
          vv[it][3][d] = vv[it][2][d] + dv[i3][d];
        }
        ++it;
      }
    }
  int nt = 0;
  for (int k3 = 0; k3 < n3; ++k3)
    for (int k2 = 0; k2 < n2; ++k2)
      for (int k1 = 0; k1 < n1; ++k1)
        for (it = 0; it < 6; ++it, ++nt)
          for (int ii = 0; ii < 4; ++ii) {
            int kv[3] = {k1 + vv[it][ii][0], k2 + vv[it][ii][1], k3 + vv[it][ii][2]};
            for (int d = 0; d < 3; ++d) kv[d] = ((kv[d] % ng[d]) + ng[d]) % ng[d];
            idx[4 * nt + ii] = kv[0] + n1 * kv[1] + n1 * n2 * kv[2];
          }
}

/* The generalized-tetrahedron weights (closed forms, chk1, calc, ttct,
 * ttrhdrn_chi0) in double: see tet_forms.h, which tet_quad.c instantiates in
 * __float128 as well. */
#define TF_Z ro_z
#define TF_LOG clog
#define TF(name) d_##name
#include "tet_forms.h"

/* The closed forms above with the reference's argument lists (g[0..3] as its
 * g0..g3; the unused ones ignored): which as in the checker's rpref_tet_sub. */
int ro_tet_sub(int which, const ro_z *g, int ne, const ro_z *em, ro_z *x) {
  for (int ie = 0; ie < ne; ++ie) {
    ro_z e = em[ie];
    switch (which) {
    case 0: x[ie] = 1.0 / ((g[0] + e) * 6.0); break;
    case 1: x[ie] = d_sub1g30g21(g[0], g[1], e); break;
    case 2: x[ie] = d_sub1g30g20(g[0], g[1], e); break;
    case 3: x[ie] = d_sub1g30(g[0], g[1], g[2], e); break;
    case 4: x[ie] = d_sub1normal(g[0], g[1], g[2], g[3], e); break;
    case 5: x[ie] = d_sub2simple(g[0], e); break;
    case 6: x[ie] = d_sub2g30g20(g[0], g[1], e); break;
    case 7: x[ie] = d_sub2g30g20(g[1], g[0], e); break; /* sub2g30g21(g0,g1,..) */
    case 8: x[ie] = d_sub2g30g10(g[0], g[2], e); break;
    case 9: x[ie] = d_sub2g10(g[0], g[2], g[3], e); break;
    case 10: x[ie] = d_sub2g30(g[0], g[1], g[2], e); break;
    case 11: x[ie] = d_sub2normal(g[0], g[1], g[2], g[3], e); break;
    default: return -1;
    }
  }
  return 0;
}

int ro_tet_series(const ro_z *gt, int ne, const ro_z *em, ro_z *w, int32_t *used) {
  int n = 0;
  for (int ie = 0; ie < ne; ++ie) {
    used[ie] = tet_series(gt, em[ie], w + 4 * ie);
    n += used[ie];
  }
  return n;
}

int ro_ttrhdrn_calc(double dmna, double dmnr, const ro_z *fk, const ro_z *gk, int ne,
                    const ro_z *em, ro_z *xo, int use_series) {
  const int c[4] = {0, 1, 2, 3};
  return d_calc(dmna, dmnr, c, fk, gk, ne, em, xo, use_series);
}

void ro_ttrhdrn_chi0(double dmna, double dmnr, int n1, int n2, int n3,
                     const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                     double f2, double delt, double dsgn, int ne, const ro_z *em,
                     ro_z *xo) {
  d_chi0(dmna, dmnr, n1, n2, n3, idx, e1, e2, f1, f2, delt, dsgn, ne, em, xo, 1);
}

void ro_ttrhdrn_chi0_closed(double dmna, double dmnr, int n1, int n2, int n3,
                            const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                            double f2, double delt, double dsgn, int ne, const ro_z *em,
                            ro_z *xo) {
  d_chi0(dmna, dmnr, n1, n2, n3, idx, e1, e2, f1, f2, delt, dsgn, ne, em, xo, 0);
}

void ro_ttrhdrn_simple(double dmna, double dmnr, int n1, int n2, int n3,
                       const int32_t *idx, const ro_z *fk, const ro_z *gk, ro_z *xo) {
  /* m_tetrahedron_20170325.F:63-97: calc over every tetrahedron, ne = 1, em = 0 */
  ro_z em0 = 0.0;
  int nt = n1 * n2 * n3 * 6;
  for (int it = 0; it < nt; ++it) {
    int c[4] = {idx[4 * it], idx[4 * it + 1], idx[4 * it + 2], idx[4 * it + 3]};
    d_calc(dmna, dmnr, c, fk, gk, 1, &em0, xo, 1);
  }
}

/* ---------------------------------------------------------- pair density */

/* One axis of a separable DFT: a[(i*stride)] over n points, sign s. */
static void dft_axis(int n, int s, ro_z *a, size_t stride, ro_z *tmp, const ro_z *w) {
  for (int k = 0; k < n; ++k) {
    ro_z acc = 0.0;
    for (int j = 0; j < n; ++j) {
      int p = (int)(((long)j * k) % n);
      ro_z t = s > 0 ? conj(w[p]) : w[p];
      acc += a[j * stride] * t;
    }
    tmp[k] = acc;
  }
  for (int k = 0; k < n; ++k) a[k * stride] = tmp[k];
}

static void dft3_inplace(int nx, int ny, int nz, int dir, ro_z *a) {
  int nm = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
  ro_z *tmp = malloc(sizeof(ro_z) * nm);
  ro_z *wx = malloc(sizeof(ro_z) * nx), *wy = malloc(sizeof(ro_z) * ny),
       *wz = malloc(sizeof(ro_z) * nz);
  for (int j = 0; j < nx; ++j) wx[j] = cos(2 * RO_PI * j / nx) - I * sin(2 * RO_PI * j / nx);
  for (int j = 0; j < ny; ++j) wy[j] = cos(2 * RO_PI * j / ny) - I * sin(2 * RO_PI * j / ny);
  for (int j = 0; j < nz; ++j) wz[j] = cos(2 * RO_PI * j / nz) - I * sin(2 * RO_PI * j / nz);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) dft_axis(nz, dir, a + x + (size_t)nx * y, (size_t)nx * ny, tmp, wz);
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) dft_axis(ny, dir, a + x + (size_t)nx * ny * z, nx, tmp, wy);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) dft_axis(nx, dir, a + (size_t)nx * (y + (size_t)ny * z), 1, tmp, wx);
  if (dir < 0) {
    double s = 1.0 / ((double)nx * ny * nz);
    for (size_t i = 0; i < (size_t)nx * ny * nz; ++i) a[i] *= s;
  }
  free(tmp); free(wx); free(wy); free(wz);
}

void ro_dft3(int nx, int ny, int nz, int dir, const ro_z *in, ro_z *out) {
  memcpy(out, in, sizeof(ro_z) * (size_t)nx * ny * nz);
  dft3_inplace(nx, ny, nz, dir, out);
}

static inline int wrapi(int g, int n) { return ((g % n) + n) % n; }

/* calc_InterStateMatrix (chiqw_sub.F:97-205): u_k = bw(C_k), u_kq = bw(C_kq),
 * both rounded to single precision; u_kq *= exp(-i 2pi r.dG/N);
 * f = conj(u_kq) u_k; M(G) = fw(f)[-G], rounded to single precision.
 * work holds 3*nwx*nwy*nwz complex. */
static void pairdens_fft(int ng_k, const int32_t *kg_k, const ro_z *c_k, int ng_kq,
                         const int32_t *kg_kq, const ro_z *c_kq, int nx, int ny,
                         int nz, int ng, const int32_t *lg, const int32_t *sg,
                         ro_z *m, ro_z *work) {
  size_t nr = (size_t)nx * ny * nz;
  ro_z *uk = work, *ukq = work + nr;
  memset(uk, 0, sizeof(ro_z) * nr);
  memset(ukq, 0, sizeof(ro_z) * nr);
  for (int ig = 0; ig < ng_k; ++ig)
    uk[wrapi(kg_k[3 * ig], nx) + nx * (wrapi(kg_k[3 * ig + 1], ny) + (size_t)ny * wrapi(kg_k[3 * ig + 2], nz))] = c_k[ig];
  for (int ig = 0; ig < ng_kq; ++ig)
    ukq[wrapi(kg_kq[3 * ig], nx) + nx * (wrapi(kg_kq[3 * ig + 1], ny) + (size_t)ny * wrapi(kg_kq[3 * ig + 2], nz))] = c_kq[ig];
  dft3_inplace(nx, ny, nz, +1, uk);
  dft3_inplace(nx, ny, nz, +1, ukq);
  double d1 = sg[0], d2 = sg[1], d3 = sg[2];
  for (int r3 = 0; r3 < nz; ++r3)
    for (int r2 = 0; r2 < ny; ++r2)
      for (int r1 = 0; r1 < nx; ++r1) {
        size_t ir = r1 + nx * (r2 + (size_t)ny * r3);
        double ph = 2.0 * RO_PI * ((double)r1 * d1 / nx + (double)r2 * d2 / ny + (double)r3 * d3 / nz);
        ro_z pf = cexp(-I * ph);
        ro_z a = fround(uk[ir]);
        ro_z b = fround(ukq[ir]) * pf;
        uk[ir] = conj(b) * a;
      }
  dft3_inplace(nx, ny, nz, -1, uk);
  for (int ig = 0; ig < ng; ++ig) {
    size_t ir = wrapi(-lg[3 * ig], nx) + nx * (wrapi(-lg[3 * ig + 1], ny) + (size_t)ny * wrapi(-lg[3 * ig + 2], nz));
    m[ig] = fround(uk[ir]);
  }
}

void ro_pairdens(int ng_k, const int32_t *kg_k, const ro_z *c_k, int ng_kq,
                 const int32_t *kg_kq, const ro_z *c_kq, int nwx, int nwy, int nwz,
                 int ng_eps, const int32_t *lg, const int32_t *shift_g, ro_z *m) {
  ro_z *work = malloc(sizeof(ro_z) * 2 * (size_t)nwx * nwy * nwz);
  pairdens_fft(ng_k, kg_k, c_k, ng_kq, kg_kq, c_kq, nwx, nwy, nwz, ng_eps, lg, shift_g, m, work);
  free(work);
}

void ro_vmab(int ng, const int32_t *kg, const ro_z *c_k, const ro_z *c_kq,
             const double *b1, const double *b2, const double *b3, ro_z *vm) {
  /* chiqw_sub.F:207-232 */
  ro_z s[3] = {0, 0, 0};
  for (int ig = 0; ig < ng; ++ig) {
    for (int d = 0; d < 3; ++d) {
      double gv = (double)kg[3 * ig] * b1[d] + (double)kg[3 * ig + 1] * b2[d] + (double)kg[3 * ig + 2] * b3[d];
      s[d] += gv * conj(c_kq[ig]) * c_k[ig];
    }
  }
  for (int d = 0; d < 3; ++d) vm[d] = s[d] * I;
}

void ro_vmaa(int ng, const int32_t *kg, const ro_z *c_k, const ro_z *c_kq,
             const double *b1, const double *b2, const double *b3, const double *skt,
             ro_z *vm) {
  /* chiqw_sub.F:234-262 */
  ro_z s[3] = {0, 0, 0};
  for (int ig = 0; ig < ng; ++ig) {
    for (int d = 0; d < 3; ++d) {
      double gv = (skt[0] + kg[3 * ig]) * b1[d] + (skt[1] + kg[3 * ig + 1]) * b2[d] +
                  (skt[2] + kg[3 * ig + 2]) * b3[d];
      s[d] += gv * conj(c_kq[ig]) * c_k[ig];
    }
  }
  for (int d = 0; d < 3; ++d) vm[d] = s[d] * I;
}

/* --------------------------------------------------------------- Dyson */

static inline double cabs1(ro_z z) { return fabs(creal(z)) + fabs(cimag(z)); }

int ro_zinv(int n, ro_z *a) {
  /* LU with partial pivoting by |Re|+|Im| (zgetrf's izamax), then the inverse
   * by triangular solves against the identity. */
  int *piv = malloc(sizeof(int) * n);
  int info = 0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    double best = cabs1(a[j + (size_t)n * j]);
    for (int i = j + 1; i < n; ++i) {
      double v = cabs1(a[i + (size_t)n * j]);
      if (v > best) { best = v; p = i; }
    }
    piv[j] = p;
    if (a[p + (size_t)n * j] == 0.0) { if (!info) info = j + 1; continue; }
    if (p != j)
      for (int k = 0; k < n; ++k) {
        ro_z t = a[j + (size_t)n * k]; a[j + (size_t)n * k] = a[p + (size_t)n * k]; a[p + (size_t)n * k] = t;
      }
    ro_z r = 1.0 / a[j + (size_t)n * j];
    for (int i = j + 1; i < n; ++i) a[i + (size_t)n * j] *= r;
    for (int k = j + 1; k < n; ++k) {
      ro_z t = a[j + (size_t)n * k];
      if (t == 0.0) continue;
      for (int i = j + 1; i < n; ++i) a[i + (size_t)n * k] -= a[i + (size_t)n * j] * t;
    }
  }
  if (info) { free(piv); return info; }
  ro_z *x = malloc(sizeof(ro_z) * (size_t)n * n);
  for (int c = 0; c < n; ++c) {
    ro_z *col = x + (size_t)n * c;
    for (int i = 0; i < n; ++i) col[i] = 0.0;
    col[c] = 1.0;
    for (int j = 0; j < n; ++j) if (piv[j] != j) { ro_z t = col[j]; col[j] = col[piv[j]]; col[piv[j]] = t; }
    for (int j = 0; j < n; ++j) { ro_z t = col[j]; if (t != 0.0) for (int i = j + 1; i < n; ++i) col[i] -= a[i + (size_t)n * j] * t; }
    for (int j = n - 1; j >= 0; --j) { col[j] /= a[j + (size_t)n * j]; ro_z t = col[j]; for (int i = 0; i < j; ++i) col[i] -= a[i + (size_t)n * j] * t; }
  }
  memcpy(a, x, sizeof(ro_z) * (size_t)n * n);
  free(x);
  free(piv);
  return 0;
}

static double qg_norm(const int32_t *g, const double *q, const double *b1,
                      const double *b2, const double *b3) {
  double v[3];
  for (int d = 0; d < 3; ++d)
    v[d] = (q[0] + (double)g[0]) * b1[d] + (q[1] + (double)g[1]) * b2[d] + (q[2] + (double)g[2]) * b3[d];
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

int ro_eps_rpa(int ng, const int32_t *lg, const double *q, const double *b1,
               const double *b2, const double *b3, const ro_z *chi, ro_z *eps) {
  /* chiqw_sub.F:386-457 */
  const double tpi = 2.0 * acos(-1.0);
  double *qn = malloc(sizeof(double) * ng);
  for (int i = 0; i < ng; ++i) qn[i] = qg_norm(lg + 3 * i, q, b1, b2, b3);
  for (int j = 0; j < ng; ++j)
    for (int i = 0; i < ng; ++i) {
      ro_z c = 2.0 * chi[i + (size_t)ng * j];
      ro_z v = c * 2.0 * tpi / qn[i] / qn[j];
      eps[i + (size_t)ng * j] = (i == j) ? 1.0 - v : -v;
    }
  free(qn);
  return ro_zinv(ng, eps);
}

int ro_eps_rpa_q0(int ng, const int32_t *lg, const double *b1, const double *b2,
                  const double *b3, int g0, double wd, double delt, double wr,
                  const ro_z *chi, ro_z *eps) {
  /* chiqw_sub.F:459-568 */
  const double tpi = 2.0 * acos(-1.0);
  const double q[3] = {0, 0, 0};
  double w = wr;
  if (w == 0.0) w = 1.0e-5;
  double *qn = malloc(sizeof(double) * ng);
  for (int i = 0; i < ng; ++i) qn[i] = qg_norm(lg + 3 * i, q, b1, b2, b3);
  /* CMPLX without KIND: the Drude term is rounded to single precision */
  ro_z drude = (double)(float)(-wd * wd / (w * w + delt * delt)) +
               I * (double)(float)(delt * wd * wd / (w * w * w + w * delt * delt));
  for (int j = 0; j < ng; ++j)
    for (int i = 0; i < ng; ++i) {
      ro_z c = 2.0 * chi[i + (size_t)ng * j];
      ro_z v;
      if (i == j && i == g0) v = 1.0 - c * 2.0 * tpi + drude;
      else if (i == j) v = 1.0 - c * 2.0 * tpi / qn[i] / qn[j];
      else if (i == g0) v = -c * 2.0 * tpi / qn[j];
      else if (j == g0) v = -c * 2.0 * tpi / qn[i];
      else v = -c * 2.0 * tpi / qn[i] / qn[j];
      eps[i + (size_t)ng * j] = v;
    }
  free(qn);
  return ro_zinv(ng, eps);
}

/* ------------------------------------------------------------- chi(q,w) */

/* Tetrahedron weights for one pair (chiqw.F:1350-1395): energies to the 3D k
 * grid (rounded to single precision as the reference's CMPLX does), dsgn = +1
 * then -1 into one accumulator, back to k order: x[ik*nen + ie]. */
static void pair_weights(int ntk, int n1, int n2, int n3, const int32_t *imt,
                         const int32_t *index_kpt, const double *eo, const double *ev,
                         double ef, double delt, double dmna, double dmnr, int nen,
                         const ro_z *em, ro_z *e1g, ro_z *e2g, ro_z *xg, ro_z *x) {
  int nk3 = n1 * n2 * n3;
  for (int g = 0; g < nk3; ++g) {
    int ik = index_kpt[g];
    e1g[g] = (double)(float)eo[ik];
    e2g[g] = (double)(float)ev[ik];
  }
  memset(xg, 0, sizeof(ro_z) * (size_t)nen * nk3);
  ro_ttrhdrn_chi0(dmna, dmnr, n1, n2, n3, imt, e1g, e2g, ef, ef, delt, 1.0, nen, em, xg);
  ro_ttrhdrn_chi0(dmna, dmnr, n1, n2, n3, imt, e1g, e2g, ef, ef, delt, -1.0, nen, em, xg);
  (void)ntk;
  for (int g = 0; g < nk3; ++g) {
    int ik = index_kpt[g];
    for (int ie = 0; ie < nen; ++ie) x[(size_t)ik * nen + ie] = xg[(size_t)g * nen + ie];
  }
}

int ro_chiqw_q(int ntk, int ntg, const int32_t *ng0, const int32_t *kg0,
               int nocc, const ro_z *c_occ, int nvir, const ro_z *c_vir,
               const double *e_occ, const double *e_vir, const ro_z *p_occ,
               const ro_z *p_vir, int npair, const int32_t *pair_occ,
               const int32_t *pair_vir, const int32_t *ikq, const int32_t *shift,
               int nwx, int nwy, int nwz, int ng, const int32_t *lg, int nkb1,
               int nkb2, int nkb3, const int32_t *imt, const int32_t *index_kpt,
               double ef, double delt, double dmna, double dmnr, int nen,
               const ro_z *em, double volume, ro_z *chi, ro_z *m_out, ro_z *x_out,
               int nthreads) {
  (void)nocc; (void)nvir;
  size_t nchi = (size_t)ng * ng * nen;
  memset(chi, 0, sizeof(ro_z) * nchi);
  size_t nr = (size_t)nwx * nwy * nwz;
  int nk3 = nkb1 * nkb2 * nkb3;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    ro_z *pc = calloc(nchi, sizeof(ro_z));
    ro_z *ism = malloc(sizeof(ro_z) * (size_t)ntk * ng); /* ism(ik, ig): ik fastest */
    ro_z *mt = malloc(sizeof(ro_z) * ng);
    ro_z *work = malloc(sizeof(ro_z) * 2 * nr);
    ro_z *x = malloc(sizeof(ro_z) * (size_t)ntk * nen);
    ro_z *xg = malloc(sizeof(ro_z) * (size_t)nk3 * nen);
    ro_z *e1g = malloc(sizeof(ro_z) * nk3), *e2g = malloc(sizeof(ro_z) * nk3);
    double *eo = malloc(sizeof(double) * ntk), *ev = malloc(sizeof(double) * ntk);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int ij = 0; ij < npair; ++ij) {
      int jb = pair_occ[ij], ib = pair_vir[ij];
      for (int ik = 0; ik < ntk; ++ik) {
        int kq = ikq[ik];
        pairdens_fft(ng0[ik], kg0 + (size_t)3 * ntg * ik, c_occ + ((size_t)jb * ntk + ik) * ntg,
                     ng0[kq], kg0 + (size_t)3 * ntg * kq, c_vir + ((size_t)ib * ntk + kq) * ntg,
                     nwx, nwy, nwz, ng, lg, shift + 3 * ik, mt, work);
        for (int ig = 0; ig < ng; ++ig) ism[ik + (size_t)ntk * ig] = mt[ig];
        eo[ik] = e_occ[(size_t)jb * ntk + ik];
        ev[ik] = e_vir[(size_t)ib * ntk + kq];
      }
      pair_weights(ntk, nkb1, nkb2, nkb3, imt, index_kpt, eo, ev, ef, delt, dmna, dmnr,
                   nen, em, e1g, e2g, xg, x);
      if (ij == 0 && m_out)
        for (int ik = 0; ik < ntk; ++ik)
          for (int ig = 0; ig < ng; ++ig) m_out[(size_t)ik * ng + ig] = ism[ik + (size_t)ntk * ig];
      if (ij == 0 && x_out) memcpy(x_out, x, sizeof(ro_z) * (size_t)ntk * nen);
      /* chiqw.F:1397-1410 */
      for (int ie = 0; ie < nen; ++ie)
        for (int jg = 0; jg < ng; ++jg)
          for (int ig = 0; ig < ng; ++ig) {
            ro_z acc = pc[((size_t)ie * ng + jg) * ng + ig];
            const ro_z *ai = ism + (size_t)ntk * ig, *aj = ism + (size_t)ntk * jg;
            for (int ik = 0; ik < ntk; ++ik) {
              int kq = ikq[ik];
              acc = acc + x[(size_t)ik * nen + ie] * conj(ai[ik]) * aj[ik] / (double)ntk *
                              (1.0 - p_occ[(size_t)jb * ntk + ik] * p_vir[(size_t)ib * ntk + kq]);
            }
            pc[((size_t)ie * ng + jg) * ng + ig] = acc;
          }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (size_t i = 0; i < nchi; ++i) chi[i] += pc[i];
    free(pc); free(ism); free(mt); free(work); free(x); free(xg); free(e1g); free(e2g);
    free(eo); free(ev);
  }
  for (size_t i = 0; i < nchi; ++i) chi[i] /= volume;
  return 0;
}

int ro_chiqw_q0(int ntk, int ntg, const int32_t *ng0, const int32_t *kg0,
                int nocc, const ro_z *c_occ, int nvir, const ro_z *c_vir,
                const double *e_occ, const double *e_vir, const ro_z *p_occ,
                const ro_z *p_vir, int npair, const int32_t *pair_occ,
                const int32_t *pair_vir, int nwx, int nwy, int nwz, int ng,
                const int32_t *lg, int g0, const double *b1, const double *b2,
                const double *b3, int nkb1, int nkb2, int nkb3, const int32_t *imt,
                const int32_t *index_kpt, double ef, double delt, double dmna,
                double dmnr, int nen, const ro_z *em, double volume, ro_z *chi3,
                int nthreads) {
  (void)nocc; (void)nvir;
  size_t nchi = (size_t)ng * ng * nen;
  memset(chi3, 0, sizeof(ro_z) * 3 * nchi);
  size_t nr = (size_t)nwx * nwy * nwz;
  int nk3 = nkb1 * nkb2 * nkb3;
  const int32_t zero[3] = {0, 0, 0};
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    ro_z *pc = calloc(3 * nchi, sizeof(ro_z));
    ro_z *ism = malloc(sizeof(ro_z) * (size_t)ntk * ng * 3);
    ro_z *mt = malloc(sizeof(ro_z) * ng);
    ro_z *work = malloc(sizeof(ro_z) * 2 * nr);
    ro_z *x = malloc(sizeof(ro_z) * (size_t)ntk * nen);
    ro_z *xg = malloc(sizeof(ro_z) * (size_t)nk3 * nen);
    ro_z *e1g = malloc(sizeof(ro_z) * nk3), *e2g = malloc(sizeof(ro_z) * nk3);
    double *eo = malloc(sizeof(double) * ntk), *ev = malloc(sizeof(double) * ntk);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int ij = 0; ij < npair; ++ij) {
      int jb = pair_occ[ij], ib = pair_vir[ij];
      for (int ik = 0; ik < ntk; ++ik) {
        const ro_z *ck = c_occ + ((size_t)jb * ntk + ik) * ntg;
        const ro_z *ckq = c_vir + ((size_t)ib * ntk + ik) * ntg;
        const int32_t *kg = kg0 + (size_t)3 * ntg * ik;
        pairdens_fft(ng0[ik], kg, ck, ng0[ik], kg, ckq, nwx, nwy, nwz, ng, lg, zero, mt, work);
        ro_z vm[3];
        ro_vmab(ng0[ik], kg, ck, ckq, b1, b2, b3, vm);
        /* chiqw.F:1787-1805 */
        double de = e_vir[(size_t)ib * ntk + ik] - e_occ[(size_t)jb * ntk + ik];
        if (fabs(de) < delt) de = delt;
        for (int ix = 0; ix < 3; ++ix) {
          vm[ix] = -I * (vm[ix] - 0.0) / de;
          for (int ig = 0; ig < ng; ++ig) ism[ik + (size_t)ntk * (ig + (size_t)ng * ix)] = mt[ig];
          ism[ik + (size_t)ntk * (g0 + (size_t)ng * ix)] = fabs(de) <= delt ? 0.0 : vm[ix];
        }
        eo[ik] = e_occ[(size_t)jb * ntk + ik];
        ev[ik] = e_vir[(size_t)ib * ntk + ik];
      }
      pair_weights(ntk, nkb1, nkb2, nkb3, imt, index_kpt, eo, ev, ef, delt, dmna, dmnr,
                   nen, em, e1g, e2g, xg, x);
      for (int ix = 0; ix < 3; ++ix)
        for (int ie = 0; ie < nen; ++ie)
          for (int jg = 0; jg < ng; ++jg)
            for (int ig = 0; ig < ng; ++ig) {
              size_t o = (((size_t)ix * nen + ie) * ng + jg) * ng + ig;
              ro_z acc = pc[o];
              const ro_z *ai = ism + (size_t)ntk * (ig + (size_t)ng * ix);
              const ro_z *aj = ism + (size_t)ntk * (jg + (size_t)ng * ix);
              for (int ik = 0; ik < ntk; ++ik)
                acc = acc + x[(size_t)ik * nen + ie] * conj(ai[ik]) * aj[ik] / (double)ntk *
                                (1.0 - p_occ[(size_t)jb * ntk + ik] * p_vir[(size_t)ib * ntk + ik]);
              pc[o] = acc;
            }
    }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (size_t i = 0; i < 3 * nchi; ++i) chi3[i] += pc[i];
    free(pc); free(ism); free(mt); free(work); free(x); free(xg); free(e1g); free(e2g);
    free(eo); free(ev);
  }
  for (size_t i = 0; i < 3 * nchi; ++i) chi3[i] /= volume;
  return 0;
}

/* ======================================================================== */
/* calc_int, calc_w3d                                                         */
/* ======================================================================== */

int ro_make_lg0(int ntg, const double *b1, const double *b2, const double *b3,
                double gcut_eps, double gcut_psi, double q1, double q2, double q3,
                int32_t *lg, int32_t *ng_eps) { /* calc_int_sub.F:447-504 */
  int n = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i1 = -100; i1 <= 100; ++i1)
      for (int i2 = -100; i2 <= 100; ++i2)
        for (int i3 = -100; i3 <= 100; ++i3) {
          double v[3];
          for (int d = 0; d < 3; ++d)
            v[d] = (q1 + i1) * b1[d] + (q2 + i2) * b2[d] + (q3 + i3) * b3[d];
          double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
          int take = pass == 0 ? (s <= gcut_eps) : (s > gcut_eps && s <= gcut_psi);
          if (!take) continue;
          if (n < ntg) { lg[3 * n] = i1; lg[3 * n + 1] = i2; lg[3 * n + 2] = i3; }
          ++n;
        }
    if (pass == 0) *ng_eps = n;
  }
  return n;
}

void ro_make_eps(int ntgq, int ne, int itrs, int ng, const int32_t *lg,
                 const int32_t *rw, const double *rginv, const int32_t *pg, int nnp,
                 int l1, int l2, int l3, const int32_t *pack, const float _Complex *epsirr,
                 float _Complex *epsmk) { /* calc_int_sub.F:222-311 */
  const double tpi = 2.0 * acos(-1.0);
  const size_t plane = (size_t)ntgq * ntgq;
  for (size_t i = 0; i < plane * ne; ++i) epsmk[i] = 0;
  int *map = malloc(sizeof(int) * (ng > 0 ? ng : 1));
  ro_z *pf = malloc(sizeof(ro_z) * (ng > 0 ? ng : 1));
  for (int ig = 0; ig < ng; ++ig) {
    int i1 = itrs * lg[3 * ig], j1 = itrs * lg[3 * ig + 1], k1 = itrs * lg[3 * ig + 2];
    int i2 = i1 + rw[0], j2 = j1 + rw[1], k2 = k1 + rw[2];
    int i3 = (int)rginv[0] * i2 + (int)rginv[3] * j2 + (int)rginv[6] * k2;
    int j3 = (int)rginv[1] * i2 + (int)rginv[4] * j2 + (int)rginv[7] * k2;
    int k3 = (int)rginv[2] * i2 + (int)rginv[5] * j2 + (int)rginv[8] * k2;
    map[ig] = pack[(i3 + l1) + (2 * l1 + 1) * ((j3 + l2) + (size_t)(2 * l2 + 1) * (k3 + l3))] - 1;
    double ph = tpi * ((double)i1 * pg[0] + (double)j1 * pg[1] + (double)k1 * pg[2]);
    pf[ig] = cexp(-I * ph / (double)nnp); /* pf1; pf2 is its conjugate */
  }
  for (int ig = 0; ig < ng; ++ig)
    for (int jg = 0; jg < ng; ++jg) {
      const int a = itrs == 1 ? map[ig] : map[jg], b = itrs == 1 ? map[jg] : map[ig];
      for (int ie = 0; ie < ne; ++ie) {
        const ro_z e = (ro_z)epsirr[(size_t)ie * plane + a + (size_t)ntgq * b];
        /* epsirr*pf1*pf2, left to right, stored in single precision */
        epsmk[(size_t)ie * plane + ig + (size_t)ntgq * jg] = (float _Complex)((e * pf[ig]) * conj(pf[jg]));
      }
    }
  free(map);
  free(pf);
}

void ro_w3d_func(int nwf, int ng_psi, int ldr, const ro_z *rho, ro_z *func) {
  /* calc_intW.F:833-847 */
  for (int iw = 0; iw < nwf; ++iw)
    for (int jw = 0; jw < nwf; ++jw) {
      ro_z s = 0;
      for (int g = 0; g < ng_psi; ++g) s += rho[(size_t)iw * ldr + g] * conj(rho[(size_t)jw * ldr + g]);
      func[jw * nwf + iw] = s;
    }
}

void ro_w3d_funcw(int nwf, int ng_eps, int ng_psi, int ldr, const ro_z *rho, int ntgq,
                  int ne, const float _Complex *epsmk, ro_z *funcw) { /* calc_intW.F:848-912 */
  const size_t plane = (size_t)ntgq * ntgq;
  /* threads over the frequencies, as the reference's OMP DO: every element is
   * summed by one thread in the reference's order */
#pragma omp parallel for schedule(dynamic)
  for (int ie = 0; ie < ne; ++ie)
    for (int iw = 0; iw < nwf; ++iw)
      for (int jw = 0; jw < nwf; ++jw) {
        ro_z s = 0;
        for (int ig = 0; ig < ng_eps; ++ig)
          for (int jg = 0; jg < ng_eps; ++jg)
            s += rho[(size_t)iw * ldr + ig] * (ro_z)epsmk[(size_t)ie * plane + ig + (size_t)ntgq * jg] *
                 conj(rho[(size_t)jw * ldr + jg]);
        for (int g = ng_eps; g < ng_psi; ++g)
          s += rho[(size_t)iw * ldr + g] * conj(rho[(size_t)jw * ldr + g]);
        funcw[((size_t)ie * nwf + jw) * nwf + iw] = s;
      }
}

void ro_w3d_vw(int nwf, int ntq, const double *sq, const ro_z *func, const ro_z *funcw,
               int ne, int na1, int na2, int na3, double volume, double chead,
               const ro_z *cheadw, ro_z *vmat, ro_z *wmat) { /* calc_intW.F:942-985 */
  const double tpi = 2.0 * acos(-1.0);
  const int n1 = 2 * na1 + 1, n2 = 2 * na2 + 1, n3 = 2 * na3 + 1;
  const size_t nr = (size_t)n1 * n2 * n3;
  for (int ia1 = -na1; ia1 <= na1; ++ia1)
    for (int ia2 = -na2; ia2 <= na2; ++ia2)
      for (int ia3 = -na3; ia3 <= na3; ++ia3) {
        const size_t r = (size_t)(ia1 + na1) + n1 * ((size_t)(ia2 + na2) + (size_t)n2 * (ia3 + na3));
        for (int iw = 0; iw < nwf; ++iw)
          for (int jw = 0; jw < nwf; ++jw) {
            ro_z s = 0;
            for (int q = 0; q < ntq; ++q) {
              double ph = tpi * (sq[3 * q] * ia1 + sq[3 * q + 1] * ia2 + sq[3 * q + 2] * ia3);
              s += func[((size_t)q * nwf + jw) * nwf + iw] * cexp(-I * ph);
            }
            vmat[(r * nwf + jw) * nwf + iw] = 2.0 * tpi * s / volume / (double)ntq + chead;
            for (int ie = 0; ie < ne; ++ie) {
              ro_z t = 0;
              for (int q = 0; q < ntq; ++q) {
                double ph = tpi * (sq[3 * q] * ia1 + sq[3 * q + 1] * ia2 + sq[3 * q + 2] * ia3);
                t += funcw[(((size_t)q * ne + ie) * nwf + jw) * nwf + iw] * cexp(-I * ph);
              }
              wmat[(((size_t)ie * nr + r) * nwf + jw) * nwf + iw] =
                  2.0 * tpi * t / volume / (double)ntq + cheadw[ie];
            }
          }
      }
}
