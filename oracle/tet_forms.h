/*
 * tet_forms.h -- generalized-tetrahedron vertex weights of ttrhdrn_chi0
 * (m_tetrahedron_20170325.F:99-503).
 *
 * TEST INFRASTRUCTURE: part of the oracle, never linked by the product.
 *
 * The closed forms are written once here and instantiated twice: in double by
 * oracle.c (the restatement) and in __float128 by tet_quad.c, which evaluates
 * the reference's formulas with 113-bit arithmetic to arbitrate between the
 * reference and the oracle where the double closed forms cancel.  Before
 * inclusion define
 *   TF_Z       the complex type of the closed forms and of the accumulators,
 *   TF_LOG(z)  its principal complex logarithm,
 *   TF(name)   the name given to an instantiated function.
 * Classification (chk1), the sub-tetrahedron cuts (ttct), the merged nodes of
 * a degenerate class and the series below are double in both instantiations:
 * the reference forms them in double, and they decide which formula applies.
 *
 * Every closed form is the exact integral
 *   W_v = int_T lambda_v / (e + sum_j lambda_j g~_j)   (T the unit simplex, |T| = 1/6)
 * for a node multiset g~: the vertex values, with each class of nearly equal
 * values (chk1) replaced by its mean.  Where |g~_j - c| is small against
 * |s| = |e + c| (c the mean node), the closed forms cancel: their terms are
 * ~24 r^-3 times W, r = max_j |g~_j - c| / |s|, and the rounding of a 1-ulp
 * logarithm is amplified by as much (1e-5 relative at 98 eV on the Fermi-
 * surface pieces of a metal, in the reference and in a faithful restatement
 * alike).  There the oracle evaluates the same integral by its series
 *   W_v = (1/s) sum_n (-1/s)^n n!/(n+4)! P_{v,n},
 *   P_{v,n} = [z^n] prod_j 1/(1 - d_j z) * 1/(1 - d_v z),   d_j = g~_j - c,
 * from the Dirichlet moments int_T lambda^a = a!/(|a|+3)!; term n is at most
 * r^n / (24 |s|), so for r < TET_SERIES_R the sum converges geometrically and
 * loses nothing.  At r >= TET_SERIES_R the closed forms lose at most ~24/r^3
 * (3e3 ulps) and are used as the reference writes them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef TET_FORMS_COMMON
#define TET_FORMS_COMMON

#define TET_SERIES_R 0.5
#define TT_NMAX 99

/* chk1 (m_tetrahedron_20170325.F:457-503): degeneracy classes of 4 vertices.
 * Vertices are compared in the reference's rotated order p = mod(i,4)+1. */
static void tet_chk1(double dmna, double dmnr, const ro_z *g, int *ia2, int *ia3,
                     int *ia4, int *i3out, int *n2) {
  int i3 = 0, i4 = 0;
  for (int v = 0; v < 4; ++v) { ia2[v] = 0; ia3[v] = -1; ia4[v] = -1; }
  for (int i1 = 1; i1 <= 3; ++i1)
    for (int i2 = i1 + 1; i2 <= 4; ++i2) {
      int p = i1 % 4, q = i2 % 4; /* 0-based form of mod(i,4)+1 */
      double d1 = cabs(g[p] - g[q]);
      double d2 = cabs(g[p] + g[q]) * 0.5;
      if (d1 < dmna || d1 / d2 < dmnr) {
        ++i3;
        if (!ia2[p]) { ia3[i4++] = p; ia2[p] = 1; }
        if (!ia2[q]) { ia3[i4++] = q; ia2[q] = 1; }
      }
    }
  int i5 = 0;
  for (int i1 = 1; i1 <= 4; ++i1) {
    int p = i1 % 4;
    if (!ia2[p]) ia4[i5++] = p;
  }
  *i3out = i3;
  *n2 = i4;
}

/* The class of a tetrahedron and its merged nodes: gt[v] is the node of
 * vertex v in the integral the reference's formula of that class evaluates
 * (calc, m_tetrahedron_20170325.F:231-300: the class means geq1/geq2, summed
 * in the reference's order).  Returns the class (0 normal, 2 pair, 3 triple,
 * 4 two pairs, 5 all equal) or -1 for a class the reference rejects. */
static int tet_nodes(double dmna, double dmnr, const ro_z *g, int *a3, int *a4, ro_z *gt) {
  int ia2[4], i3, n2;
  tet_chk1(dmna, dmnr, g, ia2, a3, a4, &i3, &n2);
  for (int v = 0; v < 4; ++v) gt[v] = g[v];
  switch (n2) {
  case 0:
    return 0;
  case 2: {
    ro_z q1 = (g[a3[0]] + g[a3[1]]) / 2.0;
    gt[a3[0]] = gt[a3[1]] = q1;
    return 2;
  }
  case 3: {
    ro_z q1 = (g[a3[0]] + g[a3[1]] + g[a3[2]]) / 3.0;
    gt[a3[0]] = gt[a3[1]] = gt[a3[2]] = q1;
    return 3;
  }
  case 4:
    if (i3 == 2) {
      ro_z q1 = (g[a3[0]] + g[a3[1]]) / 2.0, q2 = (g[a3[2]] + g[a3[3]]) / 2.0;
      gt[a3[0]] = gt[a3[1]] = q1;
      gt[a3[2]] = gt[a3[3]] = q2;
      return 4;
    }
    if (i3 >= 3 && i3 <= 6) {
      ro_z q1 = (g[a3[0]] + g[a3[1]] + g[a3[2]] + g[a3[3]]) / 4.0;
      for (int v = 0; v < 4; ++v) gt[v] = q1;
      return 5;
    }
    return -1;
  default:
    return -1;
  }
}

/* W_v (v = 0..3) of the node set gt at frequency e by the series of the file
 * header; returns 0 (W untouched) when r >= TET_SERIES_R.  The nodes share
 * one imaginary part (the broadening; a class mean may differ from it by an
 * ulp), so d_j is taken real and the common imaginary part goes into s. */
static int tet_series(const ro_z *gt, ro_z e, ro_z *W) {
  const ro_z c = (gt[0] + gt[1] + gt[2] + gt[3]) * 0.25;
  const ro_z s = e + c;
  double d[4], dm = 0.0;
  for (int j = 0; j < 4; ++j) {
    d[j] = creal(gt[j]) - creal(c);
    if (fabs(d[j]) > dm) dm = fabs(d[j]);
  }
  const double as = cabs(s);
  if (!(as > 0.0) || !(dm < TET_SERIES_R * as)) return 0;
  const double r = dm / as;
  /* terms until r^n / 24 < 2^-60 of the leading term */
  int nmax = r > 0.0 ? (int)ceil(-60.0 * log(2.0) / log(r)) : 1;
  if (nmax < 1) nmax = 1;
  const ro_z t = -1.0 / s;
  ro_z tn = 1.0 / s; /* (1/s) (-1/s)^n */
  double h1 = 1.0, h2 = 1.0, h3 = 1.0, h = 1.0; /* partial products at order n */
  double p[4] = {1.0, 1.0, 1.0, 1.0};
  ro_z w[4] = {0.0, 0.0, 0.0, 0.0};
  for (int n = 0; n <= nmax; ++n) {
    const double cf = 1.0 / ((double)(n + 1) * (n + 2) * (n + 3) * (n + 4));
    for (int v = 0; v < 4; ++v) w[v] += (cf * p[v]) * tn;
    tn *= t;
    h1 *= d[0];
    h2 = h1 + d[1] * h2;
    h3 = h2 + d[2] * h3;
    h = h3 + d[3] * h;
    for (int v = 0; v < 4; ++v) p[v] = h + d[v] * p[v];
  }
  for (int v = 0; v < 4; ++v) W[v] = w[v];
  return 1;
}

/* ttct (m_tetrahedron_20170325.F:143-207): split sub-tetrahedra by the plane
 * dvtx = dval along edges, then keep those on the imod side.
 * dwgt[t][v][j]: weight of original corner j in vertex v of sub-tetrahedron t
 * (the reference's dwgt(j, v, t)). */
static int tet_ttct(double dval, const double *dvtx, int imod, int *ntet,
                    double dwgt[][4][4], double *dvol) {
  unsigned char bvtx[TT_NMAX][4], btsk[TT_NMAX];
  memset(bvtx, 0, sizeof bvtx);
  memset(btsk, 0, sizeof btsk);
  int itet = 0;
  while (itet < *ntet) {
    ++itet;
    int t = itet - 1;
    double v0[4];
    unsigned char b0[4];
    for (int i = 0; i < 4; ++i) {
      v0[i] = dwgt[t][i][0] * dvtx[0] + dwgt[t][i][1] * dvtx[1] +
              dwgt[t][i][2] * dvtx[2] + dwgt[t][i][3] * dvtx[3];
      b0[i] = bvtx[t][i];
    }
    int b1 = 0, i1 = 0, i2 = 0;
    for (i1 = 0; i1 < 3 && !b1; ++i1) {
      if (b0[i1]) continue;
      for (i2 = i1 + 1; i2 < 4; ++i2) {
        if (b0[i2]) continue;
        b1 = (v0[i1] - dval) * (dval - v0[i2]) > 0.0;
        if (b1) break;
      }
      if (b1) break;
    }
    if (b1) {
      double d1 = (v0[i2] - dval) / (v0[i2] - v0[i1]);
      double d2 = (dval - v0[i1]) / (v0[i2] - v0[i1]);
      if (*ntet + 1 > TT_NMAX) return -1; /* "ntet.gt.nmax" */
      int n = (*ntet)++;
      memcpy(bvtx[n], bvtx[t], 4);
      memcpy(dwgt[n], dwgt[t], sizeof dwgt[0]);
      bvtx[n][i1] = 1;
      for (int j = 0; j < 4; ++j) dwgt[n][i1][j] = dwgt[t][i1][j] * d1 + dwgt[t][i2][j] * d2;
      dvol[n] = dvol[t] * d1;
      bvtx[t][i2] = 1;
      for (int j = 0; j < 4; ++j) dwgt[t][i2][j] = dwgt[t][i1][j] * d1 + dwgt[t][i2][j] * d2;
      dvol[t] = dvol[t] * d2;
      --itet;
    } else {
      for (int i = 0; i < 4; ++i) {
        if (b0[i]) continue;
        btsk[t] = btsk[t] || ((double)imod * (v0[i] - dval) > 0.0);
      }
    }
  }
  int k = 0;
  for (int i = 0; i < *ntet; ++i) {
    if (!btsk[i]) continue;
    if (k != i) { memcpy(dwgt[k], dwgt[i], sizeof dwgt[0]); dvol[k] = dvol[i]; }
    ++k;
  }
  *ntet = k;
  return 0;
}

#endif /* TET_FORMS_COMMON */

/* Closed forms of the generalized tetrahedron method (Fujiwara, Yamamoto,
 * Ishii, J. Phys. Soc. Jpn. 72, 777 (2003)) exactly as evaluated at
 * m_tetrahedron_20170325.F:304-451, for one frequency e. */
static TF_Z TF(sub1g30g21)(TF_Z g0, TF_Z g1, TF_Z e) {
  TF_Z d = g0 - g1;
  TF_Z x = (0.5 * (g0 + g1) + e) / (d * d);
  x = x - TF_LOG((g0 + e) / (g1 + e)) * (g0 + e) * (g1 + e) / (d * d * d);
  return x;
}
static TF_Z TF(sub1g30g20)(TF_Z g0, TF_Z g1, TF_Z e) {
  TF_Z d = g0 - g1;
  TF_Z x = (g0 - 3.0 * g1 - 2.0 * e) * 0.25 / (d * d);
  x = x + (g1 + e) * (g1 + e) * TF_LOG((g0 + e) / (g1 + e)) * 0.5 / (d * d * d);
  return x;
}
static TF_Z TF(sub1g30)(TF_Z g0, TF_Z g1, TF_Z g2, TF_Z e) {
  TF_Z x = 0.5 / (g0 - g2);
  TF_Z p = (g0 - g2) * (g0 - g1);
  x = x - 0.5 / (p * p) * (g0 + e) *
              (-2.0 * (g1 + e) * (g2 + e) + (g0 + e) * (g1 + g2 + 2.0 * e)) *
              TF_LOG((g0 + e) / (g1 + e));
  x = x + 0.5 / ((g0 - g2) * (g0 - g2)) * ((g2 + e) * (g2 + e)) *
              TF_LOG((g1 + e) / (g2 + e)) / (g1 - g2);
  x = x + 0.5 / ((g0 - g2) * (g0 - g1)) * (g1 + e);
  return x;
}
static TF_Z TF(sub1normal)(TF_Z g0, TF_Z g1, TF_Z g2, TF_Z g3, TF_Z e) {
  TF_Z l1 = (g1 + e) * (g1 + e) * TF_LOG((g1 + e) / (g0 + e));
  TF_Z x = (-l1) / ((g0 - g2) * (g0 - g3) * (g0 - g1)) * 0.5;
  x = x - (l1 - (g2 + e) * (g2 + e) * TF_LOG((g2 + e) / (g0 + e))) /
              ((g0 - g2) * (g1 - g2) * (g2 - g3)) * 0.5;
  x = x + (l1 - (g3 + e) * (g3 + e) * TF_LOG((g3 + e) / (g0 + e))) /
              ((g1 - g3) * (g0 - g3) * (g2 - g3)) * 0.5;
  return x;
}
static TF_Z TF(sub2simple)(TF_Z g0, TF_Z e) { return 1.0 / ((g0 + e) * 24.0); }
static TF_Z TF(sub2g30g20)(TF_Z g0, TF_Z g1, TF_Z e) {
  TF_Z d = g0 - g1, d2 = d * d;
  TF_Z x = (d * (g0 + 2.0 * g1 + 3.0 * e) - 6.0 * (g0 + e) * (g1 + e)) / (12.0 * (d2 * d));
  x = x + (g0 + e) * ((g1 + e) * (g1 + e)) * TF_LOG((g0 + e) / (g1 + e)) / (d2 * d2) * 0.5;
  return x;
}
static TF_Z TF(sub2g30g10)(TF_Z g0, TF_Z g2, TF_Z e) {
  TF_Z a = g0 + e, c = g2 + e, d = g0 - g2, d2 = d * d;
  TF_Z x = (2.0 * (a * a) - 7.0 * a * c + 11.0 * (c * c)) / (36.0 * (d2 * d));
  x = x - (c * c * c) * TF_LOG(a / c) / (6.0 * (d2 * d2));
  return x;
}
static TF_Z TF(sub2g10)(TF_Z g0, TF_Z g2, TF_Z g3, TF_Z e) {
  TF_Z a = g0 + e, c2 = g2 + e, c3 = g3 + e;
  TF_Z x = a * ((a * a) + 5.0 * c2 * c3 - 3.0 * a * (g2 + g3 + 2.0 * e)) /
           (12.0 * ((g0 - g2) * (g0 - g2)) * ((g0 - g3) * (g0 - g3)));
  TF_Z d02 = g0 - g2, d03 = g0 - g3;
  x = x - (c2 * c2 * c2) * TF_LOG(c2 / a) / (6.0 * (d02 * d02 * d02) * (g2 - g3));
  x = x + (c3 * c3 * c3) * TF_LOG(c3 / a) / (6.0 * (d03 * d03 * d03) * (g2 - g3));
  return x;
}
static TF_Z TF(sub2g30)(TF_Z g0, TF_Z g1, TF_Z g2, TF_Z e) {
  TF_Z a = g0 + e, b = g1 + e, c = g2 + e;
  TF_Z d01 = g0 - g1, d02 = g0 - g2, d12 = g1 - g2;
  TF_Z x = ((a * a) * d12 + (b * b) * d02) / (6.0 * d02 * (d01 * d01) * d12);
  x = x - (a * a) * (2.0 * b * d02 + c * d01) * TF_LOG(a / c) /
              (6.0 * (d02 * d02) * (d01 * d01 * d01));
  x = x + (b * b) * (2.0 * a * d12 - c * d01) * TF_LOG(b / c) /
              (6.0 * (d01 * d01 * d01) * (d12 * d12));
  return x;
}
static TF_Z TF(sub2normal)(TF_Z g0, TF_Z g1, TF_Z g2, TF_Z g3, TF_Z e) {
  TF_Z a = g0 + e, b = g1 + e, c = g2 + e, d = g3 + e;
  TF_Z x = -(b * b) / (6.0 * (g0 - g1) * (g1 - g2) * (g1 - g3));
  x = x - (c * c * c) * TF_LOG(c / b) / (6.0 * (g0 - g2) * ((g1 - g2) * (g1 - g2)) * (g2 - g3));
  x = x + (d * d * d) * TF_LOG(d / b) / (6.0 * (g0 - g3) * ((g1 - g3) * (g1 - g3)) * (g2 - g3));
  x = x + (a * a * a) * TF_LOG(a / b) / (6.0 * ((g0 - g1) * (g0 - g1)) * (g0 - g2) * (g0 - g3));
  return x;
}

/* calc (m_tetrahedron_20170325.F:209-302) over corner slots c[4] of arrays
 * fk/gk; xo + ne*c[v] receives the corner contributions.  use_series: the
 * series where it converges (r < TET_SERIES_R), else the closed forms only. */
static int TF(calc)(double dmna, double dmnr, const int *c, const ro_z *fk, const ro_z *gk,
                    int ne, const ro_z *em, TF_Z *xo, int use_series) {
  ro_z gd[4], gt[4];
  TF_Z g[4], f[4];
  for (int v = 0; v < 4; ++v) {
    gd[v] = gk[c[v]];
    g[v] = (TF_Z)gk[c[v]];
    f[v] = (TF_Z)fk[c[v]];
  }
  int a3[4], a4[4];
  const int cls = tet_nodes(dmna, dmnr, gd, a3, a4, gt);
  if (cls < 0) return -1;
  for (int ie = 0; ie < ne; ++ie) {
    TF_Z e = (TF_Z)em[ie];
#define XO(v) xo[(size_t)ne * c[v] + ie]
    if (use_series) {
      ro_z W[4];
      if (tet_series(gt, em[ie], W)) {
        for (int v = 0; v < 4; ++v) XO(v) += f[v] * (TF_Z)W[v];
        continue;
      }
    }
    switch (cls) {
    case 0: {
      TF_Z c1 = TF(sub2normal)(g[a4[1]], g[a4[0]], g[a4[2]], g[a4[3]], e);
      TF_Z c2 = TF(sub2normal)(g[a4[0]], g[a4[1]], g[a4[2]], g[a4[3]], e);
      TF_Z c3 = TF(sub2normal)(g[a4[0]], g[a4[2]], g[a4[1]], g[a4[3]], e);
      TF_Z c4 = TF(sub1normal)(g[a4[0]], g[a4[3]], g[a4[1]], g[a4[2]], e);
      XO(a4[0]) += f[a4[0]] * c1;
      XO(a4[1]) += f[a4[1]] * c2;
      XO(a4[2]) += f[a4[2]] * c3;
      XO(a4[3]) += f[a4[3]] * (c4 - c3 - c2 - c1);
      break;
    }
    case 2: {
      TF_Z q1 = (TF_Z)gt[a3[0]];
      TF_Z c1 = TF(sub2g10)(q1, g[a4[0]], g[a4[1]], e);
      TF_Z c2 = TF(sub2g30)(q1, g[a4[0]], g[a4[1]], e);
      TF_Z c3 = TF(sub1g30)(q1, g[a4[1]], g[a4[0]], e);
      XO(a3[0]) += f[a3[0]] * c1;
      XO(a3[1]) += f[a3[1]] * c1;
      XO(a4[0]) += f[a4[0]] * c2;
      XO(a4[1]) += f[a4[1]] * (c3 - c2 - c1 - c1);
      break;
    }
    case 3: {
      TF_Z q1 = (TF_Z)gt[a3[0]];
      TF_Z c1 = TF(sub2g30g10)(q1, g[a4[0]], e);
      TF_Z c2 = TF(sub1g30g20)(q1, g[a4[0]], e);
      XO(a3[0]) += f[a3[0]] * c1;
      XO(a3[1]) += f[a3[1]] * c1;
      XO(a3[2]) += f[a3[2]] * c1;
      XO(a4[0]) += f[a4[0]] * (c2 - c1 - c1 - c1);
      break;
    }
    case 4: {
      TF_Z q1 = (TF_Z)gt[a3[0]], q2 = (TF_Z)gt[a3[2]];
      TF_Z c1 = TF(sub2g30g20)(q1, q2, e); /* sub2g30g21(q2,q1,..) */
      TF_Z c2 = TF(sub1g30g21)(q1, q2, e);
      XO(a3[0]) += f[a3[0]] * c1;
      XO(a3[1]) += f[a3[1]] * c1;
      XO(a3[2]) += f[a3[2]] * (c2 - c1 - c1) * 0.5;
      XO(a3[3]) += f[a3[3]] * (c2 - c1 - c1) * 0.5;
      break;
    }
    default: { /* 5: all four equal */
      TF_Z c1 = TF(sub2simple)((TF_Z)gt[a3[0]], e);
      XO(a3[0]) += c1 * f[a3[0]];
      XO(a3[1]) += c1 * f[a3[1]];
      XO(a3[2]) += c1 * f[a3[2]];
      XO(a3[3]) += c1 * f[a3[3]];
      break;
    }
    }
#undef XO
  }
  return 0;
}

/* ttrhdrn_chi0 (m_tetrahedron_20170325.F:99-141): xo[ie + ne*ik] accumulated. */
static void TF(chi0)(double dmna, double dmnr, int n1, int n2, int n3, const int32_t *idx,
                     const ro_z *e1, const ro_z *e2, double f1, double f2, double delt,
                     double dsgn, int ne, const ro_z *em, TF_Z *xo, int use_series) {
  int nt = n1 * n2 * n3 * 6;
  double (*dwgt)[4][4] = malloc(sizeof(double[TT_NMAX][4][4]));
  double dvol[TT_NMAX];
  TF_Z *xxo = malloc(sizeof(TF_Z) * 4 * ne);
  const int slots[4] = {0, 1, 2, 3};
  for (int it = 0; it < nt; ++it) {
    const int32_t *c = idx + 4 * it;
    memset(dwgt[0], 0, sizeof dwgt[0]);
    for (int v = 0; v < 4; ++v) dwgt[0][v][v] = 1.0;
    dvol[0] = 1.0;
    int ntet = 1;
    double ee1[4], ee2[4];
    for (int v = 0; v < 4; ++v) ee2[v] = creal(e2[c[v]]);
    tet_ttct(f2, ee2, +1, &ntet, dwgt, dvol);
    for (int v = 0; v < 4; ++v) ee1[v] = creal(e1[c[v]]);
    tet_ttct(f1, ee1, -1, &ntet, dwgt, dvol);
    for (int t = 0; t < ntet; ++t) {
      ro_z ffk[4], ggk[4];
      for (int i = 0; i < 4 * ne; ++i) xxo[i] = 0.0;
      for (int v = 0; v < 4; ++v) {
        double s = 0.0;
        for (int j = 0; j < 4; ++j) s += (ee2[j] - ee1[j]) * dwgt[t][v][j];
        ffk[v] = dvol[t] * dsgn;
        ggk[v] = -(s - I * delt) * dsgn;
      }
      TF(calc)(dmna, dmnr, slots, ffk, ggk, ne, em, xxo, use_series);
      for (int v = 0; v < 4; ++v)
        for (int j = 0; j < 4; ++j) {
          double w = dwgt[t][v][j];
          TF_Z *dst = xo + (size_t)ne * c[j];
          for (int ie = 0; ie < ne; ++ie) dst[ie] += xxo[ne * v + ie] * w;
        }
    }
  }
  free(xxo);
  free(dwgt);
}
