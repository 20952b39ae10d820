// Restated setup routines of the reference shared by the host-side sources
// (k-mesh handling, G spheres, FFT and frequency grids, tetrahedron corners).
#include <algorithm>
#include <cmath>

#include "rp_internal.h"

namespace rp {

thread_local std::string g_err;


// ---------------------------------------------------------------- setup math
int algn235(int n) {  // est_nwx2.f90:1-42
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

void gen_grid(int nt, double e, double* em) {  // gen_grid.f90:1-39 (em interleaved)
  const double dlt = 0.05 / kAu, ratio = 0.1, eps = 1.0e-6;
  if (nt == 1) { em[0] = em[1] = 0.0; return; }
  int m = (int)(ratio * (double)nt), n = nt - m;
  double alp = e / dlt, x = 1.5, dl = 1.0;
  while (dl > eps) {
    double f = std::pow(x, n - 1) - alp * x + alp - 1.0;
    double g = (double)(n - 1) * std::pow(x, n - 2) - alp;
    double x1 = x - f / g;
    dl = std::fabs(x1 - x);
    x = x1;
  }
  double a = dlt / (x - 1.0), b = std::log(x);
  for (int i = 1; i <= n; ++i) { em[2 * (i - 1)] = a * std::exp(b * (double)(i - 1)) - a; em[2 * (i - 1) + 1] = 0; }
  double c = e * std::exp(-std::log(2.0) * (double)n / (double)m), d = std::log(2.0) / (double)m;
  for (int i = n + 1; i <= n + m; ++i) { em[2 * (i - 1)] = c * std::exp(d * (double)i); em[2 * (i - 1) + 1] = 0; }
}

// make_KG0 (chiqw_sub.F:773-812): g1 outer, g3 inner over [-100,100]^3; the
// loop is clipped to a box that contains the sphere, which keeps the order.
std::vector<int> make_kg0(const double b[3][3], const double a[3][3], double gcut, const double* q) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    double an = std::sqrt(a[d][0] * a[d][0] + a[d][1] * a[d][1] + a[d][2] * a[d][2]);
    int r = (int)std::ceil(std::sqrt(std::max(gcut, 0.0)) * an / (2.0 * M_PI) + std::fabs(q[d])) + 2;
    lo[d] = std::max(-100, -r);
    hi[d] = std::min(100, r);
  }
  std::vector<int> kg;
  for (int i1 = lo[0]; i1 <= hi[0]; ++i1)
    for (int i2 = lo[1]; i2 <= hi[1]; ++i2)
      for (int i3 = lo[2]; i3 <= hi[2]; ++i3) {
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = (q[0] + (double)i1) * b[0][d] + (q[1] + (double)i2) * b[1][d] + (q[2] + (double)i3) * b[2][d];
        double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (s <= gcut) { kg.push_back(i1); kg.push_back(i2); kg.push_back(i3); }
      }
  return kg;
}

void kcheck(double* k, int* rw) {  // chiqw_sub.F:668-716
  const double d = 1.0e-6;
  for (int i = 0; i < 3; ++i) {
    if (k[i] > 1.5 + d) { k[i] -= 2.0; rw[i] = -2; }
    if (k[i] > 0.5 + d) { k[i] -= 1.0; rw[i] = -1; }
    if (k[i] <= -1.5 + d) { k[i] += 2.0; rw[i] = 2; }
    if (k[i] <= -0.5 + d) { k[i] += 1.0; rw[i] = 1; }
  }
}
void kcheck_trs(double* k, int* rw) {  // chiqw_sub.F:718-771
  const double d = -1.0e-6;
  for (int i = 0; i < 3; ++i) {
    if (k[i] >= 1.5 + d) { k[i] -= 2.0; rw[i] = -2; }
    if (k[i] >= 0.5 + d) { k[i] -= 1.0; rw[i] = -1; }
    if (k[i] < -1.5 + d) { k[i] += 2.0; rw[i] = 2; }
    if (k[i] < -0.5 + d) { k[i] += 1.0; rw[i] = 1; }
  }
}
void search_list(const double* in, double* sq) {  // chiqw_sub.F:76-95
  const double d = 1.0e-6;
  for (int i = 0; i < 3; ++i) sq[i] = in[i];
  for (int i = 0; i < 2; ++i) {
    if (sq[i] > 1.5 + d) sq[i] -= 2.0;
    if (sq[i] > 0.5 + d) sq[i] -= 1.0;
    if (sq[i] <= -1.5 + d) sq[i] += 2.0;
    if (sq[i] <= -0.5 + d) sq[i] += 1.0;
  }
  if (sq[2] > 1.5 + d) sq[2] -= 2.0;
  if (sq[2] > 0.5 + d) sq[2] -= 1.0;
  if (sq[2] <= -0.5 + d) sq[2] += 1.0;
  if (sq[2] <= -1.5 + d) sq[2] += 2.0;
}
int search_kq(int ntk, const double* sk0, const double* q, int ik, int* sg) {  // chiqw_sub.F:1-75
  const double d = 1.0e-6;
  double s[3] = {sk0[3 * ik] + q[0], sk0[3 * ik + 1] + q[1], sk0[3 * ik + 2] + q[2]};
  sg[0] = sg[1] = sg[2] = 0;
  for (int i = 0; i < 3; ++i) {
    if (s[i] > 1.5 + d) { s[i] -= 2.0; sg[i] = 2; }
    if (s[i] > 0.5 + d) { s[i] -= 1.0; sg[i] = 1; }
    if (s[i] <= -1.5 + d) { s[i] += 2.0; sg[i] = -2; }
    if (s[i] <= -0.5 + d) { s[i] += 1.0; sg[i] = -1; }
  }
  int ikq = -1;
  for (int jk = 0; jk < ntk; ++jk)
    if (std::fabs(sk0[3 * jk] - s[0]) < 1e-6 && std::fabs(sk0[3 * jk + 1] - s[1]) < 1e-6 &&
        std::fabs(sk0[3 * jk + 2] - s[2]) < 1e-6)
      ikq = jk;
  return ikq;
}

void mkidx(int n1, int n2, int n3, const double* b, int* idx) {
  // libtetrabz corner table (m_tetrahedron_20170325.F:523-584)
  int ng[3] = {n1, n2, n3};
  double bv[3][3], b3v[4][3], l[4];
  for (int d = 0; d < 3; ++d) { bv[0][d] = b[d] / n1; bv[1][d] = b[3 + d] / n2; bv[2][d] = b[6 + d] / n3; }
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
  int v0[4] = {0, 0, 0, 0}, dv[4][4] = {{0}};
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



double reciprocal(const double a[3][3], double b[3][3]) {
  auto cross = [](const double* x, const double* y, double* z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
  };
  const double tpi = 2.0 * std::acos(-1.0);
  cross(a[1], a[2], b[0]);
  const double volume = a[0][0] * b[0][0] + a[0][1] * b[0][1] + a[0][2] * b[0][2];
  cross(a[2], a[0], b[1]);
  cross(a[0], a[1], b[2]);
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) b[i][d] = b[i][d] * tpi / volume;
  return volume;
}

void fft_grid(const double a[3][3], const double b[3][3], double ecut_psi, int nw[3]) {
  const double qwf = 2.0 * std::sqrt(ecut_psi);
  for (int d = 0; d < 3; ++d) {
    double an = std::sqrt(a[d][0] * a[d][0] + a[d][1] * a[d][1] + a[d][2] * a[d][2]);
    double dd = std::fabs((b[d][0] * a[d][0] + b[d][1] * a[d][1] + b[d][2] * a[d][2]) / an);
    nw[d] = algn235((int)(qwf / dd) + 1);
  }
}

double ecut_for_psi(const double b[3][3], int nk_irr, const double* ski, const int* ngi, int ntg,
                    const int* kgi) {
  double emx = 0;
  for (int ik = 0; ik < nk_irr; ++ik)
    for (int ig = 0; ig < ngi[ik]; ++ig) {
      const int* g = kgi + 3 * ((size_t)ik * ntg + ig);
      double v[3];
      for (int d = 0; d < 3; ++d)
        v[d] = (ski[3 * ik] + (double)g[0]) * b[0][d] + (ski[3 * ik + 1] + (double)g[1]) * b[1][d] +
               (ski[3 * ik + 2] + (double)g[2]) * b[2][d];
      emx = std::max(emx, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
  return emx + 1.0e-8;
}

}  // namespace rp
