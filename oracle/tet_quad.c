/*
 * tet_quad.c -- the generalized-tetrahedron closed forms of tet_forms.h
 * (m_tetrahedron_20170325.F:99-451) evaluated in __float128.
 *
 * TEST INFRASTRUCTURE: the arbiter of the oracle's and the reference's
 * double results where the closed forms cancel.  The formula text is the one
 * oracle.c evaluates in double; only the arithmetic is wider (113-bit
 * significand, libquadmath's clogq), so the closed forms keep ~16 digits
 * even where they amplify rounding by 1e17.  Classification, sub-tetrahedron
 * cuts and merged nodes stay double, as the reference forms them.
 */
#include <quadmath.h>

#include "oracle.h"

#define TF_Z __complex128
#define TF_LOG clogq
#define TF(name) q_##name
#include "tet_forms.h"

static ro_z round_z(__complex128 z) { return (double)crealq(z) + I * (double)cimagq(z); }

void ro_ttrhdrn_chi0_quad(double dmna, double dmnr, int n1, int n2, int n3,
                          const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                          double f2, double delt, double dsgn, int ne, const ro_z *em,
                          ro_z *xo) {
  const size_t n = (size_t)ne * n1 * n2 * n3;
  __complex128 *acc = malloc(sizeof(__complex128) * n);
  for (size_t i = 0; i < n; ++i) acc[i] = (__complex128)xo[i];
  q_chi0(dmna, dmnr, n1, n2, n3, idx, e1, e2, f1, f2, delt, dsgn, ne, em, acc, 0);
  for (size_t i = 0; i < n; ++i) xo[i] = round_z(acc[i]);
  free(acc);
}

int ro_ttrhdrn_calc_quad(double dmna, double dmnr, const ro_z *fk, const ro_z *gk, int ne,
                         const ro_z *em, ro_z *xo) {
  const int c[4] = {0, 1, 2, 3};
  __complex128 *acc = malloc(sizeof(__complex128) * 4 * ne);
  for (int i = 0; i < 4 * ne; ++i) acc[i] = (__complex128)xo[i];
  const int rc = q_calc(dmna, dmnr, c, fk, gk, ne, em, acc, 0);
  for (int i = 0; i < 4 * ne; ++i) xo[i] = round_z(acc[i]);
  free(acc);
  return rc;
}
