/*
 * oracle.h -- CPU restatement of RESPACK's cRPA polarisation path
 * (src/chiqw) and the Wannier-basis integrals (src/calc_int).
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker.  The product
 * (respackdev_amd, librespack_gpu.so) never links, loads or calls it.
 *
 * Conventions: 0-based indices; complex numbers are C99 double complex (two
 * adjacent doubles, the same bytes as Fortran complex(8)); arrays that the
 * reference stores column-major keep that element order (first index fastest).
 * Each function cites the reference routine (file:line under src/) it restates.
 */
#ifndef RESPACK_ORACLE_H
#define RESPACK_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex ro_z;

/* chiqw/est_nwx2.f90:1-13 */
int ro_algn235(int n);
/* chiqw/gen_grid.f90:1-39 */
void ro_gen_grid(int nt, double emax, ro_z *em);
/* chiqw/chiqw_sub.F:773-812; kg is kg[3*i+d]; returns NG (capped writes at ntg) */
int ro_make_kg0(int ntg, const double *b1, const double *b2, const double *b3,
                double gcut, double q1, double q2, double q3, int32_t *kg);
/* chiqw/chiqw_sub.F:668-716 and :718-771 */
void ro_kcheck(double *k, int32_t *rw);
void ro_kcheck_trs(double *k, int32_t *rw);
/* chiqw/chiqw_sub.F:76-95 */
void ro_search_list(const double *sqi, double *sq);
/* chiqw/chiqw_sub.F:1-75; returns ikq (0-based), -1 if not found */
int ro_search_kq(int ntk, const double *sk0, double q1, double q2, double q3,
                 int ik, int32_t *shift_g);
/* chiqw/est_NTK.f90:1-39 */
void ro_est_nkbi(int n, const double *sk, int32_t *nkb);
/* chiqw/chiqw.F:437-480 (unfolding; also est_NTK.f90:42-98 for the count).
 * rg[9*iop + i + 3*j] = rg(i+1,j+1,iop+1).  Outputs sized 2*nk_irr*nsym;
 * returns NTK. numirr/numrot are 0-based. */
int ro_unfold_k(int nk_irr, int nsym, const double *ski, const int32_t *rg,
                double *sk0, int32_t *numirr, int32_t *numrot, int32_t *trs,
                int32_t *rw);
/* chiqw/chiqw_sub.F:324-384; idx[ix + n1*(iy + n2*iz)] = 0-based k */
void ro_make_index_kpt(int ntk, int n1, int n2, int n3, const double *sk0,
                       int32_t *idx);
/* chiqw/chiqw_sub.F:570-613; eig[ib + ntb*ik] */
void ro_judge_fermi(int nk, int ntb, const double *eig, double ef, int32_t *win,
                    int32_t *tindx, double *eave, double *bmax, double *bmin);
/* chiqw/chiqw_sub.F:264-322; pack is (2l1+1)(2l2+1)(2l3+1), 1-based G index
 * values as in the reference (0 = absent); rginv[i + 3*j]. */
void ro_make_c0(int itrs, int ng, const int32_t *kg, const int32_t *rw,
                const double *rginv, const int32_t *pg, int nnp, int l1, int l2,
                int l3, const int32_t *pack, const ro_z *occ, ro_z *c0);

/* m_tetrahedron_20170325.F:505-584 (libtetrabz corner table); idx[4*t + v] is
 * the 0-based k index of corner v of tetrahedron t (6 per k-cube). */
void ro_ttrhdrn_mkidx(int n1, int n2, int n3, const double *b1, const double *b2,
                      const double *b3, int32_t *idx);
/* m_tetrahedron_20170325.F:99-141; xo[ie + ne*ik] is ACCUMULATED.  The
 * vertex weights by the series of tet_forms.h where it converges
 * (|g - c| < 0.2 |e + c|), else by the reference's closed forms. */
void ro_ttrhdrn_chi0(double dmna, double dmnr, int n1, int n2, int n3,
                     const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                     double f2, double delt, double dsgn, int ne, const ro_z *em,
                     ro_z *xo);
/* The same with the reference's closed forms everywhere (as it evaluates). */
void ro_ttrhdrn_chi0_closed(double dmna, double dmnr, int n1, int n2, int n3,
                            const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                            double f2, double delt, double dsgn, int ne, const ro_z *em,
                            ro_z *xo);
/* The closed forms evaluated in __float128 (tet_quad.c; cuts, classes and
 * merged nodes in double as the reference forms them); xo accumulated in
 * __float128 from its input and rounded back. */
void ro_ttrhdrn_chi0_quad(double dmna, double dmnr, int n1, int n2, int n3,
                          const int32_t *idx, const ro_z *e1, const ro_z *e2, double f1,
                          double f2, double delt, double dsgn, int ne, const ro_z *em,
                          ro_z *xo);
/* m_tetrahedron_20170325.F:63-76 (ne = 1, em = 0); xo accumulated */
void ro_ttrhdrn_simple(double dmna, double dmnr, int n1, int n2, int n3,
                       const int32_t *idx, const ro_z *fk, const ro_z *gk, ro_z *xo);
/* One tetrahedron closed form (m_tetrahedron_20170325.F:304-451) at ne
 * frequencies, g[0..3] the reference's g0..g3; which: 0 sub1simple,
 * 1 sub1g30g21, 2 sub1g30g20, 3 sub1g30, 4 sub1normal, 5 sub2simple,
 * 6 sub2g30g20, 7 sub2g30g21, 8 sub2g30g10, 9 sub2g10, 10 sub2g30, 11 sub2normal. */
int ro_tet_sub(int which, const ro_z *g, int ne, const ro_z *em, ro_z *x);
/* calc (m_tetrahedron_20170325.F:209-302) for one tetrahedron: fk[4], gk[4]
 * in, xo[ie + ne*v] accumulated (v = 0..3); use_series as ro_ttrhdrn_chi0
 * (1) or ro_ttrhdrn_chi0_closed (0).  Returns -1 for a rejected class. */
int ro_ttrhdrn_calc(double dmna, double dmnr, const ro_z *fk, const ro_z *gk, int ne,
                    const ro_z *em, ro_z *xo, int use_series);
/* The same calc in __float128, closed forms only (tet_quad.c). */
int ro_ttrhdrn_calc_quad(double dmna, double dmnr, const ro_z *fk, const ro_z *gk, int ne,
                         const ro_z *em, ro_z *xo);
/* W_v of the node set gt[4] by the series of tet_forms.h at ne frequencies:
 * w[4*ie + v]; used[ie] = 0 where it does not converge fast enough (w unset).
 * Returns the number of frequencies evaluated. */
int ro_tet_series(const ro_z *gt, int ne, const ro_z *em, ro_z *w, int32_t *used);

/* Brute-force 3D DFT on an nx*ny*nz grid (x fastest).  dir = -1 matches
 * fft3_fw (e^{-i}, scaled 1/N), dir = +1 matches fft3_bw (e^{+i}, unscaled),
 * m_fft3d_20150826.f90:1598-1720. */
void ro_dft3(int nx, int ny, int nz, int dir, const ro_z *in, ro_z *out);

/* calc_InterStateMatrix, chiqw/chiqw_sub.F:97-205, restated in G space:
 * M(G) = sum_{G1 in kg_k} C_k(G1) conj(grid_kq[wrap(G1 + G + shift)]), where
 * grid_kq holds C_kq scattered at wrap(G2).  Same aliasing as the FFT route. */
void ro_pairdens(int ng_k, const int32_t *kg_k, const ro_z *c_k, int ng_kq,
                 const int32_t *kg_kq, const ro_z *c_kq, int nwx, int nwy, int nwz,
                 int ng_eps, const int32_t *lg, const int32_t *shift_g, ro_z *m);

/* calc_VMab (chiqw_sub.F:207-232) and calc_VMaa (:234-262); vm[3] */
void ro_vmab(int ng, const int32_t *kg, const ro_z *c_k, const ro_z *c_kq,
             const double *b1, const double *b2, const double *b3, ro_z *vm);
void ro_vmaa(int ng, const int32_t *kg, const ro_z *c_k, const ro_z *c_kq,
             const double *b1, const double *b2, const double *b3, const double *skt,
             ro_z *vm);

/* In-place inverse of a complex n*n matrix (column-major) by LU with partial
 * pivoting, the algorithm of LAPACK zgetrf+zgetri called at invZGE.F90:1-23.
 * Returns 0, or k+1 if U(k,k) is exactly zero (as zgetrf's info). */
int ro_zinv(int n, ro_z *a);

/* calc_eps_rpa (chiqw_sub.F:386-457): eps = delta - 2*chi*4pi/(|q+G||q+G'|),
 * inverted.  chi (n*n, column-major) is read; epsinv written. */
int ro_eps_rpa(int ng, const int32_t *lg, const double *q, const double *b1,
               const double *b2, const double *b3, const ro_z *chi, ro_z *epsinv);
/* calc_eps_rpa_q_0 (chiqw_sub.F:459-568) with head/wings and Drude term;
 * w = Re of the frequency (the routine uses dble(wcmplx) only). */
int ro_eps_rpa_q0(int ng, const int32_t *lg, const double *b1, const double *b2,
                  const double *b3, int g0, double wd, double delt, double w,
                  const ro_z *chi, ro_z *epsinv);

/*
 * One q != 0 of chiqw (chiqw.F:1257-1476), in the reference's loop order:
 * OpenMP over band pairs, then k; per pair the tetrahedron weights for
 * dsgn = +1 and -1, then pc0d(ig,jg,ie) += x(ik,ie) conj(M(ik,ig)) M(ik,jg)
 * / NTK * (1 - P_occ P_vir).  chi is divided by the cell volume at the end.
 *
 * Wavefunctions are given already unfolded to the full mesh (make_C0 applied):
 * c_occ[(ib*ntk + ik)*ntg + g], c_vir likewise; ng0[ik], kg0[(ik*ntg+g)*3+d].
 * e_occ[ib*ntk + ik], e_vir likewise (eigenvalues of numirr(ik)).
 * p_occ[ib*ntk + ik], p_vir likewise (cRPA projector weights; zeros = fRPA).
 * pair_occ/pair_vir: the band pairs (0-based into the occ/vir arrays).
 * ikq[ik], shift[3*ik]: k+q map from search_kq.
 * chi[(ie*ng + jg)*ng + ig] = chiqw(ig,jg,ie).
 * If m_out != NULL it receives M for pair 0: m_out[ik*ng + ig].
 * If x_out != NULL it receives x for pair 0: x_out[ik*nen + ie].
 * nthreads <= 0 uses the OpenMP default.  Returns 0.
 */
int ro_chiqw_q(int ntk, int ntg, const int32_t *ng0, const int32_t *kg0,
               int nocc, const ro_z *c_occ, int nvir, const ro_z *c_vir,
               const double *e_occ, const double *e_vir, const ro_z *p_occ,
               const ro_z *p_vir, int npair, const int32_t *pair_occ,
               const int32_t *pair_vir, const int32_t *ikq, const int32_t *shift,
               int nwx, int nwy, int nwz, int ng, const int32_t *lg, int nkb1,
               int nkb2, int nkb3, const int32_t *imt, const int32_t *index_kpt,
               double ef, double delt, double dmna, double dmnr, int nen,
               const ro_z *em, double volume, ro_z *chi, ro_z *m_out, ro_z *x_out,
               int nthreads);

/*
 * The q = 0 branch (chiqw.F:1557-1985) for gapped, non-intraband pairs: three
 * accumulators (q -> 0 along b1, b2, b3) whose G = 0 entry of M is replaced by
 * -i <vir|p|occ>/(e_vir - e_occ) (calc_VMab; zero when |de| <= delt).
 * chi3[((ix*nen + ie)*ng + jg)*ng + ig].  g0 = 0-based index of G = 0.
 */
int ro_chiqw_q0(int ntk, int ntg, const int32_t *ng0, const int32_t *kg0,
                int nocc, const ro_z *c_occ, int nvir, const ro_z *c_vir,
                const double *e_occ, const double *e_vir, const ro_z *p_occ,
                const ro_z *p_vir, int npair, const int32_t *pair_occ,
                const int32_t *pair_vir, int nwx, int nwy, int nwz, int ng,
                const int32_t *lg, int g0, const double *b1, const double *b2,
                const double *b3, int nkb1, int nkb2, int nkb3, const int32_t *imt,
                const int32_t *index_kpt, double ef, double delt, double dmna,
                double dmnr, int nen, const ro_z *em, double volume, ro_z *chi3,
                int nthreads);

/* ------------------------------------------------------------------------
 * calc_int, calc_w3d (src/calc_int/calc_intW.F, calc_int_sub.F)
 * ------------------------------------------------------------------------ */
/* make_LG0 (calc_int_sub.F:447-504): the eps sphere |q+G|^2 <= gcut_eps, then
 * gcut_eps < |q+G|^2 <= gcut_psi, each over [-100,100]^3 in g1 > g2 > g3 loop
 * order; lg[3*i+d].  Returns NG_for_psi (writes capped at ntg). */
int ro_make_lg0(int ntg, const double *b1, const double *b2, const double *b3,
                double gcut_eps, double gcut_psi, double q1, double q2, double q3,
                int32_t *lg, int32_t *ng_eps);
/* make_eps (calc_int_sub.F:222-311).  epsirr/epsmk single-precision complex,
 * Fortran (NTGQ,NTGQ,ne); lg[3*i+d] (ng entries used); rginv[i + 3*j];
 * pack[(i3+l1) + (2l1+1)*((j3+l2) + (2l2+1)*(k3+l3))] = 1-based G index. */
void ro_make_eps(int ntgq, int ne, int itrs, int ng, const int32_t *lg,
                 const int32_t *rw, const double *rginv, const int32_t *pg, int nnp,
                 int l1, int l2, int l3, const int32_t *pack, const float _Complex *epsirr,
                 float _Complex *epsmk);
/* func (calc_intW.F:833-847): func[(q*nwf + jw)*nwf + iw] for one q;
 * rho[w*ldr + g] (ng_psi used). */
void ro_w3d_func(int nwf, int ng_psi, int ldr, const ro_z *rho, ro_z *func);
/* funcw (calc_intW.F:848-912) for one q, the reference's loop order;
 * epsmk (NTGQ,NTGQ,ne) single precision; funcw[(ie*nwf + jw)*nwf + iw]. */
void ro_w3d_funcw(int nwf, int ng_eps, int ng_psi, int ldr, const ro_z *rho, int ntgq,
                  int ne, const float _Complex *epsmk, ro_z *funcw);
/* V_MAT / W_MAT (calc_intW.F:942-985): func[(q*nwf + j)*nwf + i],
 * funcw[((q*ne + ie)*nwf + j)*nwf + i], sq[3*q+d]; outputs Fortran
 * V(nwf,nwf,-na1:na1,-na2:na2,-na3:na3) and W(...,ne). */
void ro_w3d_vw(int nwf, int ntq, const double *sq, const ro_z *func, const ro_z *funcw,
               int ne, int na1, int na2, int na3, double volume, double chead,
               const ro_z *cheadw, ro_z *vmat, ro_z *wmat);

#ifdef __cplusplus
}
#endif
#endif
