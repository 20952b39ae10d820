! C-callable wrappers around individual subroutines of the RESPACK reference
! (src/chiqw), compiled from the reference sources where they lie.
!
! TEST INFRASTRUCTURE ONLY.  The resulting library (oracle/_ref/librespack_ref.so)
! is used by tests/ as a checker for the oracle's C restatement; nothing in the
! product links or loads it.  The files compiled are m_fft3d_20150826.f90,
! m_tetrahedron_20170325.F, gen_grid.f90, est_nwx2.f90, est_NTK.f90,
! chiqw_sub.F and invZGE.F90; the LAPACK routines zgetrf/zgetri that invZGE
! calls (the Dyson step calc_eps_rpa / calc_eps_rpa_q_0) come from the Intel
! MKL installed in the image (/opt/conda/lib), as in the reference's own
! Makefile (src/chiqw/Makefile:10).  No MPI is needed by these files.
module rpref_shim
  use iso_c_binding
  use fft_3d
  use m_tetrahedron
  implicit none
contains

  subroutine rpref_gen_grid(nt, emax, em) bind(C, name="rpref_gen_grid")
    integer(c_int), value :: nt
    real(c_double), value :: emax
    complex(c_double_complex) :: em(nt)
    real(8) :: e
    e = emax
    call gen_grid(nt, e, em)
  end subroutine

  function rpref_algn235(n) bind(C, name="rpref_algn235") result(r)
    integer(c_int), value :: n
    integer(c_int) :: r
    integer :: algn235
    external :: algn235
    r = algn235(int(n))
  end function

  subroutine rpref_make_kg0(ntg, b1, b2, b3, gcut, q1, q2, q3, kg0, ng) &
       bind(C, name="rpref_make_kg0")
    integer(c_int), value :: ntg
    real(c_double) :: b1(3), b2(3), b3(3)
    real(c_double), value :: gcut, q1, q2, q3
    integer(c_int) :: kg0(3, ntg)
    integer(c_int) :: ng
    call make_KG0(ntg, b1, b2, b3, gcut, q1, q2, q3, kg0, ng)
  end subroutine

  subroutine rpref_search_kq(ntk, sk0, q1, q2, q3, ik, ikq, shift_g) &
       bind(C, name="rpref_search_kq")
    integer(c_int), value :: ntk, ik
    real(c_double) :: sk0(3, ntk)
    real(c_double), value :: q1, q2, q3
    integer(c_int) :: ikq, shift_g(3)
    real(8) :: a, b, c
    a = q1; b = q2; c = q3
    shift_g = 0
    ikq = 0
    call search_kq(ntk, sk0, a, b, c, ik, ikq, shift_g)
  end subroutine

  subroutine rpref_search_list(sqi, sq) bind(C, name="rpref_search_list")
    real(c_double) :: sqi(3), sq(3)
    call search_list(sqi, sq)
  end subroutine

  subroutine rpref_kcheck(ktmp, rw, trs) bind(C, name="rpref_kcheck")
    real(c_double) :: ktmp(3)
    integer(c_int) :: rw(3)
    integer(c_int), value :: trs
    rw = 0
    if (trs == 1) then
      call kcheck(ktmp, rw)
    else
      call kcheck_trs(ktmp, rw)
    end if
  end subroutine

  subroutine rpref_est_ntk(nk_irr, nsymq, ski, rg, ntk) bind(C, name="rpref_est_ntk")
    integer(c_int), value :: nk_irr, nsymq
    real(c_double) :: ski(3, nk_irr)
    integer(c_int) :: rg(3, 3, nsymq)
    integer(c_int) :: ntk
    call est_NTK(nk_irr, nsymq, ski, rg, ntk)
  end subroutine

  subroutine rpref_est_nkbi(n, sk, nkb) bind(C, name="rpref_est_nkbi")
    integer(c_int), value :: n
    real(c_double) :: sk(3, n)
    integer(c_int) :: nkb(3)
    call est_nkbi(n, sk, nkb(1), nkb(2), nkb(3))
  end subroutine

  subroutine rpref_make_index_kpt(ntk, nkb1, nkb2, nkb3, sk0, idx) &
       bind(C, name="rpref_make_index_kpt")
    integer(c_int), value :: ntk, nkb1, nkb2, nkb3
    real(c_double) :: sk0(3, ntk)
    integer(c_int) :: idx(nkb1, nkb2, nkb3)
    idx = 0
    call make_index_kpt(ntk, nkb1, nkb2, nkb3, sk0, idx)
  end subroutine

  subroutine rpref_judge_fermi(nk, ntb, eig, ef, win, tindx, eave, bmax, bmin) &
       bind(C, name="rpref_judge_fermi")
    integer(c_int), value :: nk, ntb
    real(c_double) :: eig(ntb, nk)
    real(c_double), value :: ef
    integer(c_int) :: win(ntb), tindx(ntb)
    real(c_double) :: eave(ntb), bmax(ntb), bmin(ntb)
    real(8) :: e
    e = ef
    call judge_FermiInside(nk, ntb, eig, e, win, tindx, eave, bmax, bmin)
  end subroutine

  subroutine rpref_make_c0(ntg, itrs, ng, kg, rw, rginv, pg, nnp, l1, l2, l3, &
       pack, occ, c0) bind(C, name="rpref_make_c0")
    integer(c_int), value :: ntg, itrs, ng, nnp, l1, l2, l3
    integer(c_int) :: kg(3, ntg), rw(3), pg(3)
    real(c_double) :: rginv(3, 3)
    integer(c_int) :: pack(-l1:l1, -l2:l2, -l3:l3)
    complex(c_double_complex) :: occ(ntg), c0(ntg)
    call make_C0(ntg, itrs, ng, kg, rw, rginv, pg, nnp, l1, l2, l3, pack, occ, c0)
  end subroutine

  subroutine rpref_mkidx(n1, n2, n3, idx, b1, b2, b3) bind(C, name="rpref_mkidx")
    integer(c_int), value :: n1, n2, n3
    integer(c_int) :: idx(4, n1*n2*n3*6)
    real(c_double) :: b1(3), b2(3), b3(3)
    call ttrhdrn_mkidx(n1, n2, n3, idx, b1, b2, b3)
  end subroutine

  ! xo is accumulated (intent inout), as in the reference.
  subroutine rpref_chi0(dmna, dmnr, n1, n2, n3, idx, e1, e2, f1, f2, delt, dsgn, &
       ne, em, xo) bind(C, name="rpref_chi0")
    real(c_double), value :: dmna, dmnr, f1, f2, delt, dsgn
    integer(c_int), value :: n1, n2, n3, ne
    integer(c_int) :: idx(4, n1*n2*n3*6)
    complex(c_double_complex) :: e1(n1*n2*n3), e2(n1*n2*n3), em(ne)
    complex(c_double_complex) :: xo(ne, n1*n2*n3)
    complex(8), allocatable :: ca1(:)
    allocate(ca1(4*ne))
    ca1 = 0
    call ttrhdrn_chi0(dmna, dmnr, n1, n2, n3, idx, e1, e2, f1, f2, delt, dsgn, &
         ne, em, ca1, xo)
    deallocate(ca1)
  end subroutine

  subroutine rpref_simple(dmna, dmnr, n1, n2, n3, idx, fk, gk, xo) &
       bind(C, name="rpref_simple")
    real(c_double), value :: dmna, dmnr
    integer(c_int), value :: n1, n2, n3
    integer(c_int) :: idx(4, n1*n2*n3*6)
    complex(c_double_complex) :: fk(n1*n2*n3), gk(n1*n2*n3), xo(n1*n2*n3)
    call ttrhdrn_simple(dmna, dmnr, n1, n2, n3, idx, fk, gk, xo)
  end subroutine

  ! One 3D FFT of an nx*ny*nz complex grid stored x-fastest without padding.
  ! dir = -1: fft3_fw (e^{-i}, scaled 1/N); dir = +1: fft3_bw (e^{+i}, unscaled).
  subroutine rpref_fft3(nx, ny, nz, dir, z) bind(C, name="rpref_fft3")
    integer(c_int), value :: nx, ny, nz, dir
    complex(c_double_complex) :: z(nx, ny, nz)
    type(fft3_struct) :: fs
    real(8), allocatable :: dat(:), tmp(:)
    integer :: l1, l2, l3, n, i, j, k, ind
    l1 = nx + 1; l2 = ny + 1; l3 = nz + 1
    n = l1*l2*l3
    call fft3_init(nx, ny, nz, l1, l2, l3, fs)
    allocate(dat(2*n), tmp(2*n))
    dat = 0; tmp = 0
    do k = 1, nz; do j = 1, ny; do i = 1, nx
      ind = i + (j-1)*l1 + (k-1)*l1*l2
      dat(ind) = dble(z(i,j,k)); dat(ind+n) = dimag(z(i,j,k))
    end do; end do; end do
    if (dir < 0) then
      call fft3_fw(fs, dat, tmp)
    else
      call fft3_bw(fs, dat, tmp)
    end if
    do k = 1, nz; do j = 1, ny; do i = 1, nx
      ind = i + (j-1)*l1 + (k-1)*l1*l2
      z(i,j,k) = cmplx(dat(ind), dat(ind+n), kind=8)
    end do; end do; end do
    call fft3_done(fs)
    deallocate(dat, tmp)
  end subroutine

  ! calc_InterStateMatrix for one (pair, k): C0_K at k, C0_KQ at k+q.
  subroutine rpref_pairdens(ntk, ntg, ng0, kg0, c0k, c0kq, ik, ikq, nwx, nwy, nwz, &
       lg0, ngeps, shift_g, m) bind(C, name="rpref_pairdens")
    integer(c_int), value :: ntk, ntg, ik, ikq, nwx, nwy, nwz, ngeps
    integer(c_int) :: ng0(ntk), kg0(3, ntg, ntk), lg0(3, ntg), shift_g(3)
    complex(c_double_complex) :: c0k(ntg), c0kq(ntg), m(ngeps)
    type(fft3_struct) :: fs
    real(8), allocatable :: wfunc(:), fftwk(:)
    integer :: nf1, nf2, nf3, nl
    nf1 = nwx + 1; nf2 = nwy + 1; nf3 = nwz + 1
    nl = nf1*nf2*nf3
    call fft3_init(nwx, nwy, nwz, nf1, nf2, nf3, fs)
    allocate(wfunc(2*nl), fftwk(2*nl))
    call calc_InterStateMatrix(ntk, ntg, ng0, kg0, c0k, c0kq, ik, ikq, nwx, nwy, nwz, &
         nf1, nf2, nl, wfunc, fftwk, fs, lg0, ngeps, shift_g, m)
    deallocate(wfunc, fftwk)
    call fft3_done(fs)
  end subroutine

  subroutine rpref_vmab(ntk, ntg, ng0, kg0, c0k, c0kq, ik, b1, b2, b3, vm) &
       bind(C, name="rpref_vmab")
    integer(c_int), value :: ntk, ntg, ik
    integer(c_int) :: ng0(ntk), kg0(3, ntg, ntk)
    complex(c_double_complex) :: c0k(ntg), c0kq(ntg), vm(3)
    real(c_double) :: b1(3), b2(3), b3(3)
    call calc_VMab(ntk, ntg, ng0, kg0, c0k, c0kq, ik, b1, b2, b3, vm)
  end subroutine

  subroutine rpref_vmaa(ntk, ntg, ng0, kg0, c0k, c0kq, ik, b1, b2, b3, skt, vm) &
       bind(C, name="rpref_vmaa")
    integer(c_int), value :: ntk, ntg, ik
    integer(c_int) :: ng0(ntk), kg0(3, ntg, ntk)
    complex(c_double_complex) :: c0k(ntg), c0kq(ntg), vm(3)
    real(c_double) :: b1(3), b2(3), b3(3), skt(3)
    call calc_VMaa(ntk, ntg, ng0, kg0, c0k, c0kq, ik, b1, b2, b3, skt, vm)
  end subroutine

  ! calc_eps_rpa (chiqw_sub.F:386-457): chi0 (NG x NG, doubled in place as
  ! the reference does) -> eps^-1 by zgetrf/zgetri.  Its diagonal text output
  ! goes to a scratch unit.
  subroutine rpref_eps_rpa(ntg, ng, lg0, q1, q2, q3, b1, b2, b3, chi0, eps) &
       bind(C, name="rpref_eps_rpa")
    integer(c_int), value :: ntg, ng
    integer(c_int) :: lg0(3, ntg)
    real(c_double), value :: q1, q2, q3
    real(c_double) :: b1(3), b2(3), b3(3)
    complex(c_double_complex) :: chi0(ng, ng), eps(ng, ng)
    integer :: u
    real(8) :: tpi, a, b, c
    tpi = 2.0d0*dacos(-1.0d0)   ! chiqw/config.h:20-21
    a = q1; b = q2; c = q3
    open(newunit=u, status='scratch')
    call calc_eps_rpa(ntg, ng, lg0, a, b, c, b1, b2, b3, chi0, tpi, eps, 0, u)
    close(u)
  end subroutine

  ! calc_eps_rpa_q_0 (chiqw_sub.F:459-568): head, wings, Drude term of plasma
  ! frequency wd at complex frequency wcmplx (only its real part is used).
  subroutine rpref_eps_rpa_q0(ntg, ng, lg0, b1, b2, b3, chi0, g0, wd, delt, wre, wim, eps) &
       bind(C, name="rpref_eps_rpa_q0")
    integer(c_int), value :: ntg, ng, g0
    integer(c_int) :: lg0(3, ntg)
    real(c_double) :: b1(3), b2(3), b3(3)
    real(c_double), value :: wd, delt, wre, wim
    complex(c_double_complex) :: chi0(ng, ng), eps(ng, ng)
    integer :: u, ng0
    real(8) :: tpi, z, wdd, dd
    complex(8) :: wc
    tpi = 2.0d0*dacos(-1.0d0)
    z = 0.0d0
    wdd = wd; dd = delt
    wc = cmplx(wre, wim, kind=8)
    ng0 = g0
    open(newunit=u, status='scratch')
    call calc_eps_rpa_q_0(ntg, ng, lg0, z, z, z, b1, b2, b3, chi0, tpi, ng0, eps, wdd, dd, wc, 0, u)
    close(u)
  end subroutine

  ! One closed form of the generalized tetrahedron method (the public
  ! ttrhdrn_sub* of m_tetrahedron_20170325.F:304-451) at ne frequencies:
  ! which = 0 sub1simple, 1 sub1g30g21, 2 sub1g30g20, 3 sub1g30, 4 sub1normal,
  ! 5 sub2simple, 6 sub2g30g20, 7 sub2g30g21, 8 sub2g30g10, 9 sub2g10,
  ! 10 sub2g30, 11 sub2normal.
  subroutine rpref_tet_sub(which, g, ne, em, x) bind(C, name="rpref_tet_sub")
    integer(c_int), value :: which, ne
    complex(c_double_complex) :: g(4), em(ne), x(ne)
    select case (which)
    case (0); call ttrhdrn_sub1simple(g(1), g(2), g(3), g(4), ne, em, x)
    case (1); call ttrhdrn_sub1g30g21(g(1), g(2), g(3), g(4), ne, em, x)
    case (2); call ttrhdrn_sub1g30g20(g(1), g(2), g(3), g(4), ne, em, x)
    case (3); call ttrhdrn_sub1g30(g(1), g(2), g(3), g(4), ne, em, x)
    case (4); call ttrhdrn_sub1normal(g(1), g(2), g(3), g(4), ne, em, x)
    case (5); call ttrhdrn_sub2simple(g(1), g(2), g(3), g(4), ne, em, x)
    case (6); call ttrhdrn_sub2g30g20(g(1), g(2), g(3), g(4), ne, em, x)
    case (7); call ttrhdrn_sub2g30g21(g(1), g(2), g(3), g(4), ne, em, x)
    case (8); call ttrhdrn_sub2g30g10(g(1), g(2), g(3), g(4), ne, em, x)
    case (9); call ttrhdrn_sub2g10(g(1), g(2), g(3), g(4), ne, em, x)
    case (10); call ttrhdrn_sub2g30(g(1), g(2), g(3), g(4), ne, em, x)
    case (11); call ttrhdrn_sub2normal(g(1), g(2), g(3), g(4), ne, em, x)
    end select
  end subroutine

end module rpref_shim
