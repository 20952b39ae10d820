! C-callable wrappers around subroutines of RESPACK's src/calc_int
! (calc_int_sub.F), compiled from the reference sources where they lie.
!
! TEST INFRASTRUCTURE ONLY.  The resulting library (oracle/_ref/librespack_ref_int.so)
! is a checker for the oracle's C restatement of calc_w3d; nothing in the
! product links or loads it.  It is a library of its own because calc_int_sub.F
! repeats subroutine names of chiqw_sub.F.
module rpref_shim_int
  use iso_c_binding
  use fft_3d
  implicit none
contains

  ! make_LG0 (calc_int_sub.F:447-504)
  subroutine rpref_make_lg0(ntg, b1, b2, b3, gcut_eps, gcut_psi, q1, q2, q3, lg0, ng_eps, &
       ng_psi) bind(C, name="rpref_make_lg0")
    integer(c_int), value :: ntg
    real(c_double) :: b1(3), b2(3), b3(3)
    real(c_double), value :: gcut_eps, gcut_psi, q1, q2, q3
    integer(c_int) :: lg0(3, ntg), ng_eps, ng_psi
    call make_LG0(ntg, b1, b2, b3, gcut_eps, gcut_psi, q1, q2, q3, lg0, ng_eps, ng_psi)
  end subroutine

  ! make_eps (calc_int_sub.F:222-311); epsirr/epsmk complex(4) (NTGQ,NTGQ,ne)
  subroutine rpref_make_eps(ntg, ntgq, ne, itrs, ng, lg, rw, rginv, pg, nnp, l1, l2, l3, &
       pack, epsirr, epsmk) bind(C, name="rpref_make_eps")
    integer(c_int), value :: ntg, ntgq, ne, itrs, ng, nnp, l1, l2, l3
    integer(c_int) :: lg(3, ntg), rw(3), pg(3)
    real(c_double) :: rginv(3, 3)
    integer(c_int) :: pack(-l1:l1, -l2:l2, -l3:l3)
    complex(c_float_complex) :: epsirr(ntgq, ntgq, ne), epsmk(ntgq, ntgq, ne)
    call make_eps(ntg, ntgq, ne, itrs, ng, lg, rw, rginv, pg, nnp, l1, l2, l3, pack, epsirr, epsmk)
  end subroutine

  ! calc_InterStateMatrix of calc_int (calc_int_sub.F:102-220) for one (w, k)
  subroutine rpref_int_pairdens(ntk, ntg, ng0, kg0, c0k, c0kq, ik, ikq, nwx, nwy, nwz, &
       lg0, ng, shift_g, m) bind(C, name="rpref_int_pairdens")
    integer(c_int), value :: ntk, ntg, ik, ikq, nwx, nwy, nwz, ng
    integer(c_int) :: ng0(ntk), kg0(3, ntg, ntk), lg0(3, ntg), shift_g(3)
    complex(c_double_complex) :: c0k(ntg), c0kq(ntg), m(ng)
    type(fft3_struct) :: fs
    real(8), allocatable :: wfunc(:), fftwk(:)
    integer :: nf1, nf2, nf3, nl
    nf1 = nwx + 1; nf2 = nwy + 1; nf3 = nwz + 1
    nl = nf1*nf2*nf3
    call fft3_init(nwx, nwy, nwz, nf1, nf2, nf3, fs)
    allocate(wfunc(2*nl), fftwk(2*nl))
    call calc_InterStateMatrix(ntk, ntg, ng0, kg0, c0k, c0kq, ik, ikq, nwx, nwy, nwz, &
         nf1, nf2, nl, wfunc, fftwk, fs, lg0, ng, shift_g, m)
    deallocate(wfunc, fftwk)
    call fft3_done(fs)
  end subroutine

end module rpref_shim_int
