! C-callable wrapper of the reference's mVMC export writer wrt_mvmc
! (src/calc_int/wrt_mvmc_w3d.f90 or wrt_mvmc_j3d.f90, which share its name and
! signature), compiled from the reference sources where they lie.  Linked once
! with each of the two files into oracle/_ref/librespack_ref_mvmc{w,j}.so.
!
! TEST INFRASTRUCTURE ONLY: the checker of the drop-in host's dir-mvmc files;
! nothing in the product links or loads it.  The files are written into the
! current working directory, as the reference writes them into ./dir-mvmc
! after changing into it (calc_intW.F:1150-1157, calc_intJ.F:1035-1042).
subroutine rpref_wrt_mvmc(nkb1, nkb2, nkb3, ntk, na1, na2, na3, nwf, cut, hr) &
     bind(C, name="rpref_wrt_mvmc")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nkb1, nkb2, nkb3, ntk, na1, na2, na3, nwf
  real(c_double), value :: cut
  complex(c_double_complex) :: hr(nwf, nwf, -na1:na1, -na2:na2, -na3:na3)
  integer :: k1, k2, k3, nt, a1, a2, a3, n
  real(8) :: c
  k1 = nkb1; k2 = nkb2; k3 = nkb3; nt = ntk; a1 = na1; a2 = na2; a3 = na3; n = nwf
  c = cut
  call wrt_mvmc(k1, k2, k3, nt, a1, a2, a3, n, c, hr)
end subroutine
