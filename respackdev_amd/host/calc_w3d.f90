! calc_w3d -- drop-in for RESPACK's calc_w3d (src/calc_int/calc_intW.F) with
! the rho / func / funcw / Fourier-sum work on one MI355X through
! librespack_gpu.
!
! Same process contract as the reference: the namelist &param_calc_int on
! standard input (calc_int/m_rdinput.f90:1-36), ./dir-wfn (lattice, symmetry,
! k list and G lists), ./dir-wan (dat.wan, dat.wan-center) and ./dir-eps
! (dat.chi_cutoff, dat.wgrid, qNNN/dat.sq and dat.epsqw.600, or 600/601/602 at
! Gamma) in; ./dir-intW (dat.Vmat, dat.Wmat, dat.VvsR.NNN, dat.WvsR.NNN,
! dat.UvsE.III-JJJ) and ./dir-mvmc (zCoulombIntra.def, zCoulombInter.def) out,
! in the formats of calc_intW.F:1000-1146 and wrt_mvmc_w3d.f90.  dir-mvmc is
! created when missing.
program calc_w3d
  use iso_c_binding
  use respack_gpu
  use rp_host_io
  implicit none
  real(8), parameter :: au = 27.21151d0, bohr = 0.529177d0
  integer :: calc_ifreq = 1, ix_intJ_min = 0, ix_intJ_max = 0, iy_intJ_min = 0, iy_intJ_max = 0
  integer :: iz_intJ_min = 0, iz_intJ_max = 0
  real(8) :: wcut_mvmc = 2.0d0, jcut_mvmc = 0.3d0
  namelist /param_calc_int/ calc_ifreq, ix_intJ_min, ix_intJ_max, iy_intJ_min, iy_intJ_max, &
    iz_intJ_min, iz_intJ_max, wcut_mvmc, jcut_mvmc

  type(int_system) :: sys
  real(8), allocatable :: heads(:)
  complex(8), allocatable :: vmat(:, :, :, :, :), wmat(:, :, :, :, :, :)
  integer :: nen, ntk, nwf, na(3), nkb(3)

  read(5, nml=param_calc_int)
  write(6, nml=param_calc_int)
  call load_int_system(sys)
  nen = sys%nen
  ntk = sys%ntk
  nwf = sys%nwf
  na = sys%na
  nkb = sys%nkb
  if (calc_ifreq < 1 .or. calc_ifreq > nen) then
    write(6, *) 'ERROR: calc_ifreq out of range 1..', nen
    error stop 1
  end if
  allocate(vmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3)))
  allocate(wmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3), nen))
  allocate(heads(2 + 2 * nen))
  call rp_check(rp_w3d_vw(sys%h, vmat, wmat, heads), 'rp_w3d_vw')
  call rp_check(rp_finalize(sys%h), 'rp_finalize')
  write(6, '(a,f15.10)') ' chead(HL) in eV=', heads(1) * au

  call make_dir('dir-intW')
  call write_mat('dir-intW/dat.Vmat', '#Bare Coulomb: V', vmat)
  call write_mat('dir-intW/dat.Wmat', '#Screened Coulomb: W at omega=calc_ifreq', &
                 wmat(:, :, :, :, :, calc_ifreq))
  call write_vs_r('dat.VvsR.', '#R dependence of bare Coulomb', &
                  '#1:R [AA], 2:Re[V(R)] [eV], 3:Im[V(R)] [eV]', vmat)
  call write_vs_r('dat.WvsR.', '#R dependence of screened Coulomb at omega=calc_ifreq', &
                  '#1:R [AA], 2:Re[W(R)] [eV], 3:Im[W(R)] [eV]', wmat(:, :, :, :, :, calc_ifreq))
  call write_u_vs_e()
  call make_dir('dir-mvmc')
  call write_mvmc(wmat(:, :, :, :, :, calc_ifreq), wcut_mvmc / au)
  write(6, *) 'calc_w3d: done'

contains

  ! dat.Vmat / dat.Wmat (calc_intW.F:1007-1074)
  subroutine write_mat(path, title, m)
    character(len=*), intent(in) :: path, title
    complex(8), intent(in) :: m(:, :, -na(1):, -na(2):, -na(3):)
    integer :: uu, i1, i2, i3, i, j
    open(newunit=uu, file=path)
    rewind(uu)
    write(uu, '(a)') title
    write(uu, '(a)') '#1:R1, 2:R2, 3:R3 (lattice vector)'
    if (title(2:5) == 'Bare') then
      write(uu, '(a)') '#1:i, 2:j, 3:Re(V_ij) [eV], 4:Im(V_ij) [eV]'
    else
      write(uu, '(a)') '#1:i, 2:j, 3:Re(W_ij) [eV], 4:Im(W_ij) [eV]'
    end if
    do i1 = -na(1), na(1)
      do i2 = -na(2), na(2)
        do i3 = -na(3), na(3)
          write(uu, *) i1, i2, i3
          do i = 1, nwf
            do j = 1, nwf
              write(uu, '(i5,i5,2f20.10)') i, j, m(i, j, i1, i2, i3) * au
            end do
          end do
          write(uu, *)
        end do
      end do
    end do
    close(uu)
  end subroutine

  ! dat.VvsR.NNN / dat.WvsR.NNN (calc_intW.F:1076-1127)
  subroutine write_vs_r(prefix, title, cols, m)
    character(len=*), intent(in) :: prefix, title, cols
    complex(8), intent(in) :: m(:, :, -na(1):, -na(2):, -na(3):)
    character(len=64) :: name
    integer :: uu, i1, i2, i3, i, j
    real(8) :: rij(3)
    do i = 1, nwf
      write(name, '(a,i3.3)') prefix, i
      open(newunit=uu, file='dir-intW/' // trim(name))
      rewind(uu)
      write(uu, '(a)') title
      write(uu, '(a)') cols
      do i1 = -na(1), na(1)
        do i2 = -na(2), na(2)
          do i3 = -na(3), na(3)
            do j = 1, nwf
              rij = dble(i1) * sys%a(:, 1) + dble(i2) * sys%a(:, 2) + dble(i3) * sys%a(:, 3) &
                    + (sys%coord(:, j) - sys%coord(:, i))
              write(uu, '(3F12.5)') dsqrt(rij(1)**2 + rij(2)**2 + rij(3)**2) * bohr, &
                                    m(i, j, i1, i2, i3) * au
            end do
          end do
        end do
      end do
      close(uu)
    end do
  end subroutine

  ! dat.UvsE.III-JJJ (calc_intW.F:1128-1146)
  subroutine write_u_vs_e()
    character(len=64) :: name
    integer :: uu, i, j, k
    do i = 1, nwf
      do j = i, nwf
        write(name, '("dat.UvsE.",i3.3,"-",i3.3)') i, j
        open(newunit=uu, file='dir-intW/' // trim(name))
        rewind(uu)
        write(uu, '(a)') '#Omega dependence of onsite Coulomb'
        write(uu, '(a)') '#1:Re(w) [eV], 2:Im(w) [eV], 3:Re[U(w)] [eV], 4:Im[U(w)] [eV]'
        do k = 1, nen
          write(uu, '(4F12.5)') sys%em(k) * au, wmat(i, j, 0, 0, 0, k) * au
        end do
        close(uu)
      end do
    end do
  end subroutine

  ! zCoulombIntra.def / zCoulombInter.def (wrt_mvmc_w3d.f90)
  subroutine write_mvmc(hr, cut)
    complex(8), intent(in) :: hr(:, :, -na(1):, -na(2):, -na(3):)
    real(8), intent(in) :: cut
    integer :: nsite, isite, jsite, ib, jb, i, j, k, uu
    complex(8), allocatable :: h2(:, :, :, :)
    nsite = ntk
    allocate(h2(nwf, nwf, nsite, nsite))
    call write_mvmc_sites(nkb, na, nwf, hr, h2)
    k = nsite * nwf
    open(newunit=uu, file='dir-mvmc/zCoulombIntra.def')
    write(uu, '(a30)') '=============================='
    write(uu, '(a13,i17)') 'NCoulombIntra', k
    write(uu, '(a30)') '=============================='
    write(uu, '(a30)') '=============================='
    write(uu, '(a30)') '=============================='
    do isite = 1, nsite
      do ib = 1, nwf
        i = ib + (isite - 1) * nwf
        write(uu, '(i5,2f15.10)') i - 1, dble(h2(ib, ib, isite, isite)) * au
      end do
    end do
    close(uu)
    k = 0
    do isite = 1, nsite
      do ib = 1, nwf
        i = ib + (isite - 1) * nwf
        do jsite = 1, nsite
          do jb = 1, nwf
            j = jb + (jsite - 1) * nwf
            if (i >= j) cycle
            if (abs(h2(ib, jb, isite, jsite)) < cut) cycle
            k = k + 1
          end do
        end do
      end do
    end do
    open(newunit=uu, file='dir-mvmc/zCoulombInter.def')
    write(uu, '(a30)') '=============================='
    write(uu, '(a13,i17)') 'NCoulombInter', k
    write(uu, '(a30)') '=============================='
    write(uu, '(a10,f10.5,a10)') '=== Vcut >', cut * au, 'eV ======='
    write(uu, '(a30)') '=============================='
    do isite = 1, nsite
      do ib = 1, nwf
        i = ib + (isite - 1) * nwf
        do jsite = 1, nsite
          do jb = 1, nwf
            j = jb + (jsite - 1) * nwf
            if (i >= j) cycle
            if (abs(h2(ib, jb, isite, jsite)) < cut) cycle
            write(uu, '(2i5,2f15.10)') i - 1, j - 1, dble(h2(ib, jb, isite, jsite)) * au
          end do
        end do
      end do
    end do
    close(uu)
  end subroutine

end program calc_w3d
