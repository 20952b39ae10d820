! calc_j3d -- drop-in for RESPACK's calc_j3d (src/calc_int/calc_intJ.F): the
! bare and screened exchange integrals X_ij(R), J_ij(R, w) on one MI355X
! through librespack_gpu.
!
! Same inputs and namelist as calc_w3d (&param_calc_int; the lattice vectors
! are ix/iy/iz_intJ_min..max); writes ./dir-intJ (dat.Xmat, dat.Jmat,
! dat.JvsE.III-JJJ) and ./dir-mvmc (zHund.def, zExchange.def, zPairHop.def) in
! the formats of calc_intJ.F:920-1030 and wrt_mvmc_j3d.f90, J taken at
! calc_ifreq and zero for lattice vectors outside the computed range, as in the
! reference.
program calc_j3d
  use iso_c_binding
  use respack_gpu
  use rp_host_io
  implicit none
  real(8), parameter :: au = 27.21151d0
  integer :: calc_ifreq = 1, ix_intJ_min = 0, ix_intJ_max = 0, iy_intJ_min = 0, iy_intJ_max = 0
  integer :: iz_intJ_min = 0, iz_intJ_max = 0
  real(8) :: wcut_mvmc = 2.0d0, jcut_mvmc = 0.3d0
  namelist /param_calc_int/ calc_ifreq, ix_intJ_min, ix_intJ_max, iy_intJ_min, iy_intJ_max, &
    iz_intJ_min, iz_intJ_max, wcut_mvmc, jcut_mvmc

  type(int_system) :: sys
  real(8), allocatable :: heads(:)
  integer(c_int32_t), allocatable :: rvec(:, :)
  complex(8), allocatable :: x(:, :, :), j(:, :, :, :)
  complex(8), allocatable :: xmat(:, :, :, :, :), jmat(:, :, :, :, :, :)
  integer :: nen, ntk, nwf, na(3), nkb(3), nr, r, i1, i2, i3

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
  if (ix_intJ_min < -na(1) .or. ix_intJ_max > na(1) .or. iy_intJ_min < -na(2) .or. &
      iy_intJ_max > na(2) .or. iz_intJ_min < -na(3) .or. iz_intJ_max > na(3)) then
    write(6, *) 'ERROR: intJ range outside -Na..Na'
    error stop 1
  end if
  nr = max(0, ix_intJ_max - ix_intJ_min + 1) * max(0, iy_intJ_max - iy_intJ_min + 1) * &
       max(0, iz_intJ_max - iz_intJ_min + 1)
  if (nr == 0) then
    write(6, *) 'ERROR: empty intJ range'
    error stop 1
  end if
  allocate(rvec(3, nr))
  r = 0
  do i1 = ix_intJ_min, ix_intJ_max
    do i2 = iy_intJ_min, iy_intJ_max
      do i3 = iz_intJ_min, iz_intJ_max
        r = r + 1
        rvec(:, r) = [i1, i2, i3]
      end do
    end do
  end do
  allocate(x(nwf, nwf, nr), j(nwf, nwf, nr, nen), heads(2 + 2 * nen))
  call rp_check(rp_j3d_run(sys%h, int(nr, c_int32_t), rvec, x, j, heads), 'rp_j3d_run')
  call rp_check(rp_finalize(sys%h), 'rp_finalize')
  write(6, '(a,f15.10)') ' chead(HL) in eV=', heads(1) * au
  allocate(xmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3)))
  allocate(jmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3), nen))
  xmat = (0.0d0, 0.0d0)
  jmat = (0.0d0, 0.0d0)
  do r = 1, nr
    xmat(:, :, rvec(1, r), rvec(2, r), rvec(3, r)) = x(:, :, r)
    jmat(:, :, rvec(1, r), rvec(2, r), rvec(3, r), :) = j(:, :, r, :)
  end do

  call make_dir('dir-intJ')
  call write_range('dir-intJ/dat.Xmat', '#Bare exchamge: X', &
                   '#1:i, 2:j, 3:Re(X_ij) [eV], 4:Im(X_ij) [eV]', xmat)
  call write_range('dir-intJ/dat.Jmat', '#Screened exchamge: J at omega=calc_ifreq', &
                   '#1:i, 2:j, 3:Re(J_ij) [eV], 4:Im(J_ij) [eV]', jmat(:, :, :, :, :, calc_ifreq))
  call write_j_vs_e()
  call make_dir('dir-mvmc')
  call write_mvmc_j(jmat(:, :, :, :, :, calc_ifreq), jcut_mvmc / au)
  write(6, *) 'calc_j3d: done'

contains

  ! dat.Xmat / dat.Jmat over the computed lattice vectors (calc_intJ.F:926-1000)
  subroutine write_range(path, title, cols, m)
    character(len=*), intent(in) :: path, title, cols
    complex(8), intent(in) :: m(:, :, -na(1):, -na(2):, -na(3):)
    integer :: uu, a1, a2, a3, i, jj
    open(newunit=uu, file=path)
    rewind(uu)
    write(uu, '(a)') title
    write(uu, '(a)') '#1:R1, 2:R2, 3:R3 (lattice vector)'
    write(uu, '(a)') cols
    do a1 = ix_intJ_min, ix_intJ_max
      do a2 = iy_intJ_min, iy_intJ_max
        do a3 = iz_intJ_min, iz_intJ_max
          write(uu, *) a1, a2, a3
          do i = 1, nwf
            do jj = 1, nwf
              write(uu, '(i5,i5,2f20.10)') i, jj, m(i, jj, a1, a2, a3) * au
            end do
          end do
          write(uu, *)
        end do
      end do
    end do
    close(uu)
  end subroutine

  ! dat.JvsE.III-JJJ (calc_intJ.F:1004-1022)
  subroutine write_j_vs_e()
    character(len=64) :: name
    integer :: uu, i, jj, k
    do i = 1, nwf
      do jj = i, nwf
        write(name, '("dat.JvsE.",i3.3,"-",i3.3)') i, jj
        open(newunit=uu, file='dir-intJ/' // trim(name))
        rewind(uu)
        write(uu, '(a)') '#Omega dependence of screened exchange J'
        write(uu, '(a)') '#Re(w) [eV], Im(w) [eV], Re[J(w)] [eV], Im[J(w)] [eV]'
        do k = 1, nen
          write(uu, '(4F12.5)') sys%em(k) * au, jmat(i, jj, 0, 0, 0, k) * au
        end do
        close(uu)
      end do
    end do
  end subroutine

  ! zHund.def, zExchange.def, zPairHop.def (wrt_mvmc_j3d.f90)
  subroutine write_mvmc_j(hr, cut)
    complex(8), intent(in) :: hr(:, :, -na(1):, -na(2):, -na(3):)
    real(8), intent(in) :: cut
    integer :: nsite, k, uu
    complex(8), allocatable :: h2(:, :, :, :)
    nsite = ntk
    allocate(h2(nwf, nwf, nsite, nsite))
    call write_mvmc_sites(nkb, na, nwf, hr, h2)
    k = count_pairs(h2, cut, .true.)
    open(newunit=uu, file='dir-mvmc/zHund.def')
    write(uu, '(a30)') '=============================='
    write(uu, '(a5,i25)') 'NHund', k
    write(uu, '(a30)') '=============================='
    write(uu, '(a10,f10.5,a10)') '=== Jcut >', cut * au, 'eV ======='
    write(uu, '(a30)') '=============================='
    call write_pairs(uu, h2, cut, .true.)
    close(uu)
    open(newunit=uu, file='dir-mvmc/zExchange.def')
    write(uu, '(a30)') '=============================='
    write(uu, '(a9,i21)') 'NExchange', k
    write(uu, '(a30)') '=============================='
    write(uu, '(a10,f10.5,a10)') '=== Jcut >', cut * au, 'eV ======='
    write(uu, '(a30)') '=============================='
    call write_pairs(uu, h2, cut, .true.)
    close(uu)
    k = count_pairs(h2, cut, .false.)
    open(newunit=uu, file='dir-mvmc/zPairHop.def')
    write(uu, '(a30)') '=============================='
    write(uu, '(a8,i22)') 'NPairhop', k
    write(uu, '(a30)') '=============================='
    write(uu, '(a10,f10.5,a10)') '=== Jcut >', cut * au, 'eV ======='
    write(uu, '(a30)') '=============================='
    call write_pairs(uu, h2, cut, .false.)
    close(uu)
  end subroutine

  ! site-orbital pairs (i, j) kept: i < j (upper) or i /= j (all), |H| >= cut
  integer function count_pairs(h2, cut, upper)
    complex(8), intent(in) :: h2(:, :, :, :)
    real(8), intent(in) :: cut
    logical, intent(in) :: upper
    integer :: isite, jsite, ib, jb, i, jj
    count_pairs = 0
    do isite = 1, size(h2, 3)
      do ib = 1, nwf
        i = ib + (isite - 1) * nwf
        do jsite = 1, size(h2, 4)
          do jb = 1, nwf
            jj = jb + (jsite - 1) * nwf
            if (upper .and. i >= jj) cycle
            if (.not. upper .and. i == jj) cycle
            if (abs(h2(ib, jb, isite, jsite)) < cut) cycle
            count_pairs = count_pairs + 1
          end do
        end do
      end do
    end do
  end function

  subroutine write_pairs(uu, h2, cut, upper)
    integer, intent(in) :: uu
    complex(8), intent(in) :: h2(:, :, :, :)
    real(8), intent(in) :: cut
    logical, intent(in) :: upper
    integer :: isite, jsite, ib, jb, i, jj
    do isite = 1, size(h2, 3)
      do ib = 1, nwf
        i = ib + (isite - 1) * nwf
        do jsite = 1, size(h2, 4)
          do jb = 1, nwf
            jj = jb + (jsite - 1) * nwf
            if (upper .and. i >= jj) cycle
            if (.not. upper .and. i == jj) cycle
            if (abs(h2(ib, jb, isite, jsite)) < cut) cycle
            write(uu, '(2i5,2f15.10)') i - 1, jj - 1, dble(h2(ib, jb, isite, jsite)) * au
          end do
        end do
      end do
    end do
  end subroutine

end program calc_j3d
