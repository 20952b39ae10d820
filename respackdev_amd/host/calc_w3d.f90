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

  type(wfn_data) :: w
  type(c_ptr) :: h
  integer(c_int32_t) :: iout(16), nwf32
  integer(c_int32_t), allocatable :: ng0(:), nge(:)
  real(8) :: dout(8), b(3, 3), volume, ecut_eps
  real(8), allocatable :: sk0(:, :), sqi(:, :), heads(:), coord(:, :)
  complex(8), allocatable :: em(:), wan(:, :, :), eps(:, :, :), e3(:, :, :)
  complex(8), allocatable :: vmat(:, :, :, :, :), wmat(:, :, :, :, :, :)
  character(len=64) :: dirname
  character(len=256) :: line
  integer :: u, ik, iq, ie, nen, ntk, nwf, na(3), nkb(3), ix, iw, ig, ng

  read(5, nml=param_calc_int)
  write(6, nml=param_calc_int)
  call read_dir_wfn(w, with_bands=.false.)
  call reciprocal(w%a, b, volume)
  ! dir-eps/dat.chi_cutoff, dat.wgrid (calc_intW.F:393-410), read as the reference does
  open(newunit=u, file='./dir-eps/dat.chi_cutoff', status='old', action='read')
  read(u, *) ecut_eps
  close(u)
  open(newunit=u, file='./dir-eps/dat.wgrid', status='old', action='read')
  read(u, *) nen
  allocate(em(nen))
  do ie = 1, nen
    read(u, '(2f15.10)') em(ie)
  end do
  close(u)
  if (calc_ifreq < 1 .or. calc_ifreq > nen) then
    write(6, *) 'ERROR: calc_ifreq out of range 1..', nen
    error stop 1
  end if
  ! irreducible q of dir-eps (calc_intW.F:411-433)
  allocate(sqi(3, w%nk_irr))
  do iq = 1, w%nk_irr
    write(dirname, "('dir-eps/q',i3.3)") iq
    if (.not. file_exists(trim(dirname) // '/dat.log.400')) &
      write(6, *) 'error: no dat.log.400 in ', trim(dirname)
    if (.not. file_exists(trim(dirname) // '/dat.sq')) then
      write(6, *) 'ERROR: no dat.sq in ', trim(dirname)
      error stop 1
    end if
    open(newunit=u, file=trim(dirname) // '/dat.sq', status='old', action='read')
    read(u, *) sqi(:, iq)
    close(u)
  end do
  ! dir-wan/dat.wan-center
  call rp_check(rp_init(0_c_int32_t, h), 'rp_init')
  call rp_check(rp_w3d_open(h, w%a, int(w%nsym, c_int32_t), int(w%nnp, c_int32_t), w%rg, w%pg, &
                            int(w%nk_irr, c_int32_t), w%ski, w%ngi, int(w%ntg, c_int32_t), w%kgi, &
                            ecut_eps, int(nen, c_int32_t), em), 'rp_w3d_open')
  call rp_check(rp_w3d_info(h, iout, dout), 'rp_w3d_info')
  ntk = iout(1)
  nkb = iout(2:4)
  na = iout(5:7)
  write(6, '(a,i6,a,3i4,a,3i4,a,3i4)') ' NTK=', ntk, ' nkb=', nkb, ' Na=', na, ' FFT grid=', iout(8:10)
  allocate(sk0(3, ntk), ng0(ntk))
  call rp_check(rp_w3d_kmesh(h, sk0, ng0), 'rp_w3d_kmesh')
  ! dir-wan/dat.wan: NWF, then per full-mesh k ((C0(G,w),G=1,NG0(k)),w=1,NWF)
  open(newunit=u, file='./dir-wan/dat.wan', status='old', action='read', form='unformatted')
  read(u) nwf
  allocate(wan(w%ntg, nwf, ntk))
  wan = (0.0d0, 0.0d0)
  do ik = 1, ntk
    read(u) ((wan(ig, iw, ik), ig = 1, ng0(ik)), iw = 1, nwf)
  end do
  close(u)
  allocate(coord(3, nwf))
  open(newunit=u, file='./dir-wan/dat.wan-center', status='old', action='read')
  read(u, '(a)') line
  read(u, '(a)') line
  do iw = 1, nwf
    read(u, *) coord(:, iw)
  end do
  close(u)
  nwf32 = nwf
  call rp_check(rp_w3d_set_wannier(h, nwf32, wan), 'rp_w3d_set_wannier')
  deallocate(wan)
  allocate(nge(w%nk_irr))
  call rp_check(rp_w3d_set_q(h, int(w%nk_irr, c_int32_t), sqi, nge), 'rp_w3d_set_q')
  ! eps^-1 per irreducible q; Gamma (by dat.sample-k) averages 600/601/602 in
  ! double precision (calc_intW.F:590-630)
  do iq = 1, w%nk_irr
    write(dirname, "('dir-eps/q',i3.3)") iq
    ng = nge(iq)
    if (allocated(eps)) deallocate(eps)
    allocate(eps(ng, ng, nen))
    if (all(abs(w%ski(:, iq)) < 1.0d-5)) then
      allocate(e3(ng, ng, nen))
      eps = (0.0d0, 0.0d0)
      do ix = 1, 3
        call read_epsqw(trim(dirname), 600 + ix - 1, e3)
        eps = eps + e3
      end do
      eps = eps / 3.0d0
      deallocate(e3)
    else
      call read_epsqw(trim(dirname), 600, eps)
    end if
    call rp_check(rp_w3d_set_epsinv(h, int(iq - 1, c_int32_t), eps), 'rp_w3d_set_epsinv')
  end do
  allocate(vmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3)))
  allocate(wmat(nwf, nwf, -na(1):na(1), -na(2):na(2), -na(3):na(3), nen))
  allocate(heads(2 + 2 * nen))
  call rp_check(rp_w3d_vw(h, vmat, wmat, heads), 'rp_w3d_vw')
  call rp_check(rp_finalize(h), 'rp_finalize')
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

  subroutine read_epsqw(dir, num, e)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: num
    complex(8), intent(out) :: e(:, :, :)
    character(len=32) :: name
    integer :: uu
    write(name, '("dat.epsqw.",i3.3)') num
    if (.not. file_exists(dir // '/' // trim(name))) then
      write(6, *) 'ERROR: no ', trim(name), ' in ', dir
      error stop 1
    end if
    open(newunit=uu, file=dir // '/' // trim(name), form='unformatted', status='old', action='read')
    read(uu) e
    close(uu)
  end subroutine

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
              rij = dble(i1) * w%a(:, 1) + dble(i2) * w%a(:, 2) + dble(i3) * w%a(:, 3) &
                    + (coord(:, j) - coord(:, i))
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
          write(uu, '(4F12.5)') em(k) * au, wmat(i, j, 0, 0, 0, k) * au
        end do
        close(uu)
      end do
    end do
  end subroutine

  ! zCoulombIntra.def / zCoulombInter.def (wrt_mvmc_w3d.f90): the lattice of
  ! NTK sites, W(R) folded onto site pairs by the minimum-image vector.
  subroutine write_mvmc(hr, cut)
    complex(8), intent(in) :: hr(:, :, -na(1):, -na(2):, -na(3):)
    real(8), intent(in) :: cut
    integer :: nsite, isite, jsite, ib, jb, i, j, k, uu, t(3), ia(3), ja(3), lo, d
    complex(8), allocatable :: h2(:, :, :, :)
    logical :: even
    nsite = ntk
    allocate(h2(nwf, nwf, nsite, nsite))
    even = mod(nsite, 2) == 0
    lo = 0
    if (even) lo = 1
    ! sites (a1,a2,a3) in [-Na+lo, Na]^3, a3 fastest
    do isite = 1, nsite
      call site_vec(isite, ia)
      do jsite = 1, nsite
        call site_vec(jsite, ja)
        do d = 1, 3
          t(d) = ja(d) - ia(d)
          if (t(d) < -na(d) + lo) t(d) = t(d) + nkb(d)
          if (t(d) > na(d)) t(d) = t(d) - nkb(d)
        end do
        h2(:, :, isite, jsite) = hr(:, :, t(1), t(2), t(3))
      end do
    end do
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

  ! site index (1-based, a3 fastest) -> lattice vector of the site
  subroutine site_vec(isite, a)
    integer, intent(in) :: isite
    integer, intent(out) :: a(3)
    integer :: r, lo
    lo = 0
    if (mod(ntk, 2) == 0) lo = 1
    r = isite - 1
    a(3) = mod(r, nkb(3)) - na(3) + lo
    r = r / nkb(3)
    a(2) = mod(r, nkb(2)) - na(2) + lo
    a(1) = r / nkb(2) - na(1) + lo
  end subroutine

end program calc_w3d
