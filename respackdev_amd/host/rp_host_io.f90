! rp_host_io.f90 -- file side of the drop-in calc_chiqw host: the dir-wfn and
! dir-wan readers and the dir-eps writers, in the layouts RESPACK's chiqw reads
! and writes (chiqw.F:88-329, 870-934, 1100-1160, 1184-1218, 1494-1548,
! 2000-2085; chiqw_sub.F:440-454).
module rp_host_io
  use iso_c_binding
  use respack_gpu
  implicit none
  private
  public :: wfn_data, read_dir_wfn, read_dir_wan, fold_q, make_dir, file_exists
  public :: write_chi_cutoff, write_wgrid, write_sq, write_epsqw, write_eps_diag, write_log
  public :: reciprocal, int_system, load_int_system, write_mvmc_sites

  type :: wfn_data
    real(8) :: ecut_psi = 0, ef = 0, etot = 0
    real(8) :: a(3, 3) = 0                      ! a(:,i) = lattice vector i (bohr)
    integer :: nsym = 0, nnp = 0
    integer(c_int32_t), allocatable :: rg(:, :, :), pg(:, :)
    integer :: nk_irr = 0, ntg = 0, ntb = 0
    real(8), allocatable :: ski(:, :)           ! (3, nk_irr)
    integer(c_int32_t), allocatable :: ngi(:)   ! (nk_irr)
    integer(c_int32_t), allocatable :: kgi(:, :, :)  ! (3, ntg, nk_irr)
    real(8), allocatable :: eig(:, :)           ! (ntb, nk_irr), Ha
    complex(8), allocatable :: c0(:, :, :)      ! (ntg, ntb, nk_irr)
  end type

  ! calc_int inputs as loaded into the library (calc_intW.F:8-630)
  type :: int_system
    type(c_ptr) :: h = c_null_ptr
    integer :: nen = 0, ntk = 0, nwf = 0, nkb(3) = 0, na(3) = 0
    complex(8), allocatable :: em(:)
    real(8), allocatable :: coord(:, :)      ! (3, nwf) Wannier centres (bohr)
    real(8) :: a(3, 3) = 0
  end type

  interface
    integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
      import :: c_int, c_char
      character(kind=c_char) :: path(*)
      integer(c_int), value :: mode
    end function
  end interface

contains

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(6, '(a)') 'ERROR: ' // msg
    flush(6)
    error stop 1
  end subroutine

  logical function file_exists(path)
    character(len=*), intent(in) :: path
    inquire(file=path, exist=file_exists)
  end function

  ! mkdir -p for one level below an existing directory (no shell is spawned).
  subroutine make_dir(path)
    character(len=*), intent(in) :: path
    integer(c_int) :: rc
    logical :: there
    inquire(file=trim(path) // '/.', exist=there)
    if (there) return
    rc = c_mkdir(trim(path) // c_null_char, int(o'755', c_int))
    inquire(file=trim(path) // '/.', exist=there)
    if (.not. there) call fail('cannot create directory ' // trim(path))
  end subroutine

  ! dir-wfn (qe2respack / xtapp2respack output), read as chiqw.F:88-329 and
  ! chiqw.F:1100-1160 do.
  subroutine read_dir_wfn(w, with_bands)
    type(wfn_data), intent(out) :: w
    logical, intent(in), optional :: with_bands  ! .false.: no dat.eigenvalue / dat.wfn
    integer :: u, ik, ib, ig, i, j, iop, ng, ncomp
    open(newunit=u, file='./dir-wfn/dat.bandcalc', status='old', action='read')
    read(u, *) w%ecut_psi
    read(u, *) w%ef
    read(u, *) w%etot
    close(u)
    open(newunit=u, file='./dir-wfn/dat.lattice', status='old', action='read')
    do i = 1, 3
      read(u, *) w%a(:, i)
    end do
    close(u)
    open(newunit=u, file='./dir-wfn/dat.symmetry', status='old', action='read')
    read(u, *) w%nsym
    read(u, *) w%nnp
    allocate(w%rg(3, 3, w%nsym), w%pg(3, w%nsym))
    do iop = 1, w%nsym
      read(u, *) ((w%rg(i, j, iop), i = 1, 3), j = 1, 3)
      read(u, *) (w%pg(i, iop), i = 1, 3)
    end do
    close(u)
    open(newunit=u, file='./dir-wfn/dat.sample-k', status='old', action='read')
    read(u, *) w%nk_irr
    allocate(w%ski(3, w%nk_irr))
    do ik = 1, w%nk_irr
      read(u, *) w%ski(:, ik)
    end do
    close(u)
    allocate(w%ngi(w%nk_irr))
    open(newunit=u, file='./dir-wfn/dat.nkm', status='old', action='read')
    do ik = 1, w%nk_irr
      read(u, *) w%ngi(ik)
    end do
    close(u)
    w%ntg = maxval(w%ngi)
    allocate(w%kgi(3, w%ntg, w%nk_irr))
    w%kgi = 0
    open(newunit=u, file='./dir-wfn/dat.kg', status='old', action='read')
    do ik = 1, w%nk_irr
      read(u, *) ng
      if (ng /= w%ngi(ik)) call fail('NG_for_psi should be NGI(ik) (dat.kg vs dat.nkm)')
      do ig = 1, ng
        read(u, *) w%kgi(:, ig, ik)
      end do
    end do
    close(u)
    if (present(with_bands)) then
      if (.not. with_bands) return
    end if
    open(newunit=u, file='./dir-wfn/dat.eigenvalue', status='old', action='read')
    read(u, *) w%ntb
    allocate(w%eig(w%ntb, w%nk_irr))
    do ik = 1, w%nk_irr
      do ib = 1, w%ntb
        read(u, *) w%eig(ib, ik)
      end do
    end do
    close(u)
    allocate(w%c0(w%ntg, w%ntb, w%nk_irr))
    w%c0 = (0.0d0, 0.0d0)
    open(newunit=u, file='./dir-wfn/dat.wfn', status='old', action='read', form='unformatted')
    read(u) ncomp
    if (ncomp /= 1) call fail('ncomp /= 1 in dat.wfn is not supported')
    do ik = 1, w%nk_irr
      do ib = 1, w%ntb
        read(u) (w%c0(ig, ib, ik), ig = 1, w%ngi(ik))
      end do
    end do
    close(u)
  end subroutine

  ! dir-wan/dat.ns-nb and dat.umat for band cRPA (chiqw.F:870-895).
  subroutine read_dir_wan(ntk, ns, nb, mb, nwf, umat)
    integer, intent(in) :: ntk
    integer(c_int32_t), allocatable, intent(out) :: ns(:), nb(:)
    integer, intent(out) :: mb, nwf
    complex(8), allocatable, intent(out) :: umat(:, :, :)
    integer :: u, ik, jb, jw
    allocate(ns(ntk), nb(ntk))
    open(newunit=u, file='./dir-wan/dat.ns-nb', status='old', action='read')
    do ik = 1, ntk
      read(u, *) ns(ik), nb(ik)
    end do
    close(u)
    mb = maxval(nb)
    open(newunit=u, file='./dir-wan/dat.umat', status='old', action='read')
    read(u, *) nwf
    allocate(umat(mb, nwf, ntk))
    umat = (0.0d0, 0.0d0)
    do ik = 1, ntk
      do jb = 1, nb(ik)
        read(u, *) (umat(jb, jw, ik), jw = 1, nwf)
      end do
    end do
    close(u)
  end subroutine

  ! q folded into (-1/2, 1/2] per component as search_list (chiqw_sub.F:77-95).
  function fold_q(q) result(s)
    real(8), intent(in) :: q(3)
    real(8) :: s(3)
    real(8), parameter :: d = 1.0d-6
    integer :: i
    s = q
    do i = 1, 3
      if (s(i) > 1.5d0 + d) s(i) = s(i) - 2.0d0
      if (s(i) > 0.5d0 + d) s(i) = s(i) - 1.0d0
      if (s(i) <= -1.5d0 + d) s(i) = s(i) + 2.0d0
      if (s(i) <= -0.5d0 + d) s(i) = s(i) + 1.0d0
    end do
  end function

  ! b(:,i) = 2 pi (a_j x a_k) / volume
  subroutine reciprocal(a, b, volume)
    real(8), intent(in) :: a(3, 3)
    real(8), intent(out) :: b(3, 3), volume
    real(8), parameter :: tpi = 2.0d0 * 3.141592653589793238d0
    b(:, 1) = cross(a(:, 2), a(:, 3))
    b(:, 2) = cross(a(:, 3), a(:, 1))
    b(:, 3) = cross(a(:, 1), a(:, 2))
    volume = dot_product(a(:, 1), b(:, 1))
    b = b * (tpi / volume)
  contains
    function cross(x, y) result(z)
      real(8), intent(in) :: x(3), y(3)
      real(8) :: z(3)
      z = [x(2) * y(3) - x(3) * y(2), x(3) * y(1) - x(1) * y(3), x(1) * y(2) - x(2) * y(1)]
    end function
  end subroutine

  subroutine write_chi_cutoff(ecut_eps)
    real(8), intent(in) :: ecut_eps
    integer :: u
    open(newunit=u, file='./dir-eps/dat.chi_cutoff')
    write(u, *) ecut_eps
    close(u)
  end subroutine

  subroutine write_wgrid(em)
    complex(8), intent(in) :: em(:)
    integer :: u, ie
    open(newunit=u, file='./dir-eps/dat.wgrid')
    write(u, *) size(em)
    do ie = 1, size(em)
      write(u, '(2F15.10)') em(ie)
    end do
    close(u)
  end subroutine

  subroutine write_sq(dir, q)
    character(len=*), intent(in) :: dir
    real(8), intent(in) :: q(3)
    integer :: u
    open(newunit=u, file=dir // '/dat.sq')
    rewind(u)
    write(u, *) q(1), q(2), q(3)
    close(u)
  end subroutine

  ! one unformatted record (((eps(ig,jg,ie),ig),jg),ie): dat.epsqw.<num>
  subroutine write_epsqw(dir, num, eps)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: num
    complex(8), intent(in) :: eps(:, :, :)
    character(len=32) :: name
    integer :: u
    write(name, '("dat.epsqw.",i3.3)') num
    open(newunit=u, file=dir // '/' // trim(name), form='unformatted')
    rewind(u)
    write(u) eps
    close(u)
  end subroutine

  ! diagonal of eps^-1 per frequency: dat.epsqw.<base+ie> (i6.6), |q+G| and the
  ! diagonal element (calc_eps_rpa, chiqw_sub.F:440-454)
  subroutine write_eps_diag(dir, base, eps, lg, q, b)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: base
    complex(8), intent(in) :: eps(:, :, :)
    integer(c_int32_t), intent(in) :: lg(:, :)
    real(8), intent(in) :: q(3), b(3, 3)
    character(len=32) :: name
    integer :: u, ie, ig
    real(8) :: qg(3)
    do ie = 1, size(eps, 3)
      write(name, '("dat.epsqw.",i6.6)') base + ie
      open(newunit=u, file=dir // '/' // trim(name))
      rewind(u)
      write(u, '(a)') '#epsqw diagonal term'
      write(u, '(a)') '#1:q+G [au], 2:Re(eps-1), 3:Im(eps-1)'
      do ig = 1, size(eps, 1)
        qg = (q(1) + dble(lg(1, ig))) * b(:, 1) + (q(2) + dble(lg(2, ig))) * b(:, 2) &
             + (q(3) + dble(lg(3, ig))) * b(:, 3)
        write(u, '(3F20.10)') dsqrt(qg(1)**2 + qg(2)**2 + qg(3)**2), dble(eps(ig, ig, ie)), &
                              dimag(eps(ig, ig, ie))
      end do
      close(u)
    end do
  end subroutine

  ! dat.log.<num>: the completion flag that makes a rerun skip this q
  subroutine write_log(dir, num, qnum)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: num, qnum
    character(len=32) :: name
    integer :: u
    write(name, '("dat.log.",i3.3)') num
    open(newunit=u, file=dir // '/' // trim(name))
    rewind(u)
    write(u, *) '#'
    write(u, *) '#finish chi0 evaluation for', qnum
    write(u, *) '#'
    close(u)
  end subroutine


  ! dir-wfn (no bands), dir-eps (chi cutoff, frequencies, q list, eps^-1 per
  ! irreducible q, Gamma as the average of 600/601/602) and dir-wan, handed to
  ! the library (calc_intW.F:8-630, calc_intJ.F:8-600).
  subroutine load_int_system(sys)
    type(int_system), intent(inout) :: sys
    type(wfn_data) :: w
    integer(c_int32_t) :: iout(16)
    integer(c_int32_t), allocatable :: ng0(:), nge(:)
    real(8) :: dout(8), ecut_eps
    real(8), allocatable :: sk0(:, :), sqi(:, :)
    complex(8), allocatable :: wan(:, :, :), eps(:, :, :), e3(:, :, :)
    character(len=64) :: dirname
    character(len=256) :: line
    integer :: u, ik, iq, ie, ix, iw, ig, ng, nwf
    call read_dir_wfn(w, with_bands=.false.)
    sys%a = w%a
    open(newunit=u, file='./dir-eps/dat.chi_cutoff', status='old', action='read')
    read(u, *) ecut_eps
    close(u)
    open(newunit=u, file='./dir-eps/dat.wgrid', status='old', action='read')
    read(u, *) sys%nen
    allocate(sys%em(sys%nen))
    do ie = 1, sys%nen
      read(u, '(2f15.10)') sys%em(ie)
    end do
    close(u)
    allocate(sqi(3, w%nk_irr))
    do iq = 1, w%nk_irr
      write(dirname, "('dir-eps/q',i3.3)") iq
      if (.not. file_exists(trim(dirname) // '/dat.log.400')) &
        write(6, *) 'error: no dat.log.400 in ', trim(dirname)
      if (.not. file_exists(trim(dirname) // '/dat.sq')) call fail('no dat.sq in ' // trim(dirname))
      open(newunit=u, file=trim(dirname) // '/dat.sq', status='old', action='read')
      read(u, *) sqi(:, iq)
      close(u)
    end do
    call rp_check(rp_init(0_c_int32_t, sys%h), 'rp_init')
    call rp_check(rp_w3d_open(sys%h, w%a, int(w%nsym, c_int32_t), int(w%nnp, c_int32_t), w%rg, &
                              w%pg, int(w%nk_irr, c_int32_t), w%ski, w%ngi, int(w%ntg, c_int32_t), &
                              w%kgi, ecut_eps, int(sys%nen, c_int32_t), sys%em), 'rp_w3d_open')
    call rp_check(rp_w3d_info(sys%h, iout, dout), 'rp_w3d_info')
    sys%ntk = iout(1)
    sys%nkb = iout(2:4)
    sys%na = iout(5:7)
    write(6, '(a,i6,a,3i4,a,3i4,a,3i4)') ' NTK=', sys%ntk, ' nkb=', sys%nkb, ' Na=', sys%na, &
      ' FFT grid=', iout(8:10)
    allocate(sk0(3, sys%ntk), ng0(sys%ntk))
    call rp_check(rp_w3d_kmesh(sys%h, sk0, ng0), 'rp_w3d_kmesh')
    ! dat.wan: NWF, then per full-mesh k ((C0(G,w),G=1,NG0(k)),w=1,NWF)
    open(newunit=u, file='./dir-wan/dat.wan', status='old', action='read', form='unformatted')
    read(u) nwf
    sys%nwf = nwf
    allocate(wan(w%ntg, nwf, sys%ntk))
    wan = (0.0d0, 0.0d0)
    do ik = 1, sys%ntk
      read(u) ((wan(ig, iw, ik), ig = 1, ng0(ik)), iw = 1, nwf)
    end do
    close(u)
    allocate(sys%coord(3, nwf))
    open(newunit=u, file='./dir-wan/dat.wan-center', status='old', action='read')
    read(u, '(a)') line
    read(u, '(a)') line
    do iw = 1, nwf
      read(u, *) sys%coord(:, iw)
    end do
    close(u)
    call rp_check(rp_w3d_set_wannier(sys%h, int(nwf, c_int32_t), wan), 'rp_w3d_set_wannier')
    deallocate(wan)
    allocate(nge(w%nk_irr))
    call rp_check(rp_w3d_set_q(sys%h, int(w%nk_irr, c_int32_t), sqi, nge), 'rp_w3d_set_q')
    do iq = 1, w%nk_irr
      write(dirname, "('dir-eps/q',i3.3)") iq
      ng = nge(iq)
      if (allocated(eps)) deallocate(eps)
      allocate(eps(ng, ng, sys%nen))
      if (all(abs(w%ski(:, iq)) < 1.0d-5)) then
        allocate(e3(ng, ng, sys%nen))
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
      call rp_check(rp_w3d_set_epsinv(sys%h, int(iq - 1, c_int32_t), eps), 'rp_w3d_set_epsinv')
    end do
  end subroutine

  subroutine read_epsqw(dir, num, e)
    character(len=*), intent(in) :: dir
    integer, intent(in) :: num
    complex(8), intent(out) :: e(:, :, :)
    character(len=32) :: name
    integer :: uu
    write(name, '("dat.epsqw.",i3.3)') num
    if (.not. file_exists(dir // '/' // trim(name))) call fail('no ' // trim(name) // ' in ' // dir)
    open(newunit=uu, file=dir // '/' // trim(name), form='unformatted', status='old', action='read')
    read(uu) e
    close(uu)
  end subroutine

  ! H(R) on the lattice of NTK sites (a3 fastest), folded onto site pairs by the
  ! minimum-image vector (wrt_mvmc_w3d.f90, wrt_mvmc_j3d.f90): h2(:,:,is,js).
  subroutine write_mvmc_sites(nkb, na, nwf, hr, h2)
    integer, intent(in) :: nkb(3), na(3), nwf
    complex(8), intent(in) :: hr(:, :, -na(1):, -na(2):, -na(3):)
    complex(8), intent(out) :: h2(:, :, :, :)
    integer :: nsite, isite, jsite, d, lo, t(3), ia(3), ja(3)
    nsite = nkb(1) * nkb(2) * nkb(3)
    lo = 0
    if (mod(nsite, 2) == 0) lo = 1
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
  contains
    subroutine site_vec(is, a)
      integer, intent(in) :: is
      integer, intent(out) :: a(3)
      integer :: r
      r = is - 1
      a(3) = mod(r, nkb(3)) - na(3) + lo
      r = r / nkb(3)
      a(2) = mod(r, nkb(2)) - na(2) + lo
      a(1) = r / nkb(2) - na(1) + lo
    end subroutine
  end subroutine

end module rp_host_io
