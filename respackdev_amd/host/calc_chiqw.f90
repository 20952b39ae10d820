! calc_chiqw -- drop-in for RESPACK's calc_chiqw (src/chiqw, PROGRAM
! DIELECTRIC) with the per-q work on one MI355X through librespack_gpu.
!
! Same process contract as the reference: the namelist &param_chiqw (and the
! calc_num_k list) on standard input (m_rdinput.f90:41-128), ./dir-wfn (and
! ./dir-wan for band cRPA) in the working directory, and ./dir-eps written as
! chiqw.F:559-687 and 1184-2085 do: dat.chi_cutoff, dat.wgrid, and per q
! directory qNNN dat.sq, dat.epsqw.600 (600/601/602 at q = 0), the diagonal
! files dat.epsqw.6XXNNN and the completion flag dat.log.400 (a q whose flag
! exists is skipped, as in the reference).
!
! MPI_* namelist entries are accepted and ignored (one process drives one GPU).
! Contraction_precision is an addition: 0 = FP64 (default, the reference's
! arithmetic), 1 = FP32, 2 = split-BF16 (both checked to 1e-5 against FP64).
! Batch_q_graph is another: 1 runs every remaining q as one batch captured
! into a HIP graph (all eps^-1 held on the device until written), with the
! contraction autotuned per shape; 0 (default) runs the q one after another.
program calc_chiqw
  use iso_c_binding
  use respack_gpu
  use rp_host_io
  implicit none
  integer :: N_CALC_BAND = 0
  real(8) :: Ecut_for_eps = 0.0d0, shift_ef = 0.0d0, Max_excitation_energy = 200.0d0
  real(8) :: Green_func_delt = 0.1d0, ttrhdrn_dmna = 0.001d0, ttrhdrn_dmnr = 0.001d0
  integer :: MPI_io_rank = 0, MPI_num_proc_per_qcomm = 0, MPI_num_qcomm = 0
  integer :: Num_freq_grid = 70
  real(8) :: Lower_bound_energy_window = 0.0d0, Upper_bound_energy_window = 0.0d0
  integer :: flg_cRPA = 0, flg_calc_type = 0
  integer :: file_num_log_start = 400, file_num_chi_start = 500, file_num_eps_start = 600
  integer :: file_num_chi_base_start = 500000, file_num_eps_base_start = 600000
  integer :: n_calc_q = 0
  integer :: Contraction_precision = 0
  integer :: Batch_q_graph = 0
  namelist /param_chiqw/ N_CALC_BAND, Ecut_for_eps, shift_ef, Max_excitation_energy, &
    Green_func_delt, ttrhdrn_dmna, ttrhdrn_dmnr, MPI_io_rank, MPI_num_proc_per_qcomm, &
    MPI_num_qcomm, Num_freq_grid, Lower_bound_energy_window, Upper_bound_energy_window, &
    flg_cRPA, flg_calc_type, file_num_log_start, file_num_chi_start, file_num_eps_start, &
    file_num_chi_base_start, file_num_eps_base_start, n_calc_q, Contraction_precision, &
    Batch_q_graph

  type(wfn_data) :: w
  type(c_ptr) :: h
  type(rp_chiqw_params) :: prm
  integer, allocatable :: calc_num_k(:), qlist(:)
  integer(c_int32_t) :: iout(16), ng, ngc
  integer(c_int32_t), allocatable :: ns(:), nb(:)
  integer(c_int32_t), allocatable, target :: lg(:, :)
  real(8) :: dout(8), q(3), b(3, 3), volume, timings(8), tbatch
  real(8), allocatable, target :: qall(:, :)
  complex(8), allocatable :: umat(:, :, :)
  complex(8), allocatable, target :: em(:)
  complex(8), allocatable, target :: eps(:, :, :, :)
  character(len=64) :: dirname
  integer :: ik, jk, iq, nq, ntk, nen, ndir, ix, mb, nwf, qnum

  read(5, nml=param_chiqw)
  if (flg_calc_type == 2 .and. n_calc_q == 0) then
    write(6, *) 'ERROR: n_calc_q should not be zero'
    error stop 1
  end if
  allocate(calc_num_k(max(n_calc_q, 0)))
  if (n_calc_q /= 0) read(5, *) (calc_num_k(ix), ix = 1, n_calc_q)
  write(6, nml=param_chiqw)
  if (flg_cRPA == 0) then
    write(6, *) 'THIS CALCULATION IS fRPA'
  else
    write(6, *) 'THIS CALCULATION IS cRPA'
  end if

  call read_dir_wfn(w)
  call reciprocal(w%a, b, volume)
  write(6, '(a,i6,a,i6,a,i6)') ' Nk_irr=', w%nk_irr, ' NTB=', w%ntb, ' NTG=', w%ntg

  ! dir-eps/qNNN and the list of q still to do (chiqw.F:559-650); all file
  ! system work happens before the GPU is touched
  call make_dir('dir-eps')
  do ik = 1, w%nk_irr
    write(dirname, "('dir-eps/q',i3.3)") ik
    call make_dir(trim(dirname))
  end do
  allocate(qlist(w%nk_irr + max(n_calc_q, 0)))
  nq = 0
  select case (flg_calc_type)
  case (0)
    do ik = 1, w%nk_irr
      call add_q(ik)
    end do
  case (1)
    jk = 0
    do ik = 1, w%nk_irr
      if (all(abs(w%ski(:, ik)) < 1.0d-5)) jk = ik
    end do
    if (jk == 0) then
      write(6, *) 'ERROR: no Gamma point in dat.sample-k'
      error stop 1
    end if
    call add_q(jk)
  case (2)
    do ix = 1, n_calc_q
      if (calc_num_k(ix) < 1 .or. calc_num_k(ix) > w%nk_irr) then
        write(6, *) 'ERROR: calc_num_k out of range:', calc_num_k(ix)
        error stop 1
      end if
      call add_q(calc_num_k(ix))
    end do
  case default
    write(6, *) 'ERROR: flg_calc_type must be 0, 1 or 2'
    error stop 1
  end select
  write(6, *) 'Nq_irr=', nq

  prm%n_calc_band = N_CALC_BAND
  prm%num_freq_grid = Num_freq_grid
  prm%flg_crpa = flg_cRPA
  prm%precision = Contraction_precision
  prm%ecut_for_eps = Ecut_for_eps
  prm%shift_ef = shift_ef
  prm%max_excitation_energy = Max_excitation_energy
  prm%green_func_delt = Green_func_delt
  prm%ttrhdrn_dmna = ttrhdrn_dmna
  prm%ttrhdrn_dmnr = ttrhdrn_dmnr

  call rp_check(rp_init(0_c_int32_t, h), 'rp_init')
  call rp_check(rp_chiqw_open(h, w%a, int(w%nsym, c_int32_t), int(w%nnp, c_int32_t), w%rg, w%pg, &
                              int(w%nk_irr, c_int32_t), w%ski, w%ngi, int(w%ntg, c_int32_t), &
                              w%kgi, int(w%ntb, c_int32_t), w%eig, w%c0, w%ef, prm), &
                'rp_chiqw_open')
  deallocate(w%c0)
  call rp_check(rp_chiqw_info(h, iout, dout, c_null_ptr), 'rp_chiqw_info')
  ntk = iout(1)
  nen = iout(8)
  allocate(em(nen))
  call rp_check(rp_chiqw_info(h, iout, dout, c_loc(em)), 'rp_chiqw_info')
  write(6, '(a,i6,a,3i4,a,3i4)') ' NTK=', ntk, ' nkb=', iout(2:4), ' FFT grid=', iout(5:7)
  write(6, '(a,f12.6,a,f12.6)') ' Ecut_for_psi (Ry)=', dout(1), '  Ecut_for_eps (Ry)=', dout(2)
  if (flg_cRPA == 1) then
    call read_dir_wan(ntk, ns, nb, mb, nwf, umat)
    call rp_check(rp_chiqw_set_crpa(h, int(ntk, c_int32_t), ns, nb, int(mb, c_int32_t), &
                                    int(nwf, c_int32_t), umat), 'rp_chiqw_set_crpa')
  end if
  call write_chi_cutoff(dout(2))
  call write_wgrid(em)

  if (Batch_q_graph == 1 .and. nq > 0) then
    ! every remaining q in one captured launch sequence, replayed once
    allocate(qall(3, nq))
    do iq = 1, nq
      qall(:, iq) = fold_q(w%ski(:, qlist(iq)))
    end do
    call rp_check(rp_chiqw_autotune(h, 1_c_int32_t), 'rp_chiqw_autotune')
    call rp_check(rp_chiqw_batch_plan(h, int(nq, c_int32_t), c_loc(qall), c_null_ptr, c_null_ptr, &
                                      int(Contraction_precision, c_int32_t)), 'rp_chiqw_batch_plan')
    call rp_check(rp_chiqw_batch_run(h, tbatch), 'rp_chiqw_batch_run')
    write(6, '(a,i6,a,f10.4,a)') ' batch of', nq, ' q replayed from one graph:', tbatch, ' s'
  end if
  do iq = 1, nq
    q = fold_q(w%ski(:, qlist(iq)))
    qnum = 0
    do ik = 1, w%nk_irr
      if (all(abs(fold_q(w%ski(:, ik)) - q) < 1.0d-6)) qnum = ik
    end do
    write(dirname, "('dir-eps/q',i3.3)") qnum
    if (Batch_q_graph == 1) then
      call rp_check(rp_chiqw_batch_result(h, int(iq - 1, c_int32_t), ng, ndir, c_null_ptr, 0_c_int32_t, &
                                          c_null_ptr), 'rp_chiqw_batch_result')
      if (allocated(lg)) deallocate(lg)
      allocate(lg(3, ng))
      if (allocated(eps)) deallocate(eps)
      allocate(eps(ng, ng, nen, ndir))
      call rp_check(rp_chiqw_batch_result(h, int(iq - 1, c_int32_t), ngc, ndir, c_loc(lg), ng, &
                                          c_loc(eps)), 'rp_chiqw_batch_result')
      write(6, '(a,i4,a,3f10.5,a,i6)') ' q', qnum, ' =', q, '  NG=', ng
    else
      call rp_check(rp_chiqw_prepare_q(h, q, ng, c_null_ptr, 0_c_int32_t), 'rp_chiqw_prepare_q')
      if (allocated(lg)) deallocate(lg)
      allocate(lg(3, ng))
      call rp_check(rp_chiqw_prepare_q(h, q, ngc, c_loc(lg), ng), 'rp_chiqw_prepare_q')
      ndir = 1
      if (q(1) == 0.0d0 .and. q(2) == 0.0d0 .and. q(3) == 0.0d0) ndir = 3
      if (allocated(eps)) deallocate(eps)
      allocate(eps(ng, ng, nen, ndir))
      call rp_check(rp_chiqw_q(h, q, int(Contraction_precision, c_int32_t), c_loc(eps), c_null_ptr, &
                               timings), 'rp_chiqw_q')
      write(6, '(a,i4,a,3f10.5,a,i6,a,f10.4,a)') ' q', qnum, ' =', q, '  NG=', ng, '  device ', &
        timings(5), ' s'
    end if
    call write_sq(trim(dirname), q)
    do ix = 1, ndir
      call write_eps_diag(trim(dirname), file_num_eps_base_start + (ix - 1) * 1000, &
                          eps(:, :, :, ix), lg, q, b)
      call write_epsqw(trim(dirname), file_num_eps_start + ix - 1, eps(:, :, :, ix))
    end do
    call write_log(trim(dirname), file_num_log_start, qnum)
  end do
  call rp_check(rp_finalize(h), 'rp_finalize')
  write(6, *) 'calc_chiqw: done'

contains

  subroutine add_q(k)
    integer, intent(in) :: k
    character(len=64) :: d
    write(d, "('dir-eps/q',i3.3)") k
    if (file_exists(trim(d) // '/dat.log.400')) then
      write(6, *) 'dat.log.400 exists in ', trim(d)
    else
      nq = nq + 1
      qlist(nq) = k
    end if
  end subroutine

end program calc_chiqw
