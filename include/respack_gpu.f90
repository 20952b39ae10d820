! respack_gpu.f90 -- Fortran (ISO_C_BINDING) interface to librespack_gpu.so.
!
! One interface per entry point of respack_gpu.h, with the same argument order
! and meaning.  Arrays are passed as the Fortran arrays themselves (column-major,
! complex(8) = two adjacent doubles), so a RESPACK-side caller hands over its
! own arrays unchanged: e.g. C0(NTG,NTB,Nk_irr), E_EIGI(NTB,Nk_irr),
! epsqw(NG,NG,nen).  Every function returns 0 on success; rp_check turns a
! non-zero status into an error stop carrying rp_last_error().
module respack_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: RP_PREC_FP64 = 0
  integer(c_int32_t), parameter, public :: RP_PREC_FP32 = 1
  integer(c_int32_t), parameter, public :: RP_PREC_BF16X3 = 2

  ! &param_chiqw values the library uses (m_rdinput.f90:41-72)
  type, bind(C), public :: rp_chiqw_params
    integer(c_int32_t) :: n_calc_band = 0
    integer(c_int32_t) :: num_freq_grid = 70
    integer(c_int32_t) :: flg_crpa = 0
    integer(c_int32_t) :: precision = RP_PREC_FP64
    real(c_double) :: ecut_for_eps = 0.0d0
    real(c_double) :: shift_ef = 0.0d0
    real(c_double) :: max_excitation_energy = 200.0d0
    real(c_double) :: green_func_delt = 0.1d0
    real(c_double) :: ttrhdrn_dmna = 0.001d0
    real(c_double) :: ttrhdrn_dmnr = 0.001d0
  end type

  public :: rp_init, rp_finalize, rp_last_error_string, rp_check
  public :: rp_algn235, rp_gen_grid, rp_make_kg0, rp_ttrhdrn_mkidx
  public :: rp_chiqw_open, rp_chiqw_set_crpa, rp_chiqw_info, rp_chiqw_kmesh
  public :: rp_chiqw_prepare_q, rp_pairdens_batch, rp_tetra_weights_batch
  public :: rp_chi_accumulate, rp_dyson, rp_dyson_noalloc, rp_chiqw_q, rp_chiqw_npair, rp_chiqw_set_window
  public :: rp_chiqw_batch_plan, rp_chiqw_batch_run, rp_chiqw_batch_result, rp_chiqw_batch_free
  public :: rp_chiqw_autotune, rp_chiqw_tune_count, rp_chiqw_tune_info, rp_chiqw_set_span
  public :: rp_mlp_open, rp_mlp_nparams, rp_mlp_get_params, rp_mlp_set_params, rp_mlp_forward
  public :: rp_mlp_gradient, rp_mlp_train, rp_chi_set
  public :: rp_w3d_open, rp_w3d_info, rp_w3d_kmesh, rp_w3d_set_wannier, rp_w3d_set_q
  public :: rp_w3d_set_epsinv, rp_w3d_funcs, rp_w3d_vw, rp_j3d_run, rp_w3d_stage_times

  interface
    integer(c_int) function rp_init(device, h) bind(C, name="rp_init")
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: device
      type(c_ptr) :: h
    end function
    integer(c_int) function rp_finalize(h) bind(C, name="rp_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function rp_last_error_c() bind(C, name="rp_last_error")
      import :: c_ptr
    end function
    integer(c_int32_t) function rp_algn235(n) bind(C, name="rp_algn235")
      import :: c_int32_t
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function rp_gen_grid(nt, emax, em) bind(C, name="rp_gen_grid")
      import :: c_int, c_int32_t, c_double, c_double_complex
      integer(c_int32_t), value :: nt
      real(c_double), value :: emax
      complex(c_double_complex) :: em(*)
    end function
    integer(c_int) function rp_make_kg0(ntg, b, gcut, q, kg, ng) bind(C, name="rp_make_kg0")
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: ntg
      real(c_double) :: b(3, 3)
      real(c_double), value :: gcut
      real(c_double) :: q(3)
      integer(c_int32_t) :: kg(3, *)
      integer(c_int32_t) :: ng
    end function
    integer(c_int) function rp_ttrhdrn_mkidx(n1, n2, n3, b, idx) bind(C, name="rp_ttrhdrn_mkidx")
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n1, n2, n3
      real(c_double) :: b(3, 3)
      integer(c_int32_t) :: idx(*)
    end function
    integer(c_int) function rp_chiqw_open(h, lattice, nsym, nnp, rg, pg, nk_irr, ski, ngi, ntg, &
                                          kgi, ntb, eig, wfn, ef, prm) bind(C, name="rp_chiqw_open")
      import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr, rp_chiqw_params
      type(c_ptr), value :: h
      real(c_double) :: lattice(3, 3)
      integer(c_int32_t), value :: nsym, nnp
      integer(c_int32_t) :: rg(3, 3, *), pg(3, *)
      integer(c_int32_t), value :: nk_irr
      real(c_double) :: ski(3, *)
      integer(c_int32_t) :: ngi(*)
      integer(c_int32_t), value :: ntg
      integer(c_int32_t) :: kgi(3, ntg, *)
      integer(c_int32_t), value :: ntb
      real(c_double) :: eig(ntb, *)
      complex(c_double_complex) :: wfn(ntg, ntb, *)
      real(c_double), value :: ef
      type(rp_chiqw_params) :: prm
    end function
    integer(c_int) function rp_chiqw_set_crpa(h, ntk, ns, nb, mb, nwf, umat) &
        bind(C, name="rp_chiqw_set_crpa")
      import :: c_int, c_int32_t, c_double_complex, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: ntk
      integer(c_int32_t) :: ns(*), nb(*)
      integer(c_int32_t), value :: mb, nwf
      complex(c_double_complex) :: umat(mb, nwf, *)
    end function
    ! em: c_loc of a complex(8) array of nen elements, or c_null_ptr
    integer(c_int) function rp_chiqw_info(h, iout, dout, em) bind(C, name="rp_chiqw_info")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t) :: iout(16)
      real(c_double) :: dout(8)
      type(c_ptr), value :: em
    end function
    integer(c_int) function rp_chiqw_kmesh(h, sk0, numirr, trs) bind(C, name="rp_chiqw_kmesh")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: sk0(3, *)
      integer(c_int32_t) :: numirr(*), trs(*)
    end function
    ! lg: c_loc of an integer(c_int32_t) array (3, lg_cap), or c_null_ptr
    integer(c_int) function rp_chiqw_prepare_q(h, q, ng, lg, lg_cap) bind(C, name="rp_chiqw_prepare_q")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: q(3)
      integer(c_int32_t) :: ng
      type(c_ptr), value :: lg
      integer(c_int32_t), value :: lg_cap
    end function
    integer(c_int) function rp_pairdens_batch(h, m_out) bind(C, name="rp_pairdens_batch")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: m_out
    end function
    integer(c_int) function rp_tetra_weights_batch(h, d_out) bind(C, name="rp_tetra_weights_batch")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: d_out
    end function
    integer(c_int) function rp_chi_accumulate(h, precision, chi_out) bind(C, name="rp_chi_accumulate")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: chi_out
    end function
    integer(c_int) function rp_dyson(h, epsinv_out) bind(C, name="rp_dyson")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: epsinv_out
    end function
    ! eps^-1 from chi, both in device memory, with no allocation (see respack_gpu.h)
    integer(c_int) function rp_dyson_noalloc(h, chi_dev, epsinv_dev) bind(C, name="rp_dyson_noalloc")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: chi_dev
      type(c_ptr), value :: epsinv_dev
    end function
    integer(c_int) function rp_chiqw_q(h, q, precision, epsinv_out, chi_out, timings) &
        bind(C, name="rp_chiqw_q")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: q(3)
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: epsinv_out, chi_out
      real(c_double) :: timings(8)
    end function
    integer(c_int) function rp_chiqw_npair(h, npair) bind(C, name="rp_chiqw_npair")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t) :: npair
    end function
    integer(c_int) function rp_chiqw_set_window(h, ecut_for_eps, n_vir) bind(C, name="rp_chiqw_set_window")
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), value :: ecut_for_eps
      integer(c_int32_t), value :: n_vir
    end function
    ! batched configuration under graph capture; contraction autotuner (no
    ! counterpart in the reference)
    integer(c_int) function rp_chiqw_batch_plan(h, nq, qs, ecut_for_eps, n_vir, precision) &
        bind(C, name="rp_chiqw_batch_plan")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nq, precision
      type(c_ptr), value :: qs, ecut_for_eps, n_vir
    end function
    integer(c_int) function rp_chiqw_batch_run(h, seconds) bind(C, name="rp_chiqw_batch_run")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double) :: seconds
    end function
    integer(c_int) function rp_chiqw_batch_result(h, iq, ng, ndir, lg, lg_cap, epsinv_out) &
        bind(C, name="rp_chiqw_batch_result")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: iq, lg_cap
      integer(c_int32_t) :: ng, ndir
      type(c_ptr), value :: lg, epsinv_out
    end function
    integer(c_int) function rp_chiqw_batch_free(h) bind(C, name="rp_chiqw_batch_free")
      import :: c_ptr, c_int
      type(c_ptr), value :: h
    end function
    integer(c_int) function rp_chiqw_set_span(h, span) bind(C, name="rp_chiqw_set_span")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: span
    end function
    integer(c_int) function rp_chiqw_autotune(h, enable) bind(C, name="rp_chiqw_autotune")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function rp_chiqw_tune_count(h, n) bind(C, name="rp_chiqw_tune_count")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: n
    end function
    integer(c_int) function rp_chiqw_tune_info(h, idx, iout, dout) bind(C, name="rp_chiqw_tune_info")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: idx
      integer(c_int32_t) :: iout(10)
      real(c_double) :: dout(2)
    end function
    ! calc_w3d (calc_intW.F)
    integer(c_int) function rp_w3d_open(h, lattice, nsym, nnp, rg, pg, nk_irr, ski, ngi, ntg, &
                                        kgi, ecut_for_eps, nen, em) bind(C, name="rp_w3d_open")
      import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: lattice(3, 3)
      integer(c_int32_t), value :: nsym, nnp
      integer(c_int32_t) :: rg(3, 3, *), pg(3, *)
      integer(c_int32_t), value :: nk_irr
      real(c_double) :: ski(3, *)
      integer(c_int32_t) :: ngi(*)
      integer(c_int32_t), value :: ntg
      integer(c_int32_t) :: kgi(3, ntg, *)
      real(c_double), value :: ecut_for_eps
      integer(c_int32_t), value :: nen
      complex(c_double_complex) :: em(*)
    end function
    integer(c_int) function rp_w3d_info(h, iout, dout) bind(C, name="rp_w3d_info")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t) :: iout(16)
      real(c_double) :: dout(8)
    end function
    integer(c_int) function rp_w3d_kmesh(h, sk0, ng0) bind(C, name="rp_w3d_kmesh")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: sk0(3, *)
      integer(c_int32_t) :: ng0(*)
    end function
    integer(c_int) function rp_w3d_set_wannier(h, nwf, wan) bind(C, name="rp_w3d_set_wannier")
      import :: c_int, c_int32_t, c_double_complex, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nwf
      complex(c_double_complex) :: wan(*)
    end function
    integer(c_int) function rp_w3d_set_q(h, nq_irr, sqi, ng_eps) bind(C, name="rp_w3d_set_q")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nq_irr
      real(c_double) :: sqi(3, *)
      integer(c_int32_t) :: ng_eps(*)
    end function
    integer(c_int) function rp_w3d_set_epsinv(h, iq, epsinv) bind(C, name="rp_w3d_set_epsinv")
      import :: c_int, c_int32_t, c_double_complex, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: iq
      complex(c_double_complex) :: epsinv(*)
    end function
    integer(c_int) function rp_w3d_funcs(h, func_out, funcw_out) bind(C, name="rp_w3d_funcs")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      type(c_ptr), value :: func_out, funcw_out
    end function
    integer(c_int) function rp_w3d_vw(h, vmat, wmat, heads) bind(C, name="rp_w3d_vw")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      complex(c_double_complex) :: vmat(*), wmat(*)
      real(c_double) :: heads(*)
    end function
    ! calc_j3d (calc_intJ.F) on the opened system
    integer(c_int) function rp_j3d_run(h, nr, rvec, xmat, jmat, heads) bind(C, name="rp_j3d_run")
      import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nr
      integer(c_int32_t) :: rvec(3, *)
      complex(c_double_complex) :: xmat(*), jmat(*)
      real(c_double) :: heads(*)
    end function
    ! device seconds of the last rp_w3d_funcs and rp_j3d_run (see respack_gpu.h)
    integer(c_int) function rp_w3d_stage_times(h, t) bind(C, name="rp_w3d_stage_times")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double) :: t(8)
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function rp_mlp_open(h, nin, width, depth, nout, seed) bind(C, name="rp_mlp_open")
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nin, width, depth, nout
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function rp_mlp_nparams(h, n) bind(C, name="rp_mlp_nparams")
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: h
      integer(c_int64_t) :: n
    end function
    integer(c_int) function rp_mlp_get_params(h, params) bind(C, name="rp_mlp_get_params")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float) :: params(*)
    end function
    integer(c_int) function rp_mlp_set_params(h, params) bind(C, name="rp_mlp_set_params")
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float) :: params(*)
    end function
    integer(c_int) function rp_mlp_forward(h, n, x, y) bind(C, name="rp_mlp_forward")
      import :: c_ptr, c_int, c_int64_t, c_float
      type(c_ptr), value :: h
      integer(c_int64_t), value :: n
      real(c_float) :: x(*), y(*)
    end function
    integer(c_int) function rp_mlp_gradient(h, n, x, t, grad, loss) bind(C, name="rp_mlp_gradient")
      import :: c_ptr, c_int, c_int64_t, c_float, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: n
      real(c_float) :: x(*), t(*), grad(*)
      real(c_double) :: loss
    end function
    integer(c_int) function rp_mlp_train(h, n, x, t, steps, batch, lr, seed, losses) bind(C, name="rp_mlp_train")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_float, c_double
      type(c_ptr), value :: h
      integer(c_int64_t), value :: n, seed
      real(c_float) :: x(*), t(*)
      integer(c_int32_t), value :: steps, batch
      real(c_double), value :: lr
      real(c_double) :: losses(*)
    end function
    integer(c_int) function rp_chi_set(h, chi) bind(C, name="rp_chi_set")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double) :: chi(*)
    end function
  end interface

contains

  ! rp_last_error() as a Fortran string
  function rp_last_error_string() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = rp_last_error_c()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! Non-zero status -> error stop with the library's message (the reference
  ! reports errors with write(6,*) + stop; here the status carries the text).
  subroutine rp_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= 0) then
      write(6, '(a)') 'ERROR: ' // what // ': ' // rp_last_error_string()
      flush(6)
      error stop 1
    end if
  end subroutine

end module respack_gpu
