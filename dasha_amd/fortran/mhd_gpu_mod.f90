! mhd_gpu_mod.f90 -- ISO_C_BINDING interface to the MI355X step library
! (include/mhd_gpu.h).  The Fortran host program calls these in place of
! the reference's setdt / stepon / current / diagn (t3dsdsc.f:139, 171,
! 1171, 1174-1182).
module mhd_gpu
  use iso_c_binding
  implicit none

  type, bind(C) :: mhd_config
     integer(c_int) :: mx, my, mz
     integer(c_int) :: periodic_y, hall, whistler_dt
     real(c_double) :: di, eta0, fmu0, frho0, fpr0, gamma, cfl
     ! magnetosphere mode (t3dsdsc.f:1398-1399, 1460-1548, 1692-1772, 4712)
     integer(c_int) :: magnetosphere = 0, inflow = 0, radial_hall = 0, partial_e = 0
     real(c_double) :: rcr = 0.d0
  end type mhd_config

  type, bind(C) :: mhd_diag
     real(c_double) :: cxmax, cxmin, cymax, cymin, czmax, czmin
     real(c_double) :: e_kin, e_mag, e_th, e_mag_pert
     real(c_double) :: divb_max, divb_rms
  end type mhd_diag

  ! Harris-sheet initial state built on the device (mhd_init_harris)
  type, bind(C) :: mhd_harris
     real(c_double) :: lam, beta_inf, guide, temp, amp, width, noise, x_offset
     integer(c_long_long) :: seed
  end type mhd_harris

  integer(c_int), parameter :: MHD_OK = 0, MHD_ERR_ARG = -1, MHD_ERR_HIP = -2, &
       MHD_ERR_NOMEM = -3, MHD_ERR_DT_COLLAPSE = -4, MHD_ERR_BLOWUP = -5
  integer(c_int), parameter :: MHD_ADJ_RESIDUAL = 1

  interface
     integer(c_int) function mhd_create(cfg, xx, yy, zz, dx, dy, dz, h) bind(C, name='mhd_create')
       import :: c_int, c_double, c_ptr, mhd_config
       type(mhd_config), intent(in) :: cfg
       real(c_double), intent(in) :: xx(*), yy(*), zz(*), dx(*), dy(*), dz(*)
       type(c_ptr), intent(out) :: h
     end function mhd_create

     integer(c_int) function mhd_set_equilibrium(h, rhoi, bxi, byi, bzi, pri) &
          bind(C, name='mhd_set_equilibrium')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: rhoi(*), bxi(*), byi(*), bzi(*), pri(*)
     end function mhd_set_equilibrium

     integer(c_int) function mhd_set_state(h, x) bind(C, name='mhd_set_state')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: x(*)
     end function mhd_set_state

     integer(c_int) function mhd_get_state(h, x) bind(C, name='mhd_get_state')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: x(*)
     end function mhd_get_state

     integer(c_int) function mhd_init_harris(h, p) bind(C, name='mhd_init_harris')
       import :: c_int, c_ptr, mhd_harris
       type(c_ptr), value :: h
       type(mhd_harris), intent(in) :: p
     end function mhd_init_harris

     integer(c_int) function mhd_init_residual(h) bind(C, name='mhd_init_residual')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function mhd_init_residual

     integer(c_int) function mhd_set_inflow_face(h, bxinfc, byinfc, bzinfc) bind(C, name='mhd_set_inflow_face')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: bxinfc(*), byinfc(*), bzinfc(*)
     end function mhd_set_inflow_face

     integer(c_int) function mhd_compute_dt_noalloc(h, dt) bind(C, name='mhd_compute_dt_noalloc')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: dt
     end function mhd_compute_dt_noalloc

     integer(c_int) function mhd_compute_dt(h, dt) bind(C, name='mhd_compute_dt')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: dt
     end function mhd_compute_dt

     integer(c_int) function mhd_step(h, time, dt) bind(C, name='mhd_step')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), value :: time, dt
     end function mhd_step

     integer(c_int) function mhd_get_current(h, cur) bind(C, name='mhd_get_current')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: cur(*)
     end function mhd_get_current

     integer(c_int) function mhd_diagnostics(h, d) bind(C, name='mhd_diagnostics')
       import :: c_int, c_ptr, mhd_diag
       type(c_ptr), value :: h
       type(mhd_diag), intent(out) :: d
     end function mhd_diagnostics

     type(c_ptr) function mhd_last_error_c(h) bind(C, name='mhd_last_error')
       import :: c_ptr
       type(c_ptr), value :: h
     end function mhd_last_error_c

     subroutine mhd_destroy(h) bind(C, name='mhd_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine mhd_destroy

     ! ensemble of independent runs (handles created as above), advanced together
     integer(c_int) function mhd_ensemble_create(runs, n, e) bind(C, name='mhd_ensemble_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(in) :: runs(*)
       integer(c_int), value :: n
       type(c_ptr), intent(out) :: e
     end function mhd_ensemble_create

     integer(c_int) function mhd_ensemble_advance(e, nsteps, dt0, time, elapsed_ms) &
          bind(C, name='mhd_ensemble_advance')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: e
       integer(c_int), value :: nsteps
       real(c_double), intent(in) :: dt0(*)
       real(c_double), intent(inout) :: time(*)
       real(c_double), intent(out) :: elapsed_ms
     end function mhd_ensemble_advance

     integer(c_int) function mhd_ensemble_status(e, status) bind(C, name='mhd_ensemble_status')
       import :: c_int, c_ptr
       type(c_ptr), value :: e
       integer(c_int), intent(out) :: status(*)
     end function mhd_ensemble_status

     ! discrete adjoint of nsteps steps with the frozen dt sequence dts
     ! (objective 0: magnetic energy, 1: reconnected flux); grad_x has the
     ! layout of x, ms(2) the forward / reverse times
     integer(c_int) function mhd_adjoint_run(h, nsteps, dts, objective, every, flags, value, grad_x, &
          grad_eta, ms) bind(C, name='mhd_adjoint_run')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nsteps, objective, every, flags
       real(c_double), intent(in) :: dts(*)
       real(c_double), intent(out) :: value, grad_x(*), grad_eta, ms(2)
     end function mhd_adjoint_run

     integer(c_int) function mhd_ensemble_tune(e, trial, dt0, time, best, ms) bind(C, name='mhd_ensemble_tune')
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: e
       integer(c_int), value :: trial
       real(c_double), intent(in) :: dt0(*), time(*)
       integer(c_int), intent(out) :: best
       real(c_double), intent(out) :: ms(4)
     end function mhd_ensemble_tune

     subroutine mhd_ensemble_destroy(e) bind(C, name='mhd_ensemble_destroy')
       import :: c_ptr
       type(c_ptr), value :: e
     end subroutine mhd_ensemble_destroy
  end interface

contains

  ! text of the last error of a handle
  function mhd_last_error(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    msg = ' '
    p = mhd_last_error_c(h)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do n = 1, 256
       if (s(n) == c_null_char) exit
       msg(n:n) = s(n)
    end do
  end function mhd_last_error

end module mhd_gpu
