!> ISO_C_BINDING interface to libeuler2d_hip (C-ABI: include/euler2d_hip.h).
!!
!! One interface per C entry point, with the C argument meaning; `e2d_check`
!! turns a non-zero status into `error stop` with the library's message.
module euler2d_hip

  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: E2D_IMPL_FUSED = 0, E2D_IMPL_MATERIALISED = 1, E2D_IMPL_FLUX_STORE = 2
  integer(c_int), parameter :: E2D_PREC_FP64 = 0, E2D_PREC_FP32 = 1, E2D_PREC_FP32_NATIVE = 2, &
       E2D_PREC_MIXED = 3
  integer(c_int), parameter :: E2D_RIEMANN_APPROX = 0, E2D_RIEMANN_HLLC = 2
  integer(c_int), parameter :: E2D_PROBLEM_IMPLODE = 0, E2D_PROBLEM_DIAG = 1, E2D_PROBLEM_BUMPS = 2
  integer(c_int), parameter :: E2D_TIMER_GODUNOV = 0, E2D_TIMER_BOUNDARIES = 1, E2D_TIMER_DT = 2

  !> mirrors struct e2d_params
  type, bind(C) :: e2d_params_t
     integer(c_int32_t) :: nx = 0, ny = 0
     integer(c_int32_t) :: boundary_type_xmin = 1, boundary_type_xmax = 1
     integer(c_int32_t) :: boundary_type_ymin = 1, boundary_type_ymax = 1
     integer(c_int32_t) :: slope_type = 1, niter_riemann = 10, iorder = 2
     integer(c_int32_t) :: implementation_version = 0, precision = 0, riemann_solver = 0
     real(c_double)     :: gamma0 = 1.4d0, cfl = 0.8d0, smallr = 1d-7, smallc = 1d-7
  end type e2d_params_t

  interface
     integer(c_int) function e2d_version() bind(C, name='e2d_version')
       import :: c_int
     end function e2d_version

     type(c_ptr) function e2d_last_error() bind(C, name='e2d_last_error')
       import :: c_ptr
     end function e2d_last_error

     integer(c_int) function e2d_create(p, ctx) bind(C, name='e2d_create')
       import :: c_int, c_ptr, e2d_params_t
       type(e2d_params_t), intent(in) :: p
       type(c_ptr), intent(out)       :: ctx
     end function e2d_create

     integer(c_int) function e2d_destroy(ctx) bind(C, name='e2d_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function e2d_destroy

     integer(c_int) function e2d_init_problem(ctx, problem, seed) bind(C, name='e2d_init_problem')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value        :: ctx
       integer(c_int), value     :: problem
       integer(c_int64_t), value :: seed
     end function e2d_init_problem

     integer(c_int) function e2d_upload(ctx, buf, h) bind(C, name='e2d_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value         :: ctx
       integer(c_int), value      :: buf
       real(c_double), intent(in) :: h(*)
     end function e2d_upload

     integer(c_int) function e2d_download(ctx, buf, h) bind(C, name='e2d_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: buf
       real(c_double), intent(out) :: h(*)
     end function e2d_download

     integer(c_int) function e2d_compute_dt(ctx, useU, dt) bind(C, name='e2d_compute_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: useU
       real(c_double), intent(out) :: dt
     end function e2d_compute_dt

     integer(c_int) function e2d_make_boundaries(ctx, buf) bind(C, name='e2d_make_boundaries')
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: buf
     end function e2d_make_boundaries

     integer(c_int) function e2d_godunov_unsplit(ctx, nStep, dt) bind(C, name='e2d_godunov_unsplit')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       integer(c_int), value :: nStep
       real(c_double), value :: dt
     end function e2d_godunov_unsplit

     integer(c_int) function e2d_run_steps(ctx, nsteps, nStep) bind(C, name='e2d_run_steps')
       import :: c_int, c_ptr
       type(c_ptr), value            :: ctx
       integer(c_int), value         :: nsteps
       integer(c_int), intent(inout) :: nStep
     end function e2d_run_steps

     !> no counterpart in the reference: the e2d_run_steps loop on caller-owned
     !! device buffers for callers that cannot allocate device memory
     integer(c_int) function e2d_run_steps_noalloc(p, d_u, d_u2, pitch, nsteps, nStep, d_invdt, stream) &
          bind(C, name='e2d_run_steps_noalloc')
       import :: c_int, c_int64_t, c_ptr, e2d_params_t
       type(e2d_params_t), intent(in) :: p
       type(c_ptr), value             :: d_u, d_u2
       integer(c_int64_t), value      :: pitch
       integer(c_int), value          :: nsteps, nStep
       type(c_ptr), value             :: d_invdt, stream
     end function e2d_run_steps_noalloc

     integer(c_int) function e2d_get_dt(ctx, dt) bind(C, name='e2d_get_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: dt
     end function e2d_get_dt

     integer(c_int) function e2d_get_invdt(ctx, invdt) bind(C, name='e2d_get_invdt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       real(c_double), intent(out) :: invdt
     end function e2d_get_invdt

     integer(c_int) function e2d_set_invdt(ctx, invdt) bind(C, name='e2d_set_invdt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx
       real(c_double), value :: invdt
     end function e2d_set_invdt

     integer(c_int) function e2d_sums(ctx, buf, out) bind(C, name='e2d_sums')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx
       integer(c_int), value       :: buf
       real(c_double), intent(out) :: out(4)
     end function e2d_sums

     integer(c_int) function e2d_synchronize(ctx) bind(C, name='e2d_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function e2d_synchronize

     integer(c_int) function e2d_enable_timers(ctx, on) bind(C, name='e2d_enable_timers')
       import :: c_int, c_ptr
       type(c_ptr), value    :: ctx
       integer(c_int), value :: on
     end function e2d_enable_timers

     integer(c_int) function e2d_get_timers(ctx, ms, cnt) bind(C, name='e2d_get_timers')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value              :: ctx
       real(c_double), intent(out)     :: ms(3)
       integer(c_int64_t), intent(out) :: cnt(3)
     end function e2d_get_timers

     ! -- discrete adjoint (no counterpart in the reference); device fields are
     !    FP64 arrays in the state's device layout, passed as type(c_ptr)
     integer(c_int) function e2d_field_alloc(ctx, d_field) bind(C, name='e2d_field_alloc')
       import :: c_int, c_ptr
       type(c_ptr), value       :: ctx
       type(c_ptr), intent(out) :: d_field
     end function e2d_field_alloc

     integer(c_int) function e2d_field_free(ctx, d_field) bind(C, name='e2d_field_free')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, d_field
     end function e2d_field_free

     integer(c_int) function e2d_field_upload(ctx, d_field, h) bind(C, name='e2d_field_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value         :: ctx, d_field
       real(c_double), intent(in) :: h(*)
     end function e2d_field_upload

     integer(c_int) function e2d_field_download(ctx, d_field, h) bind(C, name='e2d_field_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: ctx, d_field
       real(c_double), intent(out) :: h(*)
     end function e2d_field_download

     integer(c_int) function e2d_tangent_step(ctx, buf, dt, d_v_in, d_v_out) bind(C, name='e2d_tangent_step')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx, d_v_in, d_v_out
       integer(c_int), value :: buf
       real(c_double), value :: dt
     end function e2d_tangent_step

     integer(c_int) function e2d_adjoint_step(ctx, buf, dt, d_lam_in, d_lam_out) bind(C, name='e2d_adjoint_step')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: ctx, d_lam_in, d_lam_out
       integer(c_int), value :: buf
       real(c_double), value :: dt
     end function e2d_adjoint_step

     integer(c_int) function e2d_rollout_forward(ctx, nStep0, nsteps, h_dts, nsnap) &
          bind(C, name='e2d_rollout_forward')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value         :: ctx
       integer(c_int), value      :: nStep0, nsteps, nsnap
       real(c_double), intent(in) :: h_dts(*)
     end function e2d_rollout_forward

     integer(c_int) function e2d_adjoint_rollout(ctx, nStep0, nsteps, h_dts, nsnap, d_lam_n, d_lam_0, &
          n_forward) bind(C, name='e2d_adjoint_rollout')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value              :: ctx, d_lam_n, d_lam_0
       integer(c_int), value           :: nStep0, nsteps, nsnap
       real(c_double), intent(in)      :: h_dts(*)
       integer(c_int64_t), intent(out) :: n_forward
     end function e2d_adjoint_rollout

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  !> Message of the library's last failure.
  function e2d_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = e2d_last_error()
    msg = ''
    if (.not. c_associated(p)) return
    n = int(c_strlen(p))
    if (n <= 0) return
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
       msg(k:k) = chars(k)
    end do
  end function e2d_error_message

  !> Stop with the library's message when rc is not E2D_OK.
  subroutine e2d_check(rc, where)
    integer(c_int), intent(in)   :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
       write(*,'(a,a,a,i0,a,a)') 'euler2d_hip: ', where, ' failed (', rc, '): ', e2d_error_message()
       error stop 1
    end if
  end subroutine e2d_check

end module euler2d_hip
