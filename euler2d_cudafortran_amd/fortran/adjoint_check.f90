!> The discrete adjoint from Fortran, through the bind(C) interfaces of module
!! euler2d_hip: the gradient of J = sum of rho over the interior after nsteps
!! frozen-dt steps of the diagonal-discontinuity problem with respect to the
!! initial state, by a recorded forward (e2d_rollout_forward) and the
!! checkpointed reverse sweep (e2d_adjoint_rollout).
!!
!! usage: e2d_adjoint_check [nx ny nsteps nsnap]   (defaults 64 48 20 6)
!! Prints the forward steps recomputed and, per conserved variable, the sum
!! of the gradient over the interior.
program adjoint_check
  use, intrinsic :: iso_c_binding
  use euler2d_hip
  implicit none

  type(e2d_params_t) :: p
  type(c_ptr) :: ctx, lam_n, lam_0
  integer :: nx, ny, nsteps, nsnap, n, i, j, v, isize, jsize
  integer(c_int64_t) :: nfwd
  real(c_double) :: dt
  real(c_double), allocatable :: dts(:), lam(:, :, :)
  character(len=32) :: arg

  nx = 64; ny = 48; nsteps = 20; nsnap = 6
  if (command_argument_count() >= 4) then
     call get_command_argument(1, arg); read(arg, *) nx
     call get_command_argument(2, arg); read(arg, *) ny
     call get_command_argument(3, arg); read(arg, *) nsteps
     call get_command_argument(4, arg); read(arg, *) nsnap
  end if
  isize = nx + 4
  jsize = ny + 4

  p%nx = nx
  p%ny = ny
  call e2d_check(e2d_create(p, ctx), 'e2d_create')

  ! the dt sequence of the reference loop, then frozen
  allocate(dts(nsteps))
  call start_state()
  do n = 0, nsteps - 1
     call e2d_check(e2d_compute_dt(ctx, mod(n, 2), dt), 'e2d_compute_dt')
     dts(n + 1) = dt
     call e2d_check(e2d_godunov_unsplit(ctx, n, dt), 'e2d_godunov_unsplit')
  end do

  ! forward again, keeping the reverse sweep's first checkpoints
  call start_state()
  call e2d_check(e2d_rollout_forward(ctx, 0, nsteps, dts, nsnap), 'e2d_rollout_forward')

  ! dJ/dU_n: 1 on the interior density
  allocate(lam(isize, jsize, 4))
  lam = 0.0d0
  lam(3:nx + 2, 3:ny + 2, 1) = 1.0d0
  call e2d_check(e2d_field_alloc(ctx, lam_n), 'e2d_field_alloc')
  call e2d_check(e2d_field_alloc(ctx, lam_0), 'e2d_field_alloc')
  call e2d_check(e2d_field_upload(ctx, lam_n, lam), 'e2d_field_upload')

  call start_state()
  call e2d_check(e2d_adjoint_rollout(ctx, 0, nsteps, dts, nsnap, lam_n, lam_0, nfwd), 'e2d_adjoint_rollout')
  call e2d_check(e2d_field_download(ctx, lam_0, lam), 'e2d_field_download')

  write(*, '(a,i0)') 'forward_steps_recomputed ', nfwd
  do v = 1, 4
     write(*, '(a,i0,a,es26.17e3)') 'gradient_sum_var', v, ' ', sum(lam(3:nx + 2, 3:ny + 2, v))
  end do

  call e2d_check(e2d_field_free(ctx, lam_n), 'e2d_field_free')
  call e2d_check(e2d_field_free(ctx, lam_0), 'e2d_field_free')
  call e2d_check(e2d_destroy(ctx), 'e2d_destroy')

contains

  subroutine start_state()
    call e2d_check(e2d_init_problem(ctx, E2D_PROBLEM_DIAG, 0_c_int64_t), 'e2d_init_problem')
    call e2d_check(e2d_make_boundaries(ctx, 0), 'e2d_make_boundaries')
  end subroutine start_state

end program adjoint_check
