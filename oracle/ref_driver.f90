! Test infrastructure only (oracle build): never shipped, never on the product path.
!
! Driver program around the reference serial solver.  It links the reference
! modules compiled unmodified from serial_ref/ and reproduces the time loop of
! serial_ref/main.f90:36-86 (initial compute_dt, make_boundaries(u), then per
! step compute_dt + godunov_unsplit with the u/u2 ping-pong), writing what the
! GPU path is compared against:
!   * the dt of every step,
!   * the output buffer after selected steps (ghosts and corners included),
!   * q, fx, fy after step 0 (module arrays of serial_ref/m_run.f90:14-17),
!   * both state buffers at the end.
! It also times the step loop with a wall clock (system_clock), for the CPU
! baseline.  The VTK output of the reference loop is not exercised here.
!
! Options come from an extra namelist group &ORACLE in the same parameter file
! (the reference reads RUN, MESH, HYDRO and ignores later groups):
!   initial_condition 'implode' : the reference initial condition (init_implode,
!                           serial_ref/m_run.f90:524-555)
!               'diag'    : x+y<1 -> rho=1,p=1 else rho=0.125,p=0.1, u=v=0,
!                           x=(i-3)/nx, y=(j-3)/ny (lower-left cell corner,
!                           ghosts included, as init_implode indexes them)
!   dump_file   output path ('' = no dump)
!   dump_steps  step indices (0-based) after which the output buffer is dumped
!   dump_stages dump q, fx, fy after step 0
!   timing      print one wall-clock timing line for the loop
!
! Record layout (stream access, little endian):
!   int32 tag, int32 step, int64 count, count x real64
!   tag 1 = dt used by `step`, 2 = output buffer after `step`,
!   3/4/5 = q/fx/fy after step 0, 6 = u at end, 7 = u2 at end.
program ref_driver

  use m_run
  use m_monitoring

  implicit none

  real(fp_kind) :: t = 0.0
  real(fp_kind) :: dt = 0.0
  integer       :: nStep = 0
  integer       :: k, i, j, unit_out, ios
  logical       :: dumping = .false.
  integer(8)    :: c0, c1, crate
  real(8)       :: wall

  character(len=20)  :: initial_condition = 'implode'
  character(len=256) :: dump_file = ''
  integer            :: dump_steps(64) = -1
  logical            :: dump_stages = .false.
  logical            :: timing = .false.
  character(len=256) :: nmlfile
  real(fp_kind)      :: x, y

  namelist/oracle/initial_condition, dump_file, dump_steps, dump_stages, timing

  call initHydroParameters()
  call get_command_argument(1, nmlfile)
  open(11, file=nmlfile, status='old')
  read(11, nml=oracle, iostat=ios)
  close(11)

  call run_init()

  if (trim(initial_condition) == 'diag') then
     do j = 1, hydroParams%jsize
        do i = 1, hydroParams%isize
           x = real(i - ghostWidth - 1, fp_kind) / real(hydroParams%nx, fp_kind)
           y = real(j - ghostWidth - 1, fp_kind) / real(hydroParams%ny, fp_kind)
           u(i,j,IU) = 0.0_fp_kind
           u(i,j,IV) = 0.0_fp_kind
           if (x + y < 1.0_fp_kind) then
              u(i,j,ID) = 1.0_fp_kind
              u(i,j,IP) = 1.0_fp_kind / (hydroParams%gamma0 - 1.0_fp_kind)
           else
              u(i,j,ID) = 0.125_fp_kind
              u(i,j,IP) = 0.1_fp_kind / (hydroParams%gamma0 - 1.0_fp_kind)
           end if
        end do
     end do
  end if
  ! u2 is never written by the reference before its first use as an output;
  ! fresh allocations come back as zero pages, which this makes explicit.
  u2 = 0.0_fp_kind

  dumping = .false.
  if (len_trim(dump_file) > 0) then
     dumping = .true.
     open(newunit=unit_out, file=trim(dump_file), access='stream', form='unformatted', &
          status='replace', action='write')
  end if

  call compute_dt(hydroParams, dt, modulo(nStep,2))
  call make_boundaries(hydroParams, u)

  call system_clock(c0, crate)
  do while (t < tEnd .and. nStep < nStepmax)
     call compute_dt(hydroParams, dt, modulo(nStep,2))
     if (dumping) call put_scalar(1, nStep, dt)
     call godunov_unsplit(nStep, dt)
     if (dumping) then
        if (nStep == 0 .and. dump_stages) then
           call put_array(3, 0, q)
           call put_array(4, 0, fx)
           call put_array(5, 0, fy)
        end if
        do k = 1, size(dump_steps)
           if (dump_steps(k) == nStep) then
              if (modulo(nStep,2) == 0) then
                 call put_array(2, nStep, u2)
              else
                 call put_array(2, nStep, u)
              end if
           end if
        end do
     end if
     nStep = nStep + 1
  end do
  call system_clock(c1)
  wall = real(c1 - c0, 8) / real(crate, 8)

  if (dumping) then
     call put_array(6, nStep, u)
     call put_array(7, nStep, u2)
     close(unit_out)
  end if

  if (timing) then
     write(*,'(a,i0,a,i0,a,i0,a,es16.8,a,es16.8)') 'ORACLE_TIMING nx=', hydroParams%nx, &
          ' ny=', hydroParams%ny, ' steps=', nStep, ' wall_s=', wall, &
          ' mcell_per_s=', real(nStep,8)*real(hydroParams%nx,8)*real(hydroParams%ny,8)/wall*1.0d-6
  end if

  call run_cleanup()

contains

  subroutine put_scalar(tag, step, val)
    integer, intent(in) :: tag, step
    real(fp_kind), intent(in) :: val
    write(unit_out) int(tag,4), int(step,4), int(1,8), real(val,8)
  end subroutine put_scalar

  subroutine put_array(tag, step, arr)
    integer, intent(in) :: tag, step
    real(fp_kind), intent(in) :: arr(:,:,:)
    write(unit_out) int(tag,4), int(step,4), int(size(arr),8), real(arr,8)
  end subroutine put_array

end program ref_driver
