!> euler2d_hip: 2-D Euler solver on the GPU, host flow of the reference program
!! (serial_ref/main.f90:22-97, cuf/main_gpu.cuf:18-97).
!!
!! usage: euler2d_hip <file.nml>
!! Outputs euler2d_<step>.vti every nOutput steps and after the last step
!! (buffer u, as serial_ref/main.f90:79-81 writes it), then the timer table and
!! the Perf line in Mcell-updates/s (reference formula: nStep*isize*jsize over
!! the total minus I/O time; also reported over interior cells).  Timers are
!! wall clock for total and I/O, GPU events for the kernels.
program euler2d_hip_main

  use, intrinsic :: iso_c_binding
  use euler2d_hip
  use HydroParameters
  use HydroRun
  implicit none

  real(8)    :: t = 0.0d0
  real(8)    :: dt = 0.0d0
  integer(8) :: c0, c1, crate, i0, i1
  real(8)    :: total, io_time, compute_time
  real(c_double)     :: ms(3)
  integer(c_int64_t) :: cnt(3)
  integer :: k
  integer(8) :: step0
  logical :: overlapped

  call initHydroParameters()
  call printHydroParameters()

  call initHydroRun()
  call e2d_check(e2d_enable_timers(ctx, 1_c_int), 'e2d_enable_timers')

  call compute_dt(dt, modulo(int(nStep), 2))
  write(*,*) 'Initial value for dt', dt

  call make_boundaries(0)

  write(*,*) 'Start computation....'
  io_time = 0.0d0
  call system_clock(c0, crate)

  do while (t < tEnd .and. nStep < nStepmax)
     overlapped = .false.
     if (modulo(int(nStep), nOutput) == 0) then
        write(*,*) 'Output results at step ', nStep, 'dt ', dt
        call system_clock(i0)
        call download(modulo(int(nStep), 2))
        if (device_loop .and. async_output) then
           ! the steps up to the next output run on the GPU while the host writes
           step0 = nStep
           k = min(nOutput - modulo(int(nStep), nOutput), nStepmax - int(nStep))
           call run_steps(k)
           overlapped = .true.
           call saveVTK(h_u, int(step0))
        else
           call saveVTK(h_u, int(nStep))
        end if
        call system_clock(i1)
        io_time = io_time + real(i1 - i0, 8) / real(crate, 8)
     end if
     if (device_loop) then
        ! up to the next output without a host round trip per step
        if (.not. overlapped) then
           k = min(nOutput - modulo(int(nStep), nOutput), nStepmax - int(nStep))
           call run_steps(k)
        end if
        call e2d_check(e2d_get_dt(ctx, dt), 'e2d_get_dt')
     else
        call compute_dt(dt, modulo(int(nStep), 2))
        call godunov_unsplit(dt)
        nStep = nStep + 1
     end if
  end do

  write(*,*) 'Output results at step ', nStep, 'dt ', dt
  call download(0)
  call saveVTK(h_u, int(nStep))

  call e2d_check(e2d_synchronize(ctx), 'e2d_synchronize')
  call system_clock(c1)
  total = real(c1 - c0, 8) / real(crate, 8)
  call e2d_check(e2d_get_timers(ctx, ms, cnt), 'e2d_get_timers')

  call cleanupHydroRun()

  write(*,*) 'total      time : ', total, 'secondes'
  write(*,*) 'compute    time : ', ms(1)*1d-3, 'secondes', 100*ms(1)*1d-3/total, '%'
  write(*,*) 'time step  time : ', ms(3)*1d-3, 'secondes', 100*ms(3)*1d-3/total, '%'
  write(*,*) 'io         time : ', io_time, 'secondes', 100*io_time/total, '%'
  write(*,*) 'boundaries time : ', ms(2)*1d-3, 'secondes', 100*ms(2)*1d-3/total, '%'
  if (device_loop .and. async_output) then
     ! output overlapped the computation: total - io no longer measures it, the GPU timers do
     compute_time = (ms(1) + ms(2) + ms(3)) * 1d-3
  else
     compute_time = total - io_time
  end if
  write(*,*) 'Perf             : ', 1.0d0*nStep*isize*jsize/compute_time*1d-6, &
       ' number of Mcell-updates/s'
  write(*,*) 'Perf (interior)  : ', 1.0d0*nStep*nx*ny/compute_time*1d-6, &
       ' number of Mcell-updates/s'

end program euler2d_hip_main
