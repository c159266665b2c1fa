!> Run parameters read from the reference's parameter file
!! (cuf/m_HydroParameters.cuf:17-158, serial_ref/m_run.f90:48-147).
!!
!! Namelist groups &RUN, &MESH, &HYDRO as in the reference and &OTHER with
!! implementationVersion (0 fused, 1 materialised, 2 flux-store) plus this implementation's
!! options: precision ('fp64' | 'fp32' | 'fp32_native' | 'mixed'), problem ('implode' | 'diag'
!! | 'bumps'), riemann_solver (0 two-shock, 2 HLLC), seed, device_loop
!! (keep dt on the device between outputs), async_output (with device_loop:
!! launch the steps up to the next output before writing the current one, so
!! the file is written while the GPU computes) and vtk_format ('ascii' |
!! 'binary').  &OTHER is optional; a file written for serial_ref runs as is.
module HydroParameters

  use, intrinsic :: iso_c_binding
  use euler2d_hip
  implicit none

  ! run parameters
  integer :: nStepmax = 0
  real(8) :: tEnd = 0.0d0
  integer :: nOutput = 1

  ! geometry (ghostWidth = 2, domain [0,1]^2; serial_ref/m_run.f90:28-37)
  integer, parameter :: ghostWidth = 2
  integer :: nx = 0, ny = 0
  integer :: imin = 0, imax = 0, jmin = 0, jmax = 0
  integer :: isize = 0, jsize = 0
  real(8) :: xmin = 0.0d0, xmax = 1.0d0, ymin = 0.0d0, ymax = 1.0d0
  real(8) :: dx = 0.0d0, dy = 0.0d0
  integer :: boundary_type_xmin = 1, boundary_type_xmax = 1
  integer :: boundary_type_ymin = 1, boundary_type_ymax = 1

  ! hydro
  real(8) :: gamma0 = 1.4d0, cfl = 0.8d0, smallr = 1d-7, smallc = 1d-7
  integer :: niter_riemann = 10, iorder = 2, slope_type = 2

  ! other
  integer :: implementationVersion = 0
  character(len=16) :: precision = 'fp64'
  character(len=20) :: problem = 'implode'
  integer :: riemann_solver = 0
  integer(8) :: seed = 0
  logical :: device_loop = .false.
  logical :: async_output = .false.
  character(len=16) :: vtk_format = 'ascii'

  type(e2d_params_t) :: params

contains

  subroutine initHydroParameters()
    integer :: narg, ios
    character(len=256) :: inputFilename

    namelist/run/tEnd, nStepmax, nOutput
    namelist/mesh/nx, ny, boundary_type_xmin, boundary_type_xmax, boundary_type_ymin, boundary_type_ymax
    namelist/hydro/gamma0, cfl, smallr, smallc, niter_riemann, iorder, slope_type
    namelist/other/implementationVersion, precision, problem, riemann_solver, seed, device_loop, async_output, &
         & vtk_format

    narg = command_argument_count()
    if (narg /= 1) then
       write(*,*) 'You should provide an input parameter file (with a namelist)'
       write(*,*) 'Usage:'
       write(*,*) '   ./euler2d_hip param.nml'
       stop
    end if
    call get_command_argument(1, inputFilename)

    open(1, file=inputFilename, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*,*) 'Cannot open parameter file ', trim(inputFilename)
       error stop 2
    end if
    read(1, nml=run)
    read(1, nml=mesh)
    read(1, nml=hydro)
    read(1, nml=other, iostat=ios)
    close(1)

    imin = 1
    jmin = 1
    imax = nx + 2*ghostWidth
    jmax = ny + 2*ghostWidth
    isize = imax - imin + 1
    jsize = jmax - jmin + 1
    if (nx > 0 .and. ny > 0) then
       dx = (xmax - xmin) / nx
       dy = (ymax - ymin) / ny
    end if

    if ((implementationVersion /= 0) .and. (implementationVersion /= 1) .and. (implementationVersion /= 2)) then
       write(*,*) 'The implementation version parameter should 0, 1 or 2 !!!'
       write(*,*) 'Check your namelist file, section OTHER'
       stop
    else
       write(*,*) 'Using implementation version', implementationVersion
    end if

    params%nx = nx
    params%ny = ny
    params%boundary_type_xmin = boundary_type_xmin
    params%boundary_type_xmax = boundary_type_xmax
    params%boundary_type_ymin = boundary_type_ymin
    params%boundary_type_ymax = boundary_type_ymax
    params%slope_type = slope_type
    params%niter_riemann = niter_riemann
    params%iorder = iorder
    params%implementation_version = implementationVersion
    params%riemann_solver = riemann_solver
    params%gamma0 = gamma0
    params%cfl = cfl
    params%smallr = smallr
    params%smallc = smallc
    select case (trim(precision))
    case ('fp64')
       params%precision = E2D_PREC_FP64
    case ('fp32')
       params%precision = E2D_PREC_FP32
    case ('fp32_native')
       params%precision = E2D_PREC_FP32_NATIVE
    case ('mixed')
       params%precision = E2D_PREC_MIXED
    case default
       write(*,*) 'precision should be fp64, fp32, fp32_native or mixed; got ', trim(precision)
       stop
    end select
  end subroutine initHydroParameters

  !> serial_ref/m_run.f90:131-147
  subroutine printHydroParameters()
    write(*,*) 'Simulation run parameters:'
    write(*,*) 'nx : ', nx
    write(*,*) 'ny : ', ny
    write(*,*) 'dx : ', dx
    write(*,*) 'dy : ', dy
    write(*,*) 'imin : ', imin, 'imax : ', imax
    write(*,*) 'jmin : ', jmin, 'jmax : ', jmax
    write(*,*) 'nStepmax,tEnd,nOutput : ', nStepmax, tEnd, nOutput
    write(*,*) 'gamma0,cfl : ', gamma0, cfl
    write(*,*) 'smallr,smallc,niter_riemann :', smallr, smallc, niter_riemann
    write(*,*) 'iorder,slope_type :', iorder, slope_type
    write(*,*) 'precision, problem : ', trim(precision), ' ', trim(problem)
  end subroutine printHydroParameters

end module HydroParameters
