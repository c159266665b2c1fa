!> Device state of a run and the procedures of the reference's HydroRun module
!! (cuf/m_HydroRun_gpu.cuf, serial_ref/m_run.f90), backed by libeuler2d_hip.
!!
!! Buffer 0 is u, buffer 1 is u2; step nStep reads buffer modulo(nStep,2) and
!! writes the other one (serial_ref/m_run.f90:197-201).  h_u is the host copy
!! used for output, in the reference's layout u(isize, jsize, 4).
module HydroRun

  use, intrinsic :: iso_c_binding
  use euler2d_hip
  use HydroParameters
  implicit none

  real(c_double), allocatable, target :: h_u(:,:,:)
  type(c_ptr) :: ctx = c_null_ptr
  integer(c_int) :: nStep = 0

  character(len=3), dimension(4), parameter :: varNames = (/ 'rho', '  E', ' mx', ' my' /)

contains

  !> allocation + initial condition (cuf/m_HydroRun_gpu.cuf:83-120); the
  !! initial condition is computed on the device in the run's precision.
  subroutine initHydroRun()
    integer(c_int) :: code
    allocate(h_u(isize, jsize, 4))
    h_u = 0.0d0
    call e2d_check(e2d_create(params, ctx), 'e2d_create')
    select case (trim(problem))
    case ('implode')
       code = E2D_PROBLEM_IMPLODE
    case ('diag')
       code = E2D_PROBLEM_DIAG
    case ('bumps')
       code = E2D_PROBLEM_BUMPS
    case default
       write(*,*) 'Unknown problem; default to implode'
       code = E2D_PROBLEM_IMPLODE
    end select
    call e2d_check(e2d_init_problem(ctx, code, int(seed, c_int64_t)), 'e2d_init_problem')
    nStep = 0
  end subroutine initHydroRun

  subroutine cleanupHydroRun()
    call e2d_check(e2d_destroy(ctx), 'e2d_destroy')
    ctx = c_null_ptr
    if (allocated(h_u)) deallocate(h_u)
  end subroutine cleanupHydroRun

  !> compute_dt (serial_ref/m_run.f90:248-312): running maximum kept by the library
  subroutine compute_dt(dt, useU)
    real(8), intent(out) :: dt
    integer, intent(in)  :: useU
    call e2d_check(e2d_compute_dt(ctx, int(useU, c_int), dt), 'e2d_compute_dt')
  end subroutine compute_dt

  subroutine make_boundaries(buf)
    integer, intent(in) :: buf
    call e2d_check(e2d_make_boundaries(ctx, int(buf, c_int)), 'e2d_make_boundaries')
  end subroutine make_boundaries

  !> one step with the host's dt (godunov_unsplit, serial_ref/m_run.f90:187-205)
  subroutine godunov_unsplit(dt)
    real(8), intent(in) :: dt
    call e2d_check(e2d_godunov_unsplit(ctx, nStep, dt), 'e2d_godunov_unsplit')
  end subroutine godunov_unsplit

  !> n steps with dt kept on the device; advances nStep
  subroutine run_steps(n)
    integer, intent(in) :: n
    call e2d_check(e2d_run_steps(ctx, int(n, c_int), nStep), 'e2d_run_steps')
  end subroutine run_steps

  subroutine download(buf)
    integer, intent(in) :: buf
    call e2d_check(e2d_download(ctx, int(buf, c_int), h_u), 'e2d_download')
  end subroutine download

  logical function isBigEndian()
    integer(kind=2) :: probe
    probe = 1_2
    isBigEndian = (iachar(transfer(probe, 'a')) == 0)
  end function isBigEndian

  !> VTK ImageData output in the reference's ASCII layout
  !! (serial_ref/m_run.f90:653-720): interior cells only, one DataArray per
  !! variable, list-directed rows, values in the run's precision.
  subroutine saveVTK(data, iStep)
    real(c_double), intent(in) :: data(isize, jsize, 4)
    integer, intent(in) :: iStep
    integer :: j, iVar, error
    character(len=80) :: filename
    character(len=8)  :: filenum
    character(1), parameter :: endl = char(10)
    character(500) :: charBuf
    character(7) :: floatType
    logical :: dbl

    if (trim(vtk_format) == 'binary') then
       call saveVTK_binary(data, iStep)
       return
    end if
    dbl = (params%precision == E2D_PREC_FP64)
    if (dbl) then
       write(floatType,'(a)') 'Float64'
    else
       write(floatType,'(a)') 'Float32'
    end if
    write(filenum,'(i8.8)') iStep
    filename = 'euler2d_' // filenum // '.vti'
    open(10, file=filename, iostat=error)

    write(10,'(a)') '<?xml version="1.0"?>'//endl
    if (isBigEndian()) then
       write(10,'(a)') '<VTKFile type="ImageData" version="0.1" byte_order="BigEndian">'//endl
    else
       write(10,'(a)') '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">'//endl
    end if
    write(charBuf,fmt='(6(I7))',iostat=error) 1, nx+1, 1, ny+1, 1, 2
    write(10,'(a)') repeat(' ',2)//'<ImageData WholeExtent="'//trim(charBuf)//'"'
    write(10,'(a)') ' Origin="0 0 0" Spacing="1 1 1">'//endl
    write(10,'(a)') repeat(' ',2)//'<Piece Extent="'//trim(charBuf)//'">'//endl
    write(10,'(a)') repeat(' ',3)//'<PointData>'//endl
    write(10,'(a)') repeat(' ',4)//'</PointData>'//endl
    write(10,'(a)') repeat(' ',4)//'<CellData>'//endl
    do iVar = 1, 4
       write(10,'(a)') repeat(' ',4)//'<DataArray type="'//trim(floatType)// &
            & '" Name="'//varNames(iVar)//'" format="ascii" >'//endl
       do j = jmin+ghostWidth, jmax-ghostWidth
          if (dbl) then
             write(10,*) data(imin+ghostWidth:imax-ghostWidth, j, iVar)
          else
             write(10,*) real(data(imin+ghostWidth:imax-ghostWidth, j, iVar), 4)
          end if
       end do
       write(10,'(a)') repeat(' ',4)//'</DataArray>'//endl
    end do
    write(10,'(a)') repeat(' ',4)//'</CellData>'//endl
    write(10,'(a)') repeat(' ',2)//'</Piece>'//endl
    write(10,'(a)') repeat(' ',2)//'</ImageData>'//endl
    write(10,'(a)') '</VTKFile>'//endl
    close(10)
  end subroutine saveVTK

  !> Same content as saveVTK in VTK's appended raw binary encoding (one
  !! UInt64 byte count before each array), for large grids.
  subroutine saveVTK_binary(data, iStep)
    real(c_double), intent(in) :: data(isize, jsize, 4)
    integer, intent(in) :: iStep
    integer :: iVar, u
    integer(8) :: nbytes, offset
    character(len=80) :: filename
    character(len=8)  :: filenum
    character(len=64) :: ext, off
    character(1), parameter :: endl = char(10)
    character(7) :: floatType
    logical :: dbl
    real(4), allocatable :: buf4(:,:)

    dbl = (params%precision == E2D_PREC_FP64)
    floatType = merge('Float64', 'Float32', dbl)
    nbytes = int(nx, 8) * int(ny, 8) * merge(8_8, 4_8, dbl)
    write(filenum,'(i8.8)') iStep
    filename = 'euler2d_' // filenum // '.vti'
    write(ext,'(6(1x,i0))') 1, nx+1, 1, ny+1, 1, 2
    open(newunit=u, file=filename, access='stream', form='unformatted', status='replace')
    write(u) '<?xml version="1.0"?>'//endl
    write(u) '<VTKFile type="ImageData" version="0.1" byte_order="' // &
         merge('BigEndian   ', 'LittleEndian', isBigEndian()) // '" header_type="UInt64">'//endl
    write(u) '  <ImageData WholeExtent="'//trim(adjustl(ext))//'" Origin="0 0 0" Spacing="1 1 1">'//endl
    write(u) '  <Piece Extent="'//trim(adjustl(ext))//'">'//endl
    write(u) '    <CellData>'//endl
    offset = 0
    do iVar = 1, 4
       write(off,'(i0)') offset
       write(u) '    <DataArray type="'//floatType//'" Name="'//trim(adjustl(varNames(iVar)))// &
            '" format="appended" offset="'//trim(off)//'"/>'//endl
       offset = offset + 8 + nbytes
    end do
    write(u) '    </CellData>'//endl
    write(u) '  </Piece>'//endl
    write(u) '  </ImageData>'//endl
    write(u) '  <AppendedData encoding="raw">'//endl
    write(u) '_'
    if (.not. dbl) allocate(buf4(nx, ny))
    do iVar = 1, 4
       write(u) nbytes
       if (dbl) then
          write(u) data(imin+ghostWidth:imax-ghostWidth, jmin+ghostWidth:jmax-ghostWidth, iVar)
       else
          buf4 = real(data(imin+ghostWidth:imax-ghostWidth, jmin+ghostWidth:jmax-ghostWidth, iVar), 4)
          write(u) buf4
       end if
    end do
    write(u) endl//'  </AppendedData>'//endl
    write(u) '</VTKFile>'//endl
    close(u)
  end subroutine saveVTK_binary

end module HydroRun
