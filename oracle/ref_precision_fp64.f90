! Test infrastructure only (oracle build).
!
! Precision-selection module used to build the reference serial solver in
! double precision.  The reference ships `serial_ref/m_precision.f90` with
! `fp_kind = sp_kind` (line 25, "default: use single precison"); the FP64
! oracle build compiles every other reference source unmodified from where it
! lies and links this module in its place.  The interface (module name,
! kind parameters, the two helper functions used by `saveVTK`,
! serial_ref/m_run.f90:670,684) is kept identical so the other reference
! modules compile against it unchanged; only the selected kind differs.
module m_precision

  integer, parameter :: sp_kind  = kind(1.0)
  integer, parameter :: dp_kind  = kind(1.d0)
  integer, parameter :: int_kind = kind(1)

  ! the one deliberate difference from the shipped module: double precision
  integer, parameter :: fp_kind = dp_kind

contains

  logical function isBigEndian()
    integer(kind=2) :: probe
    probe = 1_2
    isBigEndian = (iachar(transfer(probe, 'a')) == 0)
  end function isBigEndian

  logical function useDoublePrecision()
    useDoublePrecision = (fp_kind == dp_kind)
  end function useDoublePrecision

end module m_precision
