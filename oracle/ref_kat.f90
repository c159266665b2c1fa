! Test infrastructure only (oracle build): never shipped, never on the product path.
!
! Known-answer driver for the reference pointwise numerics
! (serial_ref/m_utils.f90).  It reads a batch of input vectors, calls one
! reference routine per vector, and writes the outputs, all as real64.
!
! Usage: ref_kat <in.bin> <out.bin>
! in.bin  : int32 hdr(8) = [func, n, slope_type, niter_riemann, nin, nout, 0, 0]
!           real64 par(8) = [gamma0, smallr, smallc, 0, 0, 0, 0, 0]
!           real64 x(nin, n)
! out.bin : real64 y(nout, n)
! func 1 computePrimitives  (m_utils.f90:35-64)   in u(4)            -> q(4), c
!      2 slope_unsplit_hydro_2d (:268-342)        in q,q+x,q-x,q+y,q-y -> dq(1,:), dq(2,:)
!      3 trace_unsplit_2d_along_dir (:177-254)    in q,dqX,dqY,dtdx,dtdy,face -> qface(4)
!      4 riemann_2d + cmpflx (:349-526)           in qleft, qright  -> qgdnv(4), flux(4)
!      5 riemann_2d_hllc (:534-644)               in qleft, qright  -> qgdnv(4), flux(4)
!      6 cmpflx (:498-526)                        in qgdnv(4)       -> flux(4)
!      7 eos (:15-29)                             in rho, eint      -> p, c
! Derived parameters follow initHydroParameters (serial_ref/m_run.f90:122-124).
program ref_kat

  use m_precision
  use m_constants
  use m_parameters
  use m_utils

  implicit none

  integer(4) :: hdr(8)
  real(8)    :: par(8)
  real(8), allocatable :: x(:,:), y(:,:)
  character(len=512) :: fin, fout
  integer :: n, nin, nout, k, func, uin, uout
  type(HydroPar) :: p
  real(fp_kind) :: a(4), b(4), qn(4,5), dq(2,4), qg(4), fl(4), c, pr
  integer(int_kind) :: face

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(uin) hdr
  read(uin) par
  func = hdr(1); n = hdr(2); nin = hdr(5); nout = hdr(6)
  allocate(x(nin, n), y(nout, n))
  read(uin) x
  close(uin)
  y = 0.0d0

  p%gamma0        = real(par(1), fp_kind)
  p%smallr        = real(par(2), fp_kind)
  p%smallc        = real(par(3), fp_kind)
  p%slope_type    = hdr(3)
  p%niter_riemann = hdr(4)
  p%iorder        = 2
  p%cfl           = 0.8
  p%smallp  = p%smallc * p%smallc / p%gamma0
  p%smallpp = p%smallr * p%smallp
  p%gamma6  = (p%gamma0 + ONE_F) / (TWO_F * p%gamma0)

  do k = 1, n
     select case (func)
     case (1)
        a = real(x(1:4,k), fp_kind)
        call computePrimitives(a, b, c, p)
        y(1:4,k) = real(b, 8); y(5,k) = real(c, 8)
     case (2)
        qn = reshape(real(x(1:20,k), fp_kind), (/4,5/))
        call slope_unsplit_hydro_2d(qn(:,1), qn(:,2), qn(:,3), qn(:,4), qn(:,5), dq, p)
        y(1:4,k) = real(dq(1,:), 8); y(5:8,k) = real(dq(2,:), 8)
     case (3)
        face = int(x(15,k), int_kind)
        call trace_unsplit_2d_along_dir(real(x(1:4,k), fp_kind), real(x(5:8,k), fp_kind), &
             real(x(9:12,k), fp_kind), real(x(13,k), fp_kind), real(x(14,k), fp_kind), face, b, p)
        y(1:4,k) = real(b, 8)
     case (4)
        call riemann_2d(real(x(1:4,k), fp_kind), real(x(5:8,k), fp_kind), qg, fl, p)
        y(1:4,k) = real(qg, 8); y(5:8,k) = real(fl, 8)
     case (5)
        call riemann_2d_hllc(real(x(1:4,k), fp_kind), real(x(5:8,k), fp_kind), qg, fl, p)
        y(1:4,k) = real(qg, 8); y(5:8,k) = real(fl, 8)
     case (6)
        call cmpflx(real(x(1:4,k), fp_kind), fl, p)
        y(1:4,k) = real(fl, 8)
     case (7)
        pr = 0.0
        call eos(real(x(1,k), fp_kind), real(x(2,k), fp_kind), pr, c, p)
        y(1,k) = real(pr, 8); y(2,k) = real(c, 8)
     case default
        write(*,*) 'ref_kat: unknown function', func
        stop 2
     end select
  end do

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(uout) y
  close(uout)

end program ref_kat
