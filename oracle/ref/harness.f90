! Test harness around the reference MAGIC operators (compiled together with the
! reference's own Fortran sources by oracle/ref/Makefile into oracle/_ref/).
!
! This file is test infrastructure written for this repository: it contains no
! reference code.  It only
!   * brings up a 1-rank MPI world and the 1x1x1 Cartesian communicator that the
!     reference's bc3 (bc3_mpi.f:86) and claw3 (mpilib/claw3_mpi.f:843) expect in
!     the commons /mpicomm/ and /mpi_proc_info/ (the reference driver does this
!     in mpilib/claw3ez_mpi_driver.f:49-57 and 362-382);
!   * sizes the work arrays the reference step3/claw3 need
!     (mpilib/claw3_mpi.f:689-728, lib/step3.f:56-89) and forwards to them with
!     the reference's own rpn3/rpt3/rptt3/bc3/src3/b4step3 as the EXTERNAL
!     operator arguments.
! Every entry point uses the F77 calling convention (all arguments by
! reference, lowercase name + trailing underscore), so Python's ctypes calls it
! directly.

subroutine refh_mpi_setup(mthbc)
  implicit none
  include 'mpif.h'
  integer, intent(in) :: mthbc(6)
  integer :: mpi_comm_3d, lx, ly, lz, mtotal(3), mstart(3)
  common /mpicomm/ mpi_comm_3d, lx, ly, lz, mtotal, mstart
  integer :: np, id
  common /mpi_proc_info/ np, id
  logical :: inited, periods(3)
  integer :: ierr, ndims(3), i
  call mpi_initialized(inited, ierr)
  if (.not. inited) call mpi_init(ierr)
  call mpi_comm_size(MPI_COMM_WORLD, np, ierr)
  call mpi_comm_rank(MPI_COMM_WORLD, id, ierr)
  ndims = 1
  do i = 1, 3
     periods(i) = (mthbc(2*i) .eq. 2)
  end do
  call mpi_cart_create(MPI_COMM_WORLD, 3, ndims, periods, .true., &
                       mpi_comm_3d, ierr)
  lx = 1
  ly = 1
  lz = 1
  mstart = 0
end subroutine refh_mpi_setup

! Set the time/space common read by rpn3 (rpn3euX.f:155-158); claw3 sets it at
! mpilib/claw3_mpi.f:748-752 before every step3 call.
subroutine refh_set_comxyzt(dt, dx, dy, dz, t)
  implicit none
  double precision, intent(in) :: dt, dx, dy, dz, t
  double precision :: dtcom, dxcom, dycom, dzcom, tcom
  integer :: icom, jcom, kcom
  common /comxyzt/ dtcom, dxcom, dycom, dzcom, tcom, icom, jcom, kcom
  dtcom = dt
  dxcom = dx
  dycom = dy
  dzcom = dz
  tcom = t
end subroutine refh_set_comxyzt

! One call of the reference step3 (lib/step3.f:3-7) with method = (1,2,22,*,1,0,maux),
! mthlim = 0 and the reference Riemann operators.  qnew must enter equal to qold.
subroutine refh_step3(meqn, mbc, mx, my, mz, maux, qold, qnew, aux, &
                      dx, dy, dz, dt, cfl)
  implicit none
  integer, intent(in) :: meqn, mbc, mx, my, mz, maux
  double precision, intent(in) :: qold(*), aux(*)
  double precision, intent(inout) :: qnew(*)
  double precision, intent(in) :: dx, dy, dz, dt
  double precision, intent(out) :: cfl
  external rpn3, rpt3, rptt3
  integer, parameter :: mwaves = 3
  integer :: method(7), mthlim(mwaves), maxm, n1, mwork
  logical :: use_fwave
  double precision, allocatable :: qadd(:), fadd(:), gadd(:), hadd(:), q1d(:), &
       dtdx1d(:), dtdy1d(:), dtdz1d(:), aux1(:), aux2(:), aux3(:), work(:)
  method = (/1, 2, 22, 0, 1, 0, maux/)
  mthlim = 0
  use_fwave = .true.
  maxm = max(mx, my, mz)
  n1 = maxm + 2*mbc
  mwork = n1*(32*meqn + mwaves + meqn*mwaves) + 16
  allocate(qadd(n1*meqn), fadd(n1*meqn), gadd(6*n1*meqn), hadd(6*n1*meqn), &
           q1d(n1*meqn), dtdx1d(n1), dtdy1d(n1), dtdz1d(n1), &
           aux1(3*n1*max(maux,1)), aux2(3*n1*max(maux,1)), &
           aux3(3*n1*max(maux,1)), work(mwork))
  call refh_set_comxyzt(dt, dx, dy, dz, 0.d0)
  call step3(maxm, meqn, mwaves, mbc, mx, my, mz, qold, qnew, aux, dx, dy, dz, &
             dt, method, mthlim, cfl, qadd, fadd, gadd, hadd, q1d, dtdx1d, &
             dtdy1d, dtdz1d, aux1, aux2, aux3, maux, work, mwork, rpn3, rpt3, &
             rptt3, use_fwave)
  deallocate(qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d, aux1, aux2, &
             aux3, work)
end subroutine refh_step3

! One call of the reference claw3 (mpilib/claw3_mpi.f:4-7) with the reference
! operators.  With tend < tstart (outstyle 3, mpilib/claw3ez_mpi.f:143) it takes
! exactly one accepted step; dtv/cflv/nv come back as claw3 leaves them.
subroutine refh_claw3(meqn, mbc, mx, my, mz, maux, q, aux, xlower, ylower, &
                      zlower, dx, dy, dz, tstart, tend, dtv, cflv, nv, method, &
                      mthbc, info)
  implicit none
  integer, intent(in) :: meqn, mbc, mx, my, mz, maux
  double precision, intent(inout) :: q(*)
  double precision, intent(inout) :: aux(*)
  double precision, intent(in) :: xlower, ylower, zlower, dx, dy, dz, tstart
  double precision, intent(inout) :: tend, dtv(5), cflv(4)
  integer, intent(inout) :: nv(2), method(7), mthbc(6)
  integer, intent(out) :: info
  external bc3, rpn3, rpt3, rptt3, src3, b4step3
  integer, parameter :: mwaves = 3
  integer :: mthlim(mwaves), maxm, mwork, narray
  logical :: use_fwave
  double precision, allocatable :: work(:)
  mthlim = 0
  use_fwave = .true.
  maxm = max(mx, my, mz)
  narray = 1
  if (method(5) .ge. 2) narray = 2
  mwork = (maxm + 2*mbc)*(46*meqn + mwaves + meqn*mwaves + 9*maux + 3) &
        + narray*(mx + 2*mbc)*(my + 2*mbc)*(mz + 2*mbc)*meqn
  allocate(work(mwork))
  work = 0.d0
  call claw3(meqn, mwaves, mbc, mx, my, mz, maux, q, aux, xlower, ylower, &
             zlower, dx, dy, dz, tstart, tend, dtv, cflv, nv, method, mthlim, &
             mthbc, work, mwork, use_fwave, info, bc3, rpn3, rpt3, rptt3, &
             src3, b4step3)
  deallocate(work)
end subroutine refh_claw3
