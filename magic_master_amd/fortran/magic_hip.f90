! iso_c_binding interface of the MI355X MAGIC step (C ABI: include/magic_hip.h).
!
! The reference's time integrator calls its operators as F77 externals
! (mpilib/claw3_mpi.f:765-905).  This module binds the device-resident
! equivalents so a Fortran host can drive them; claw3_hip.f90 builds the
! drop-in claw3 / step3 on top of it.
module magic_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MAGIC_OK = 0, MAGIC_ERHO = 1, MAGIC_EA2 = 2, &
       MAGIC_EBC = 3, MAGIC_EUNSUPPORTED = 4, MAGIC_EARG = 5, MAGIC_EHIP = 6, &
       MAGIC_ENOMEM = 7

  ! setprob commons (setprob.f:17-25) in the order setprob.f:31-61 reads them
  type, bind(C) :: magic_params
     integer(c_int) :: loadprof, nspecadv
     real(c_double) :: gamma, grav, dens, pres, dmolec0, dtherm0
     real(c_double) :: forcemth, omega, amplitude, propx, propz
     real(c_double) :: xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth, vsrcx
     real(c_double) :: amp0, amp1, omega1, prop1z, phi, z1width, z1pos, z1Gpos, windscale1
  end type magic_params

  interface
     integer(c_int) function magic_create(mx, my, mz, meqn, mbc, maux, xlower, ylower, &
          zlower, dx, dy, dz, ctx) bind(C, name="magic_create")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: mx, my, mz, meqn, mbc, maux
       real(c_double), value :: xlower, ylower, zlower, dx, dy, dz
       type(c_ptr), intent(out) :: ctx
     end function magic_create

     integer(c_int) function magic_destroy(ctx) bind(C, name="magic_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function magic_destroy

     type(c_ptr) function magic_strerror(status) bind(C, name="magic_strerror")
       import :: c_int, c_ptr
       integer(c_int), value :: status
     end function magic_strerror

     integer(c_int) function magic_set_params(ctx, p) bind(C, name="magic_set_params")
       import :: c_int, c_ptr, magic_params
       type(c_ptr), value :: ctx
       type(magic_params), intent(in) :: p
     end function magic_set_params

     integer(c_int) function magic_set_aux(ctx, aux) bind(C, name="magic_set_aux")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: aux(*)
     end function magic_set_aux

     integer(c_int) function magic_set_q(ctx, q) bind(C, name="magic_set_q")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: q(*)
     end function magic_set_q

     integer(c_int) function magic_get_q(ctx, q) bind(C, name="magic_get_q")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: q(*)
     end function magic_get_q

     integer(c_int) function magic_bc3(ctx, mthbc) bind(C, name="magic_bc3")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(in) :: mthbc(6)
     end function magic_bc3

     integer(c_int) function magic_b4step3(ctx) bind(C, name="magic_b4step3")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function magic_b4step3

     integer(c_int) function magic_backup(ctx) bind(C, name="magic_backup")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function magic_backup

     integer(c_int) function magic_restore(ctx) bind(C, name="magic_restore")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function magic_restore

     integer(c_int) function magic_step3(ctx, dt, cfl) bind(C, name="magic_step3")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       real(c_double), intent(out) :: cfl
     end function magic_step3

     integer(c_int) function magic_src3(ctx, t, dt, mthbc, ndtr, ndtrc) bind(C, name="magic_src3")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: t, dt
       integer(c_int), intent(in) :: mthbc(6)
       integer(c_int), intent(out) :: ndtr, ndtrc
     end function magic_src3

     integer(c_int) function magic_claw3(ctx, tstart, tend, dtv, cflv, nv, method, mthbc, &
          log_dt, log_cfl, log_acc, maxlog, nlog) bind(C, name="magic_claw3")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: tstart
       real(c_double), intent(inout) :: tend, dtv(5), cflv(4)
       integer(c_int), intent(inout) :: nv(2)
       integer(c_int), intent(in) :: method(7), mthbc(6)
       type(c_ptr), value :: log_dt, log_cfl, log_acc
       integer(c_int), value :: maxlog
       integer(c_int), intent(out) :: nlog
     end function magic_claw3

     integer(c_int) function magic_error_location(ctx, sweep, i, j, k) &
          bind(C, name="magic_error_location")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: sweep, i, j, k
     end function magic_error_location
  end interface

contains

  ! Stop like the reference does on a fatal condition, with the C status text.
  subroutine magic_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == MAGIC_OK) return
    call c_f_pointer(magic_strerror(status), msg, [256])
    n = 0
    do while (msg(n + 1) /= c_null_char .and. n < 256)
       n = n + 1
    end do
    write(6, '(3A,I0,3A)') '*** ', where, ': status ', status, ' (', &
         transfer(msg(1:n), repeat(' ', n)), ')'
    stop 1
  end subroutine magic_check

end module magic_hip
