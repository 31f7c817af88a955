! Drop-in Fortran operators on the MI355X device path.
!
! claw3, step3, bc3, b4step3 and src3 below keep the reference's exact dummy
! argument lists (mpilib/claw3_mpi.f:4-7, lib/step3.f:3-7, bc3_mpi.f:5-6,
! b4step3.f:2-3, src3.f90:203-205), so a caller written for the reference
! links against them unchanged: one that keeps its own claw3 and calls
! these operators, or one that calls claw3 itself.  They run on the GPU through the
! C ABI (module magic_hip): the device context for the grid is created on the
! first call, aux is uploaded whenever a different aux array is passed, and q
! crosses to the device on entry and back on exit of each call.  The
! rpn3/rpt3/rptt3/bc3/src3/b4step3 procedure arguments are accepted and not
! called: the device kernels carry the same algebra (rpn3euX.f, rpt3euX.f,
! rptt3euX.f, bc3_mpi.f branches 1-3, src3.f90, b4step3.f).
!
! Problem parameters come from the setprob commons (setprob.f:17-25), which
! the host fills (magic_setprob below reads setprob.data in the order of
! setprob.f:31-61).
module magic_hip_state
  use iso_c_binding
  use magic_hip
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: dims(6) = 0
  ! xlower, ylower, zlower, dx, dy, dz of the context; the lower edges are
  ! unknown while only step3 (whose arguments do not carry them) has run
  double precision, save :: geom(6) = 0.d0
  logical, save :: lowers_known = .false.
  integer(c_intptr_t), save :: aux_addr = 0
contains

  subroutine load_params()
    implicit double precision (a-h, o-z)
    common /param/ gamma, gamma1, grav, dens, pres, speed, dtherm0, dmolec0
    common /forcing/ omega, amplitude, propx, propz, w_n, w_o, vsrcx, forcemth
    common /wind/ omega1, amp0, amp1, prop1z, z1width, z1pos, z1Gpos, phi, windscale1
    common /gauss/ xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth
    integer :: loadprof, nspecadv
    common /options/ loadprof, nspecadv
    type(magic_params) :: p
    p%loadprof = loadprof; p%nspecadv = nspecadv
    p%gamma = gamma; p%grav = grav; p%dens = dens; p%pres = pres
    p%dmolec0 = dmolec0; p%dtherm0 = dtherm0
    p%forcemth = forcemth; p%omega = omega; p%amplitude = amplitude
    p%propx = propx; p%propz = propz
    p%xpos = xpos; p%ypos = ypos; p%zpos = zpos; p%xwidth = xwidth; p%ywidth = ywidth
    p%zwidth = zwidth; p%tcenter = tcenter; p%twidth = twidth; p%vsrcx = vsrcx
    p%amp0 = amp0; p%amp1 = amp1; p%omega1 = omega1; p%prop1z = prop1z; p%phi = phi
    p%z1width = z1width; p%z1pos = z1pos; p%z1Gpos = z1Gpos; p%windscale1 = windscale1
    call magic_check(magic_set_params(ctx, p), 'magic_set_params')
  end subroutine load_params

  ! Create (or reuse) the device context of this grid and make sure the
  ! device holds this aux array.  known: the caller's arguments carry the
  ! grid's lower edges (every operator but step3); a context created without
  ! them is created again once they are known.  The state q is not kept
  ! between calls, so a new context loses nothing.
  subroutine ensure_ctx(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz, aux, known)
    integer, intent(in) :: mx, my, mz, meqn, mbc, maux
    double precision, intent(in) :: xlower, ylower, zlower, dx, dy, dz
    double precision, intent(in), target :: aux(*)
    logical, intent(in) :: known
    integer :: d(6)
    logical :: same
    d = (/mx, my, mz, meqn, mbc, maux/)
    if (c_associated(ctx)) then
       same = all(d == dims) .and. geom(4) == dx .and. geom(5) == dy .and. geom(6) == dz
       if (known) same = same .and. lowers_known .and. geom(1) == xlower .and. geom(2) == ylower &
            .and. geom(3) == zlower
       if (.not. same) then
          call magic_check(magic_destroy(ctx), 'magic_destroy')
          ctx = c_null_ptr
       end if
    end if
    if (.not. c_associated(ctx)) then
       call magic_check(magic_create(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, &
            dx, dy, dz, ctx), 'magic_create')
       dims = d
       geom = (/xlower, ylower, zlower, dx, dy, dz/)
       lowers_known = known
       aux_addr = 0
    end if
    call load_params()
    if (transfer(c_loc(aux(1)), aux_addr) /= aux_addr .or. aux_addr == 0) then
       call magic_check(magic_set_aux(ctx, aux), 'magic_set_aux')
       aux_addr = transfer(c_loc(aux(1)), aux_addr)
    end if
  end subroutine ensure_ctx

end module magic_hip_state


! claw3: mpilib/claw3_mpi.f:4-7 (same arguments, same meaning).
subroutine claw3(meqn, mwaves, mbc, mx, my, mz, maux, q, aux, xlower, ylower, zlower, &
     dx, dy, dz, tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, mwork, &
     use_fwave, info, bc3, rpn3, rpt3, rptt3, src3, b4step3)
  use magic_hip_state
  implicit double precision (a-h, o-z)
  external bc3, rpn3, rpt3, rptt3, src3, b4step3
  dimension q(*), aux(*), work(*)
  dimension dtv(5), cflv(4), nv(2), method(7), mthlim(mwaves), mthbc(6)
  logical use_fwave
  integer(c_int) :: st, nlog
  if (.not. use_fwave .or. mwaves /= 3 .or. any(mthlim(1:mwaves) > 0)) then
     write(6, *) 'claw3: the device path is the f-wave solver with rpn3 limiting only'
     stop 1
  end if
  call ensure_ctx(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz, aux, .true.)
  call magic_check(magic_set_q(ctx, q), 'magic_set_q')
  st = magic_claw3(ctx, tstart, tend, dtv, cflv, nv, method, mthbc, c_null_ptr, c_null_ptr, &
       c_null_ptr, 0, nlog)
  if (st == MAGIC_ERHO .or. st == MAGIC_EA2) call report_location()
  call magic_check(st, 'claw3')
  call magic_check(magic_get_q(ctx, q), 'magic_get_q')
  info = 0
contains
  subroutine report_location()
    integer(c_int) :: s, i, j, k
    call magic_check(magic_error_location(ctx, s, i, j, k), 'magic_error_location')
    write(6, '(A,I0,A,3I6)') '*** failure in sweep ', s, ' at ', i, j, k
  end subroutine report_location
end subroutine claw3


! step3: lib/step3.f:3-7 (same arguments).  qold and qnew must be equal on
! entry, as the reference requires; qnew returns the stepped state.
subroutine step3(maxm, meqn, mwaves, mbc, mx, my, mz, qold, qnew, aux, dx, dy, dz, dt, &
     method, mthlim, cfl, qadd, fadd, gadd, hadd, q1d, dtdx1d, dtdy1d, dtdz1d, &
     aux1, aux2, aux3, maux, work, mwork, rpn3, rpt3, rptt3, use_fwave)
  use magic_hip_state
  implicit double precision (a-h, o-z)
  external rpn3, rpt3, rptt3
  dimension qold(*), qnew(*), aux(*), method(7), mthlim(mwaves)
  dimension qadd(*), fadd(*), gadd(*), hadd(*), q1d(*), dtdx1d(*), dtdy1d(*), dtdz1d(*)
  dimension aux1(*), aux2(*), aux3(*), work(*)
  logical use_fwave
  real(c_double) :: c
  call ensure_ctx(mx, my, mz, meqn, mbc, maux, 0.d0, 0.d0, 0.d0, dx, dy, dz, aux, .false.)
  call magic_check(magic_set_q(ctx, qold), 'magic_set_q')
  call magic_check(magic_step3(ctx, dt, c), 'step3')
  call magic_check(magic_get_q(ctx, qnew), 'magic_get_q')
  cfl = c
end subroutine step3


! bc3: bc3_mpi.f:5-6 (same arguments): the ghost layers of q, physical
! branches 1-3 of every face (mthbc 0 or 4 stop: they need MPI or a user
! routine, bc3_mpi.f:100-105).
subroutine bc3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, maux, aux, t, dt, mthbc)
  use magic_hip_state
  implicit double precision (a-h, o-z)
  dimension q(*), aux(*), mthbc(6)
  call ensure_ctx(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz, aux, .true.)
  call magic_check(magic_set_q(ctx, q), 'magic_set_q')
  call magic_check(magic_bc3(ctx, mthbc), 'bc3')
  call magic_check(magic_get_q(ctx, q), 'magic_get_q')
end subroutine bc3


! b4step3: b4step3.f:2-3 (same arguments): tracers 6..meqn clamped to [0, 1]
! on every cell.
subroutine b4step3(mbc, mx, my, mz, meqn, q, xlower, ylower, zlower, dx, dy, dz, t, dt, maux, aux)
  use magic_hip_state
  implicit double precision (a-h, o-z)
  dimension q(*), aux(*)
  call ensure_ctx(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz, aux, .true.)
  call magic_check(magic_set_q(ctx, q), 'magic_set_q')
  call magic_check(magic_b4step3(ctx), 'b4step3')
  call magic_check(magic_get_q(ctx, q), 'magic_get_q')
end subroutine b4step3


! src3: src3.f90:203-205 (same arguments): the source pass with diffuse.f
! and conduct.f substeps and the plume forcing, evaluated at t over dt.
! Prints the substep line the reference prints (src3.f90:370-374).
subroutine src3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, maux, aux, t, dt, mthbc)
  use magic_hip_state
  implicit double precision (a-h, o-z)
  dimension q(*), mthbc(6)
  dimension aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)
  integer(c_int) :: ndtr, ndtrc
  call ensure_ctx(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, dx, dy, dz, aux, .true.)
  call magic_check(magic_set_q(ctx, q), 'magic_set_q')
  call magic_check(magic_src3(ctx, t, dt, mthbc, ndtr, ndtrc), 'src3')
  call magic_check(magic_get_q(ctx, q), 'magic_get_q')
  ! the Courant numbers of the substeps, as src3.f90:313-325 forms them
  dxyzmin2 = min(dx, dy, dz)**2
  difcfl = 0.d0
  concfl = 0.d0
  if (ndtr > 0) difcfl = (4.d0/3.d0)*aux(11, 1, 1, mz)*(dt/ndtr)/dxyzmin2
  if (ndtrc > 0) concfl = aux(12, 1, 1, mz)*((dt/ndtr)/ndtrc)/dxyzmin2
  write(6, 601) ndtr, ndtrc, difcfl, concfl
601 format('SRC3...  Vis. Steps', i5, '  Con. Sub-steps', i5, ' VisCFL =', f6.3, ' ConCFL =', f6.3)
end subroutine src3


! setprob.data reader: fills the commons in the order of setprob.f:31-61 and
! derives speed, w_n, w_o for loadprof = 0 (setprob.f:63-68).
subroutine magic_setprob(fname)
  implicit double precision (a-h, o-z)
  character(len=*), intent(in) :: fname
  common /param/ gamma, gamma1, grav, dens, pres, speed, dtherm0, dmolec0
  common /forcing/ omega, amplitude, propx, propz, w_n, w_o, vsrcx, forcemth
  common /wind/ omega1, amp0, amp1, prop1z, z1width, z1pos, z1Gpos, phi, windscale1
  common /gauss/ xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth
  common /options/ loadprof, nspecadv
  open(unit=7, file=fname, status='old', form='formatted')
  read(7, *) loadprof
  read(7, *) nspecadv
  read(7, *) gamma
  read(7, *) grav
  read(7, *) dens
  read(7, *) pres
  read(7, *) dmolec0
  read(7, *) dtherm0
  read(7, *) forcemth
  read(7, *) omega
  read(7, *) amplitude
  read(7, *) propx
  read(7, *) propz
  read(7, *) xpos
  read(7, *) ypos
  read(7, *) zpos
  read(7, *) xwidth
  read(7, *) ywidth
  read(7, *) zwidth
  read(7, *) tcenter
  read(7, *) twidth
  read(7, *) vsrcx
  read(7, *) amp0
  read(7, *) amp1
  read(7, *) omega1
  read(7, *) prop1z
  read(7, *) phi
  read(7, *) z1width
  read(7, *) z1pos
  read(7, *) z1Gpos
  read(7, *) windscale1
  close(7)
  gamma1 = gamma - 1.d0
  speed = dsqrt(pres*gamma/dens)
  w_n = dsqrt(((grav**2)/(speed**2))*(gamma1))
  w_o = dsqrt((gamma**2)*(grav**2)/(4*(speed**2)))
end subroutine magic_setprob
