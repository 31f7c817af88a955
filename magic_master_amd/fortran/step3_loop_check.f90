! Link-level check of the drop-in operators step3, bc3, b4step3 and src3 for
! a caller that keeps its own time integrator: the step control of
! claw3 (mpilib/claw3_mpi.f:736-927, method (1, 2, 22, 0, 1, 0)) written out
! below, calling the device operators through the reference's argument
! lists (lib/step3.f:3-7, bc3_mpi.f:5-6, b4step3.f:2-3, src3.f90:203-205)
! with q and the copyq3 backup in host memory.
!
!   step3_loop_check <dims file> <aux file> <q in> <q out>
! dims file (text): mx my mz meqn mbc maux / xlower ylower zlower dx dy dz /
!                   nsteps first_dt / mthbc(6) / setprob.data path
! Prints one line per attempted step: n, dt, cfl, accepted.
program step3_loop_check
  implicit double precision (a-h, o-z)
  external rpn3, rpt3, rptt3
  double precision, allocatable :: q(:), qold(:), aux(:), work(:)
  integer :: method(7), mthlim(3), mthbc(6)
  double precision :: dtv(5), cflv(4)
  character(len=512) :: fdims, faux, fqin, fqout, fprob
  logical :: use_fwave
  call get_command_argument(1, fdims)
  call get_command_argument(2, faux)
  call get_command_argument(3, fqin)
  call get_command_argument(4, fqout)
  open(10, file=fdims, status='old')
  read(10, *) mx, my, mz, meqn, mbc, maux
  read(10, *) xlower, ylower, zlower, dx, dy, dz
  read(10, *) nsteps, dt0
  read(10, *) mthbc
  read(10, '(A)') fprob
  close(10)
  call magic_setprob(trim(fprob))
  ncell = (mx + 2*mbc) * (my + 2*mbc) * (mz + 2*mbc)
  allocate(q(meqn * ncell), qold(meqn * ncell), aux(maux * ncell), work(1))
  open(11, file=faux, access='stream', form='unformatted', status='old')
  read(11) aux
  close(11)
  open(12, file=fqin, access='stream', form='unformatted', status='old')
  read(12) q
  close(12)
  method = (/1, 2, 22, 0, 1, 0, maux/)
  mthlim = 0
  dtv = (/1.d0, 10.d0, 0.d0, 0.d0, 0.d0/)
  cflv = (/0.95d0, 0.9d0, 0.d0, 0.d0/)
  use_fwave = .true.
  maxm = max(mx, my, mz)
  t = 0.d0
  dtv(1) = dt0
  dt = dtv(1)
  do n = 1, nsteps
     told = t
     do
        call bc3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, maux, aux, told, dt, mthbc)
        call b4step3(mbc, mx, my, mz, meqn, q, xlower, ylower, zlower, dx, dy, dz, told, dt, maux, aux)
        qold = q                                            ! copyq3, claw3_mpi.f:807
        call step3(maxm, meqn, 3, mbc, mx, my, mz, qold, q, aux, dx, dy, dz, dt, method, mthlim, cfl, &
             work, work, work, work, work, work, work, work, work, work, work, maux, work, 1, &
             rpn3, rpt3, rptt3, use_fwave)
        t = told + dt
        write(6, '(I4,2ES24.16,L2)') n, dt, cfl, cfl <= cflv(1)
        if (cfl <= cflv(1)) exit                            ! claw3_mpi.f:861
        t = told                                            ! rejected: claw3_mpi.f:866-880
        q = qold
        dt = min(dtv(2), dt * cflv(2) / cfl)
     end do
     call src3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, maux, aux, told, dt, mthbc)
     if (cfl == 0.d0) then                                  ! claw3_mpi.f:910-919
        dt = dtv(2)
     else
        dt = min(dtv(2), dt * cflv(2) / cfl)
     end if
  end do
  open(13, file=fqout, access='stream', form='unformatted', status='replace')
  write(13) q
  close(13)
  write(6, '(A,ES24.16,A,ES24.16)') 't = ', t, ' next dt = ', dt
end program step3_loop_check

! The Riemann operators the reference passes to step3 as EXTERNAL; the
! device step3 does not call them.
subroutine rpn3()
end subroutine rpn3
subroutine rpt3()
end subroutine rpt3
subroutine rptt3()
end subroutine rptt3
