! Link-level check of the drop-in claw3: calls claw3 with the reference's
! argument list and EXTERNAL operators exactly as mpilib/claw3ez_mpi.f:142-145
! does, on q/aux read from raw files, and writes q back.
!
!   dropin_check <dims file> <aux file> <q in> <q out>
! dims file (text): mx my mz meqn mbc maux / xlower ylower zlower dx dy dz /
!                   nsteps / mthbc(6) / setprob.data path
program dropin_check
  implicit double precision (a-h, o-z)
  external bc3, rpn3, rpt3, rptt3, src3, b4step3
  double precision, allocatable :: q(:), aux(:), work(:)
  integer :: method(7), mthlim(3), mthbc(6), nv(2)
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
  read(10, *) nsteps
  read(10, *) mthbc
  read(10, '(A)') fprob
  close(10)
  call magic_setprob(trim(fprob))
  ncell = (mx + 2*mbc) * (my + 2*mbc) * (mz + 2*mbc)
  allocate(q(meqn * ncell), aux(maux * ncell), work(1))
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
  t = 0.d0
  do n = 1, nsteps
     tstart = t
     tend = tstart - 1.d0
     nv = (/10000, 0/)
     call claw3(meqn, 3, mbc, mx, my, mz, maux, q, aux, xlower, ylower, zlower, dx, dy, dz, &
          tstart, tend, dtv, cflv, nv, method, mthlim, mthbc, work, 1, use_fwave, info, &
          bc3, rpn3, rpt3, rptt3, src3, b4step3)
     t = tend
     dtv(1) = dtv(5)
  end do
  open(13, file=fqout, access='stream', form='unformatted', status='replace')
  write(13) q
  close(13)
  write(6, '(A,ES24.16,A,ES24.16)') 't = ', t, ' next dt = ', dtv(5)
end program dropin_check

! The Riemann operators the reference passes as EXTERNAL; the device claw3
! does not call them (bc3, src3 and b4step3 come from the drop-in library).
subroutine rpn3()
end subroutine rpn3
subroutine rpt3()
end subroutine rpt3
subroutine rptt3()
end subroutine rptt3
