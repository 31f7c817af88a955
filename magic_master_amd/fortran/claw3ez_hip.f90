! claw3ez_hip: Fortran host program of the MI355X MAGIC step.
!
! Reads claw3ez.data and setprob.data in the reference's formats
! (mpilib/claw3ez_mpi_driver.f:81-145, setprob.f:31-61), builds the FP64
! background aux and q = aux (qinit.f:74-82), and runs the claw3ez frame loop
! (mpilib/claw3ez_mpi.f:136-188) with q resident on the GPU: each claw3 call
! goes straight to magic_claw3, and q comes back to the host only at output
! frames.  One rank, one GPU: lx = ly = lz must be 1 and mthbc 4 is rejected.
!
! Backgrounds (setprob loadprof):
!   0  isothermal, rho = dens*exp(-z*gamma*grav/speed**2), R = 287
!      (the formulas of setaux.f:55-87)
!   1  the vertical profile of profile.data in the working directory, as
!      setaux.f:38-212 builds it (with its analytic x wind when amp0 or
!      amp1 is set)
!   2  the standard temperature column integrated hydrostatically
!      (same as magic_master_amd/problem.py:setaux_column)
! Frames are written by out3 as the reference's HDF5 files
! fort.qqXXXXid0000.h5 (mpilib/out3_mpi_h5sep.f90; magic_master_amd/fortran/out3_h5.f90).
program claw3ez_hip
  use iso_c_binding
  use magic_hip
  implicit double precision (a-h, o-z)
  common /param/ gamma, gamma1, grav, dens, pres, speed, dtherm0, dmolec0
  common /options/ loadprof, nspecadv
  double precision, allocatable :: q(:, :, :, :), aux(:, :, :, :), tout(:)
  integer :: method(7), mthlim(3), mthbc(6), nv(2)
  double precision :: dtv(5), cflv(4)
  integer(c_int) :: nlog
  type(c_ptr) :: ctx
  integer(8) :: c0, c1, crate

  open(55, file='claw3ez.data', status='old', form='formatted')
  read(55, *) mx
  read(55, *) my
  read(55, *) mz
  read(55, *) nout
  read(55, *) outstyle_i
  allocate(tout(max(nout, 1)))
  tfinal = 0.d0
  nstepout = 1
  if (outstyle_i == 1) then
     read(55, *) tfinal
  else if (outstyle_i == 2) then
     read(55, *) (tout(i), i = 1, nout)
  else
     read(55, *) nstepout, nstop
     nout = nstop
  end if
  read(55, *) dtv(1)
  read(55, *) dtv(2)
  read(55, *) cflv(1)
  read(55, *) cflv(2)
  read(55, *) nv(1)
  do i = 1, 7
     read(55, *) method(i)
  end do
  maux = method(7)
  read(55, *) meqn
  read(55, *) mwaves
  read(55, *) (mthlim(i), i = 1, mwaves)
  read(55, *) t0
  read(55, *) xlower
  read(55, *) xupper
  read(55, *) ylower
  read(55, *) yupper
  read(55, *) zlower
  read(55, *) zupper
  read(55, *) mbc
  do i = 1, 6
     read(55, *) mthbc(i)
  end do
  read(55, *) lx
  read(55, *) ly
  read(55, *) lz
  close(55)
  if (lx * ly * lz /= 1) then
     write(6, *) 'claw3ez_hip: one GPU holds the whole domain; set lx = ly = lz = 1'
     stop 1
  end if

  call magic_setprob('setprob.data')
  dx = (xupper - xlower) / float(mx)
  dy = (yupper - ylower) / float(my)
  dz = (zupper - zlower) / float(mz)
  allocate(q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc))
  allocate(aux(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc))
  call background(aux)
  do k = 1-mbc, mz+mbc
     do j = 1-mbc, my+mbc
        do i = 1-mbc, mx+mbc
           q(1:meqn, i, j, k) = aux(1:meqn, i, j, k)
        end do
     end do
  end do

  call magic_check(magic_create(mx, my, mz, meqn, mbc, maux, xlower, ylower, zlower, &
       dx, dy, dz, ctx), 'magic_create')
  call push_params()
  call magic_check(magic_set_aux(ctx, aux), 'magic_set_aux')
  call magic_check(magic_set_q(ctx, q), 'magic_set_q')
  call write_frame(0, t0)

  call system_clock(c0, crate)
  tend = t0
  do n = 1, nout
     tstart = tend
     if (outstyle_i == 1) tend = tstart + (tfinal - t0) / float(nout)
     if (outstyle_i == 2) tend = tout(n)
     if (outstyle_i == 3) tend = tstart - 1.d0
     nv(2) = 0
     ist = magic_claw3(ctx, tstart, tend, dtv, cflv, nv, method, mthbc, c_null_ptr, &
          c_null_ptr, c_null_ptr, 0, nlog)
     call magic_check(ist, 'claw3')
     dtv(1) = dtv(5)
     iframe = n / nstepout
     if (iframe * nstepout == n) then
        call magic_check(magic_get_q(ctx, q), 'magic_get_q')
        call write_frame(iframe, tend)
        write(6, '(A,I4,A,ES12.4)') 'CLAW3EZ: Frame ', iframe, ' done at time t = ', tend
     end if
  end do
  call system_clock(c1)
  write(6, '(A,F10.3)') 'Elapsed time = ', dble(c1 - c0) / dble(crate)
  call magic_check(magic_destroy(ctx), 'magic_destroy')

contains

  subroutine push_params()
    implicit double precision (a-h, o-z)
    common /forcing/ omega, amplitude, propx, propz, w_n, w_o, vsrcx, forcemth
    common /wind/ omega1, amp0, amp1, prop1z, z1width, z1pos, z1Gpos, phi, windscale1
    common /gauss/ xpos, ypos, zpos, xwidth, ywidth, zwidth, tcenter, twidth
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
  end subroutine push_params

  double precision function tcol(z)
    double precision, intent(in) :: z
    double precision :: zk, t11, t47, th0
    zk = z / 1000.d0
    t11 = 288.d0 - 6.5d0 * 11.d0
    if (zk <= 11.d0) then
       tcol = 288.d0 - 6.5d0 * zk
    else if (zk <= 20.d0) then
       tcol = t11
    else if (zk <= 47.d0) then
       tcol = t11 + (zk - 20.d0)
    else
       t47 = t11 + 27.d0
       th0 = tanh((47.d0 - 110.d0) / 20.d0)
       tcol = t47 + (1000.d0 - t47) * (tanh((zk - 110.d0) / 20.d0) - th0) / (1.d0 - th0)
    end if
  end function tcol

  subroutine background(a)
    double precision, intent(out) :: a(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)
    double precision :: rho(1-mbc:mz+mbc), e(1-mbc:mz+mbc), vg(1-mbc:mz+mbc)
    double precision :: nu(1-mbc:mz+mbc), kap(1-mbc:mz+mbc), p(1-mbc:mz+mbc), tt(1-mbc:mz+mbc)
    double precision :: rho1, z, pl, pr, aa, dz0
    a = 0.d0
    if (loadprof == 1) then
       call profile_background(a)
       return
    end if
    if (loadprof == 0) then
       rho1 = dens * dexp(-(zlower + 0.5d0 * dz) * gamma * grav / (speed**2))
       do k = 1-mbc, mz+mbc
          z = zlower + (k - 0.5d0) * dz
          rho(k) = dens * dexp(-z * gamma * grav / (speed**2))
          e(k) = pres * dexp(-z * gamma * grav / (speed**2)) / gamma1
          nu(k) = dmolec0 * (rho1 / rho(k))
          kap(k) = dtherm0 * (rho1 / rho(k)) * gamma
       end do
       rgas = 287.d0
    else if (loadprof == 2) then
       rgas = 287.d0
       aa = grav * dz / (2.d0 * rgas)
       do k = 1-mbc, mz+mbc
          tt(k) = tcol(zlower + (k - 0.5d0) * dz)
       end do
       dz0 = zlower + 0.5d0 * dz
       p(1) = 101325.d0 * (1.d0 - grav * dz0 / (2.d0 * rgas * tcol(0.d0))) / &
            (1.d0 + grav * dz0 / (2.d0 * rgas * tt(1)))
       do k = 2, mz+mbc
          p(k) = p(k-1) * (1.d0 - aa / tt(k-1)) / (1.d0 + aa / tt(k))
       end do
       do k = 0, 1-mbc, -1
          p(k) = p(k+1) * (1.d0 + aa / tt(k+1)) / (1.d0 - aa / tt(k))
       end do
       do k = 1-mbc, mz+mbc
          rho(k) = p(k) / (rgas * tt(k))
          e(k) = p(k) / (gamma - 1.d0)
          nu(k) = 1.8d-5 / rho(k)
          kap(k) = gamma * 1.8d-5 / (0.7d0 * rho(k))
       end do
    else
       write(6, *) 'claw3ez_hip: loadprof must be 0 (isothermal), 1 (profile.data) or 2 (standard column)'
       stop 1
    end if
    ! virtual gravity (setaux.f:176-187)
    vg = 0.d0
    do k = 2-mbc, mz+mbc
       pl = (gamma - 1.d0) * e(k-1)
       pr = (gamma - 1.d0) * e(k)
       vg(k) = (pr - pl) / (.5d0 * (rho(k) + rho(k-1)))
    end do
    vg(1-mbc) = vg(2-mbc)
    do k = 1-mbc, mz+mbc
       do j = 1-mbc, my+mbc
          do i = 1-mbc, mx+mbc
             a(1, i, j, k) = rho(k)
             a(5, i, j, k) = e(k)
             a(11, i, j, k) = nu(k)
             a(12, i, j, k) = kap(k)
             a(17, i, j, k) = vg(k)
             a(18, i, j, k) = rgas
             a(19, i, j, k) = gamma
          end do
       end do
    end do
  end subroutine background

  ! loadprof = 1 (setaux.f:38-212): for each cell
  ! centre height the profile.data row with exactly that height in km (rows
  ! read in order; a missing height stops the run), columns height, O, N2,
  ! O2 number densities (CGS), density, temperature, H and electron number
  ! densities, x and y wind.  Gas constant and gamma from the composition,
  ! density from the number densities, energy from the temperature,
  ! composition ratios (slots 6-10), molecular and thermal diffusivities
  ! (slots 11, 12; layers above the 400th take the 400th's) and the virtual
  ! gravity from the pressure jumps (slot 17; 0 on the lowest ghost layer).
  ! With amp0 or amp1 set, the analytic x wind replaces the profile's winds
  ! (setaux.f:128-134): rho*wind_x + amp0*rho (amp0 enters twice, as in the
  ! reference) and a y slot of 0.
  ! The literals keep the reference's kinds (3.43, 0.69, ... are default
  ! REAL), so the values are the reference's bit for bit.
  subroutine profile_background(a)
    double precision, intent(inout) :: a(maux, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)
    common /wind/ omega1, amp0, amp1, prop1z, z1width, z1pos, z1Gpos, phi, windscale1
    double precision :: col(19, 1-mbc:mz+mbc)
    double precision :: height, dox, dnit2, dox2, rhof, temp, dhyd, dne, windx, windy
    double precision :: zc, rho, gm1, cnv, cp, dmolec, dtherm, pl, pr, scale
    integer :: k, ios
    col = 0.d0
    cnv = 1.d3                  ! profile.data in CGS (setaux.f:20-21, 111-116)
    open(7, file='profile.data', status='old', form='formatted')
    do k = 1-mbc, mz+mbc
       zc = zlower + (k - 0.5d0) * dz
       do
          read(7, *, iostat=ios) height, dox, dnit2, dox2, rhof, temp, dhyd, dne, windx, windy
          if (ios /= 0) then
             write(6, *) 'claw3ez_hip: profile.data has no row at height', zc / 1.d3, 'km'
             stop 1
          end if
          if (height >= zc / 1.d3) exit
       end do
       if (height > zc / 1.d3) then
          write(6, *) 'claw3ez_hip: profile.data height', height, '.GT.', zc / 1.d3
          stop 1
       end if
       col(18, k) = 8.31d3 / ((dox*16.d0 + dnit2*28.d0 + dox2*32.d0) / (dox + dnit2 + dox2))
       col(19, k) = (7*(dox2 + dnit2) + 5*dox) / (5*(dox2 + dnit2) + 3*dox)
       gm1 = col(19, k) - 1.d0
       rho = (dox*16.d0 + dox2*32.d0 + dnit2*28.d0) * cnv
       rho = rho / (6.022d23)
       col(1, k) = rho
       if (amp0 == 0.d0 .and. amp1 == 0.d0) then
          col(2, k) = rho * windx
          col(3, k) = rho * windy
       else
          ! col(1, 1) is still 0 in the ghost layers below k = 1, where the
          ! reference reads aux(1,i,j,1) before setting it (a zeroed array)
          scale = (col(1, 1) / rho)**0.5d0
          col(2, k) = rho * wind_x(zc, scale)
          col(2, k) = col(2, k) + amp0 * rho
          col(3, k) = 0.d0
       end if
       col(4, k) = 0.d0
       col(5, k) = rho * temp * col(18, k) / gm1 + .5*(col(2, k)**2) / rho + .5*(col(3, k)**2) / rho
       col(6, k) = dox*16.d0*cnv / (6.022d23*rho)
       col(7, k) = dox2*32.d0*cnv / (6.022d23*rho)
       col(8, k) = dne / (dox + dox2 + dnit2)
       col(9, k) = dhyd / (dox + dox2 + dnit2)
       col(10, k) = 0.d0
       cp = 1/(1 - 1/col(19, k)) * col(18, k)
       dmolec = (3.43*dnit2 + 4.03*dox2 + 3.90*dox) * 1d-7
       dmolec = dmolec * (temp**(0.69)) / (dox + dnit2 + dox2)
       dtherm = (56.0*(dnit2 + dox2) + 75.9*dox) * 1d-5
       dtherm = dtherm * (temp**(0.69)) / (dox + dnit2 + dox2)
       col(11, k) = dmolec * (1/rho)
       col(12, k) = dtherm * (1/(rho*cp)) * col(19, k)
       if (k > 400) col(11:12, k) = col(11:12, 400)
       if (k > 1-mbc) then
          pl = (col(19, k-1) - 1.d0) * (col(5, k-1) - 0.5d0*(col(2, k-1)**2 + col(3, k-1)**2 &
               + col(4, k-1)**2) / col(1, k-1))
          pr = (col(19, k) - 1.d0) * (col(5, k) - 0.5d0*(col(2, k)**2 + col(3, k)**2 &
               + col(4, k)**2) / col(1, k))
          col(17, k) = (pr - pl) / (.5d0*(col(1, k) + col(1, k-1)))
       end if
    end do
    close(7)
    do k = 1-mbc, mz+mbc
       do j = 1-mbc, my+mbc
          do i = 1-mbc, mx+mbc
             a(1:19, i, j, k) = col(:, k)
          end do
       end do
    end do
  end subroutine profile_background

  ! The analytic background wind in x (fcns.f:100-116) at t = 0, which is
  ! what setaux asks for (it passes the integer constant 0 for t).
  double precision function wind_x(z, scale)
    double precision, intent(in) :: z, scale
    common /wind/ omega1, amp0, amp1, prop1z, z1width, z1pos, z1Gpos, phi, windscale1
    wind_x = amp1
    wind_x = wind_x * cos(prop1z * (z - z1pos) - omega1 * 0.d0 - phi)
    wind_x = wind_x * exp(-.5d0 * (((z - z1Gpos)**2) / z1width**2))
    if (windscale1 == 1.d0) wind_x = wind_x * scale
    wind_x = wind_x + amp0
  end function wind_x

  subroutine write_frame(iframe, t)
    integer, intent(in) :: iframe
    double precision, intent(in) :: t
    call out3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, t, iframe)
  end subroutine write_frame

end program claw3ez_hip
