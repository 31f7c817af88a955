! out3 with the reference's argument list (mpilib/out3_mpi_h5sep.f90:2-3),
! writing the HDF5 frame fort.qq<iframe>id0000.h5 through the C writer of
! include/magic_h5.h (same dataset, chunking, compression and attribute).
subroutine out3(meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, t, iframe)
  use iso_c_binding
  implicit none
  integer, intent(in) :: meqn, mbc, mx, my, mz, iframe
  double precision, intent(in) :: xlower, ylower, zlower, dx, dy, dz, t
  double precision, intent(in) :: q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)
  interface
     integer(c_int) function magic_out3_h5(dir, meqn, mbc, mx, my, mz, xlower, ylower, zlower, &
          dx, dy, dz, q, t, iframe) bind(C, name="magic_out3_h5")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: dir
       integer(c_int), value :: meqn, mbc, mx, my, mz, iframe
       real(c_double), value :: xlower, ylower, zlower, dx, dy, dz, t
       real(c_double), intent(in) :: q(*)
     end function magic_out3_h5
  end interface
  integer(c_int) :: ist
  ist = magic_out3_h5(c_null_ptr, meqn, mbc, mx, my, mz, xlower, ylower, zlower, dx, dy, dz, q, t, iframe)
  if (ist /= 0) then
     write(6, *) 'out3: HDF5 write failed, status ', ist
     stop 1
  end if
end subroutine out3
