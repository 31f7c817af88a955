/*
 * magic_h5.h -- HDF5 frame output with the layout of the reference's out3
 * (mpilib/out3_mpi_h5sep.f90:2-195), single rank.  Host-only library
 * (libmagic_h5.so, HDF5 C API); independent of the GPU library.
 */
#ifndef MAGIC_H5_H
#define MAGIC_H5_H

#ifdef __cplusplus
extern "C" {
#endif

#define MAGIC_H5_OK    0
#define MAGIC_H5_EARG  5   /* invalid argument */
#define MAGIC_H5_EIO   8   /* HDF5 call failed */

/* "fort.qq<iframe:4>id<id:4>.h5" (out3_mpi_h5sep.f90:69-79); n >= 24. */
int magic_h5_frame_name(int iframe, int id, char *out, int n);

/* out3(meqn,mbc,mx,my,mz,xlower,ylower,zlower,dx,dy,dz,q,t,iframe)
 * (out3_mpi_h5sep.f90:2-3): writes the interior of q (reference layout,
 * ghost cells included, FP64) to <dir>/fort.qq<iframe>id0000.h5 (dir NULL or
 * "" = current directory): dataset "Pid", Fortran dims (meqn, mx, my, mz),
 * chunk (1, mx, my, mz), gzip 6, early allocation, attribute
 * "Parameters"[25] as out3_mpi_h5sep.f90:151-175 with lx = ly = lz = 1. */
int magic_out3_h5(const char *dir, int meqn, int mbc, int mx, int my, int mz,
                  double xlower, double ylower, double zlower,
                  double dx, double dy, double dz, const double *q, double t, int iframe);

/* Read a frame back: C-order dims {mz, my, mx, meqn}, chunk dims, gzip
 * level (-1 if none), the 25 parameters and the data (C order, ndata
 * doubles available).  Any output pointer may be NULL. */
int magic_in3_h5(const char *path, int *dims4, int *chunk4, int *deflate_level,
                 double *params25, double *data, long long ndata);

#ifdef __cplusplus
}
#endif
#endif
