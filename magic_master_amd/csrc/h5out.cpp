// HDF5 frame writer and reader with the layout of the reference's
// out3 (mpilib/out3_mpi_h5sep.f90:2-195), for one rank.
//
// File  fort.qq<frame:4>id<rank:4>.h5, one dataset "Pid" of native doubles
// with Fortran dimensions (meqn, mx, my, mz) -- C dimensions
// {mz, my, mx, meqn} -- holding the interior of q, chunked one component at
// a time (Fortran chunk (1, mx, my, mz)), gzip level 6, space allocated
// early, and a 25-double attribute "Parameters" on the dataset
// (out3_mpi_h5sep.f90:151-175).  The reference writes q(1:10, ...) whatever
// meqn is (out3_mpi_h5sep.f90:191), which equals q(1:meqn, ...) only for
// meqn = 10; this writer writes q(1:meqn, ...).
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/magic_h5.h"

namespace {

struct H5Closer {
    hid_t id;
    herr_t (*fn)(hid_t);
    ~H5Closer()
    {
        if (id >= 0) fn(id);
    }
};

void frame_name(char *out, size_t n, int iframe, int id)
{
    snprintf(out, n, "fort.qq%04did%04d.h5", iframe % 10000, id % 10000);
}

}  // namespace

extern "C" int magic_h5_frame_name(int iframe, int id, char *out, int n)
{
    if (!out || n < 24 || iframe < 0 || id < 0) return MAGIC_H5_EARG;
    frame_name(out, (size_t)n, iframe, id);
    return MAGIC_H5_OK;
}

extern "C" int magic_out3_h5(const char *dir, int meqn, int mbc, int mx, int my, int mz, double xlower,
                             double ylower, double zlower, double dx, double dy, double dz, const double *q,
                             double t, int iframe)
{
    if (!q || meqn < 1 || mbc < 0 || mx < 1 || my < 1 || mz < 1 || iframe < 0) return MAGIC_H5_EARG;
    char name[64];
    frame_name(name, sizeof(name), iframe, 0);
    std::string path = dir && dir[0] ? std::string(dir) + "/" + name : std::string(name);

    // interior of q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc, 1-mbc:mz+mbc)
    const long long NX = mx + 2 * mbc, NY = my + 2 * mbc;
    std::vector<double> buf((size_t)meqn * mx * my * mz);
    size_t o = 0;
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) {
            const double *row = q + (((long long)(k + mbc) * NY + (j + mbc)) * NX + mbc) * meqn;
            memcpy(buf.data() + o, row, (size_t)mx * meqn * sizeof(double));
            o += (size_t)mx * meqn;
        }

    // out3_mpi_h5sep.f90:151-175 (single rank: mtotal = (mx, my, mz), lx = ly = lz = 1)
    double attr[25];
    attr[0] = 1.0;                 // ngrids_out
    attr[1] = 3.0;                 // nDim
    attr[2] = t;
    attr[3] = meqn;
    attr[4] = 1.0;
    attr[5] = mx; attr[6] = my; attr[7] = mz;
    attr[8] = 0.0;
    attr[9] = xlower; attr[10] = ylower; attr[11] = zlower;
    attr[12] = 0.0;
    attr[13] = xlower + mx * dx; attr[14] = ylower + my * dy; attr[15] = zlower + mz * dz;
    attr[16] = 0.0;
    attr[17] = dx; attr[18] = dy; attr[19] = dz;
    attr[20] = 0.0;
    attr[21] = 1.0; attr[22] = 1.0; attr[23] = 1.0;   // lx, ly, lz
    attr[24] = iframe;

    H5Closer f{H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
    if (f.id < 0) return MAGIC_H5_EIO;
    const hsize_t dims[4] = {(hsize_t)mz, (hsize_t)my, (hsize_t)mx, (hsize_t)meqn};
    const hsize_t chunk[4] = {(hsize_t)mz, (hsize_t)my, (hsize_t)mx, 1};
    H5Closer space{H5Screate_simple(4, dims, nullptr), H5Sclose};
    H5Closer dcpl{H5Pcreate(H5P_DATASET_CREATE), H5Pclose};
    if (space.id < 0 || dcpl.id < 0) return MAGIC_H5_EIO;
    if (H5Pset_chunk(dcpl.id, 4, chunk) < 0 || H5Pset_deflate(dcpl.id, 6) < 0 ||
        H5Pset_alloc_time(dcpl.id, H5D_ALLOC_TIME_EARLY) < 0)
        return MAGIC_H5_EIO;
    H5Closer ds{H5Dcreate2(f.id, "Pid", H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, dcpl.id, H5P_DEFAULT),
                H5Dclose};
    if (ds.id < 0) return MAGIC_H5_EIO;
    const hsize_t adims[1] = {25};
    H5Closer atype{H5Tcopy(H5T_NATIVE_DOUBLE), H5Tclose};
    H5Closer aspace{H5Screate_simple(1, adims, nullptr), H5Sclose};
    if (atype.id < 0 || aspace.id < 0) return MAGIC_H5_EIO;
    H5Closer a{H5Acreate2(ds.id, "Parameters", atype.id, aspace.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
    if (a.id < 0 || H5Awrite(a.id, atype.id, attr) < 0) return MAGIC_H5_EIO;
    if (H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) return MAGIC_H5_EIO;
    return MAGIC_H5_OK;
}

extern "C" int magic_in3_h5(const char *path, int *dims4, int *chunk4, int *deflate_level, double *params25,
                            double *data, long long ndata)
{
    if (!path) return MAGIC_H5_EARG;
    H5Closer f{H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
    if (f.id < 0) return MAGIC_H5_EIO;
    H5Closer ds{H5Dopen2(f.id, "Pid", H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return MAGIC_H5_EIO;
    H5Closer sp{H5Dget_space(ds.id), H5Sclose};
    if (sp.id < 0 || H5Sget_simple_extent_ndims(sp.id) != 4) return MAGIC_H5_EIO;
    hsize_t d[4];
    H5Sget_simple_extent_dims(sp.id, d, nullptr);
    if (dims4)
        for (int k = 0; k < 4; k++) dims4[k] = (int)d[k];
    H5Closer pl{H5Dget_create_plist(ds.id), H5Pclose};
    if (pl.id < 0) return MAGIC_H5_EIO;
    if (chunk4) {
        hsize_t c[4] = {0, 0, 0, 0};
        if (H5Pget_layout(pl.id) == H5D_CHUNKED) H5Pget_chunk(pl.id, 4, c);
        for (int k = 0; k < 4; k++) chunk4[k] = (int)c[k];
    }
    if (deflate_level) {
        *deflate_level = -1;
        const int nf = H5Pget_nfilters(pl.id);
        for (int k = 0; k < nf; k++) {
            unsigned flags = 0, cd[4] = {0, 0, 0, 0};
            size_t ncd = 4;
            const H5Z_filter_t id = H5Pget_filter2(pl.id, (unsigned)k, &flags, &ncd, cd, 0, nullptr, nullptr);
            if (id == H5Z_FILTER_DEFLATE && ncd >= 1) *deflate_level = (int)cd[0];
        }
    }
    if (params25) {
        H5Closer a{H5Aopen(ds.id, "Parameters", H5P_DEFAULT), H5Aclose};
        if (a.id < 0 || H5Aread(a.id, H5T_NATIVE_DOUBLE, params25) < 0) return MAGIC_H5_EIO;
    }
    if (data) {
        const long long n = (long long)d[0] * d[1] * d[2] * d[3];
        if (ndata < n) return MAGIC_H5_EARG;
        if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return MAGIC_H5_EIO;
    }
    return MAGIC_H5_OK;
}
