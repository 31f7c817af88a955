"""HDF5 frames with the reference's out3 layout (mpilib/out3_mpi_h5sep.f90).

The reference writer needs the HDF5 and MPI Fortran modules of the image's
conda HDF5, which were compiled by gfortran (absent here), so it cannot be
built as a checker.  The layout is pinned by reading our files back through
the HDF5 C API against what out3_mpi_h5sep.f90 states (dataset, Fortran
dims, chunk, filter, attribute order) and what the reference's reader
processing.m expects (file name, attribute indices)."""
import os

import numpy as np
import pytest

from magic_master_amd import h5
from magic_master_amd import problem as P


def test_frame_name_matches_reference_pattern():
    # out3_mpi_h5sep.f90:69-79; processing.m: strcat('fort.qq', num2str(Frame,'%04i'), 'id0000', '.h5')
    assert h5.frame_name(1) == "fort.qq0001id0000.h5"
    assert h5.frame_name(123, 7) == "fort.qq0123id0007.h5"


@pytest.mark.parametrize("meqn", [6, 10])
def test_out3_layout_and_data(tmp_path, meqn):
    g = P.Grid(mx=7, my=5, mz=4, meqn=meqn, mbc=2, maux=32, xlower=-3.0, ylower=2.0, zlower=0.0,
               dx=2000.0, dy=1500.0, dz=1000.0)
    rng = np.random.default_rng(meqn)
    q = rng.standard_normal(g.shape_q)
    path = h5.out3(g, q, 123.25, 17, str(tmp_path))
    assert os.path.basename(path) == "fort.qq0017id0000.h5"
    f = h5.in3(path)
    # Fortran dims (meqn, mx, my, mz) are C dims (mz, my, mx, meqn)
    assert f["dims"] == (g.mz, g.my, g.mx, meqn)
    # Fortran chunk (1, mx, my, mz) (out3_mpi_h5sep.f90:86-89)
    assert f["chunk"] == (g.mz, g.my, g.mx, 1)
    assert f["deflate"] == 6                                   # out3_mpi_h5sep.f90:117
    m = g.mbc
    np.testing.assert_array_equal(f["data"], q[m:m + g.mz, m:m + g.my, m:m + g.mx, :])
    a = f["params"]
    # out3_mpi_h5sep.f90:151-175, one rank (mtotal = (mx, my, mz), lx = ly = lz = 1)
    expect = [1, 3, 123.25, meqn, 1, g.mx, g.my, g.mz, 0, g.xlower, g.ylower, g.zlower, 0,
              g.xlower + g.mx * g.dx, g.ylower + g.my * g.dy, g.zlower + g.mz * g.dz, 0,
              g.dx, g.dy, g.dz, 0, 1, 1, 1, 17]
    np.testing.assert_array_equal(a, np.array(expect, dtype=np.float64))
    # processing.m reads mx = attr(6)/attr(22), dz = attr(20), t = attr(3), meqn = attr(4)
    assert a[5] / a[21] == g.mx and a[19] == g.dz and a[2] == 123.25 and a[3] == meqn


def test_out3_rejects_bad_arguments(tmp_path):
    g = P.Grid(mx=3, my=3, mz=3, meqn=6, mbc=2, maux=32, xlower=0.0, ylower=0.0, zlower=0.0,
               dx=1.0, dy=1.0, dz=1.0)
    with pytest.raises(ValueError):
        h5.out3(g, np.zeros((2, 2, 2, 6)), 0.0, 1, str(tmp_path))
    with pytest.raises(RuntimeError):
        h5.out3(g, np.zeros(g.shape_q), 0.0, 1, str(tmp_path / "missing_dir"))
