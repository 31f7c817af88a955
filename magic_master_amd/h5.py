"""ctypes binding of the HDF5 frame writer/reader (include/magic_h5.h):
frames with the layout of the reference's out3 (mpilib/out3_mpi_h5sep.f90)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .problem import Grid

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmagic_h5.so")
_L = None


def load_library():
    global _L
    if _L is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built (python -m magic_master_amd.build)")
        L = C.CDLL(_LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.magic_h5_frame_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.magic_out3_h5.argtypes = ([C.c_char_p] + [C.c_int] * 5 + [C.c_double] * 6 +
                                    [dp, C.c_double, C.c_int])
        L.magic_in3_h5.argtypes = [C.c_char_p, ip, ip, ip, dp, dp, C.c_longlong]
        _L = L
    return _L


def frame_name(iframe: int, rank: int = 0) -> str:
    buf = C.create_string_buffer(64)
    if load_library().magic_h5_frame_name(iframe, rank, buf, 64) != 0:
        raise ValueError("bad frame number")
    return buf.value.decode()


def out3(g: Grid, q: np.ndarray, t: float, iframe: int, directory: str = "") -> str:
    """Write q (reference layout, ghosts included) as frame iframe; returns the path."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    if q.shape != g.shape_q:
        raise ValueError(f"q shape {q.shape} != {g.shape_q}")
    st = load_library().magic_out3_h5(directory.encode(), g.meqn, g.mbc, g.mx, g.my, g.mz, g.xlower,
                                      g.ylower, g.zlower, g.dx, g.dy, g.dz,
                                      q.ctypes.data_as(C.POINTER(C.c_double)), t, iframe)
    if st != 0:
        raise RuntimeError(f"magic_out3_h5 failed with status {st}")
    return os.path.join(directory, frame_name(iframe)) if directory else frame_name(iframe)


def in3(path: str) -> dict:
    """Read a frame: data (mz, my, mx, meqn), Parameters[25], chunk, gzip level."""
    L = load_library()
    dims = (C.c_int * 4)()
    chunk = (C.c_int * 4)()
    lvl = C.c_int()
    par = (C.c_double * 25)()
    st = L.magic_in3_h5(path.encode(), dims, chunk, C.byref(lvl), par, None, 0)
    if st != 0:
        raise RuntimeError(f"magic_in3_h5 failed with status {st}")
    data = np.empty(tuple(dims), dtype=np.float64)
    st = L.magic_in3_h5(path.encode(), None, None, None, None,
                        data.ctypes.data_as(C.POINTER(C.c_double)), data.size)
    if st != 0:
        raise RuntimeError(f"magic_in3_h5 failed with status {st}")
    return {"data": data, "params": np.array(par[:]), "chunk": tuple(chunk), "deflate": lvl.value,
            "dims": tuple(dims)}
