"""magic_global_sums_noalloc (static device scratch, no allocation) against
magic_global_sums, through the C ABI: equal bit for bit on two calls in a
row, the second on a smaller grid, in both forms; a grid beyond the
scratch's layer limit is refused before anything runs on the device."""
import ctypes as C

import numpy as np
import pytest

from magic_master_amd import problem as P

pytestmark = pytest.mark.gpu


def _noalloc(L):
    f = L.magic_global_sums_noalloc       # argument types set by load_library
    f.restype = C.c_int
    return f


def _sums(L, f, h, meqn):
    out = np.zeros(meqn)
    st = f(h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert st == 0, st
    return out


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_noalloc_sums_equal_allocating_sums(form):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64, MagicHIP, load_library
    L = load_library()
    noalloc = _noalloc(L)
    fm = FORM_FP64 if form == "fp64" else FORM_FP32P
    # first a larger grid, then a smaller one (fewer layers: the scratch keeps
    # the first call's partial sums beyond them)
    for (mx, my, mz), seed in (((40, 30, 36), 1), ((12, 9, 7), 2)):
        print(f"[noalloc {form}] {mx}x{my}x{mz}", flush=True)
        c, p, g, aux, q = P.config_plain(mx=mx, my=my, mz=mz, noise=1e-3, seed=seed)
        rng = np.random.default_rng(seed)
        q[..., 1:4] += 0.3 * rng.standard_normal(q.shape[:3] + (3,)) * q[..., :1]
        d = MagicHIP(g, p, aux, q, form=fm)
        try:
            a = _sums(L, noalloc, d.h, g.meqn)
            b = d.global_sums()
            a2 = _sums(L, noalloc, d.h, g.meqn)
        finally:
            d.close()
        assert np.array_equal(a, b) and np.array_equal(a2, b), (a, b, a2)
        assert b[0] > 0 and b[4] > 0


def test_noalloc_refuses_more_layers_than_its_scratch():
    from magic_master_amd.hip import MAGIC_OK, MagicHIP, load_library
    L = load_library()
    noalloc = _noalloc(L)
    # 1025 layers, one more than the scratch holds (mx, my such that the
    # context accepts mz > mx)
    c, p, g, aux, q = P.config_plain(mx=128, my=10, mz=1025, noise=0.0)
    d = MagicHIP(g, p, aux, q)
    try:
        out = np.full(g.meqn, -1.0)
        st = noalloc(d.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        assert st != MAGIC_OK
        assert np.all(out == -1.0)
        assert d.global_sums()[0] > 0      # the allocating form serves it
    finally:
        d.close()
