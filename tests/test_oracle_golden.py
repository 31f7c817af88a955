"""The C oracle against golden fixtures written by the compiled reference
(tests/golden/make_golden.py): bit-for-bit equality."""
import os

import numpy as np
import pytest

from magic_master_amd import problem as P
from oracle.oracle import COracle
from tests.helpers import CASES, perturbed, unshuffle_planes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
AUX_SLOTS = [1, 2, 3, 4, 5, 11, 12, 17, 18, 19]


def _load(name):
    return np.load(os.path.join(GOLD, name))


def unpack_aux(packed, maux=32):
    aux = np.zeros(packed.shape[:3] + (maux,))
    for k, s in enumerate(AUX_SLOTS):
        aux[..., s - 1] = packed[..., k]
    return aux


@pytest.fixture(scope="module")
def O():
    return COracle()


@pytest.mark.parametrize("key", [f"m{m}_x{x}" for m in (5, 6, 7) for x in (1, 2, 3)])
def test_rpn3_known_answers(key, O):
    z = _load("rpn3_kat.npz")
    q = z[key + "_q"]
    aux = z[key + "_aux"]
    st, fw, s, am, ap, roe = O.rpn3(int(key[-1]), q, aux[:, 18], aux[:, 16], 3.0, 2000.0,
                                    int(z[key + "_nspec"]))
    assert st == 0
    N = q.shape[0]
    # interfaces 0..n+2 are defined (Fortran i = 2-mbc..mx+mbc)
    sl = slice(1, N)
    assert np.array_equal(s[sl], z[key + "_s"][sl])
    assert np.array_equal(am[sl], z[key + "_amdq"][sl])
    assert np.array_equal(ap[sl], z[key + "_apdq"][sl])
    assert np.array_equal(roe[:, sl], z[key + "_roe"][:, sl])
    # limited f-waves on the interfaces the limiter visits (1..n+1)
    fl = slice(2, N - 1)
    assert np.array_equal(fw[fl], z[key + "_fwave"][fl])


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_step3_bc3_src3(ci, O):
    z = _load("ops_small.npz")
    (mx, my, mz), mthbc = CASES[ci]
    c, p, g, _, _ = perturbed(mx, my, mz, mthbc, seed=ci)
    q = z[f"c{ci}_q"]
    aux = unpack_aux(z[f"c{ci}_aux"])
    qn, cfl, st = O.step3(g, q, aux, 2.0, p.nspecadv)
    assert st == 0
    assert cfl == float(z[f"c{ci}_step3_cfl"])
    assert np.array_equal(qn, z[f"c{ci}_step3_q"])
    qb = q.copy()
    assert O.bc3(g, qb, aux, mthbc) == 0
    O.b4step3(g, qb)
    assert np.array_equal(qb, z[f"c{ci}_bc3_q"])
    c2, p2, g2, _, _ = perturbed(mx, my, mz, mthbc, seed=ci, visc=3e10)
    p2.tcenter = 0.0
    qs = z[f"c{ci}_src3_qin"].copy()
    aux2 = unpack_aux(z[f"c{ci}_src3_aux"])
    st, ndtr, ndtrc = O.src3(g2, p2, qs, aux2, 0.0, 5.0, mthbc)
    assert st == 0 and ndtr > 1
    assert np.array_equal(qs, z[f"c{ci}_src3_q"])


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_claw3_ten_steps(ci, O):
    z = _load("claw3_small.npz")
    (mx, my, mz), mthbc = CASES[ci]
    c, p, g, _, _ = perturbed(mx, my, mz, mthbc, seed=ci, vel=0.05)
    q = z[f"c{ci}_q"].copy()
    aux = unpack_aux(z[f"c{ci}_aux"])
    dtv = [50.0, c.dtv[1], 0, 0, 0]
    t = 0.0
    dts = []
    for n in range(10):
        st, t, dtv, cflv, nv, log = O.claw3(g, p, q, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                            c.method, mthbc)
        assert st == 0
        dts.append(dtv[4])
        dtv[0] = dtv[4]
    assert t == float(z[f"c{ci}_t"])
    assert np.array_equal(np.array(dts), z[f"c{ci}_next_dt"])
    assert np.array_equal(q, z[f"c{ci}_qfinal"])


def test_plain_workload_first_steps(O):
    """64x64x48 isothermal workload: the first 10 steps of the 200-step
    reference run (dt sequence and every interior cell, bit-for-bit)."""
    z = _load("plain_200.npz")
    c, p, g, aux, q = P.config_plain()
    m = g.mbc
    q0 = q[m:m + g.mz, m:m + g.my, m:m + g.mx].copy()
    dtv = list(c.dtv)
    t = 0.0
    for n in range(1, 11):
        st, t, dtv, cflv, nv, log = O.claw3(g, p, q, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                            c.method, c.mthbc)
        assert st == 0
        assert dtv[4] == z["next_dt"][n - 1]
        assert cflv[3] == z["cfl"][n - 1]
        dtv[0] = dtv[4]
        if n in (1, 10):
            assert np.array_equal(q[m:m + g.mz, m:m + g.my, m:m + g.mx] - q0,
                                  unshuffle_planes(z[f"step{n}_dq"], q0.shape))
