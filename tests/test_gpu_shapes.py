"""Tile shapes and the autotuners (magic_set_tile_shapes, magic_tune,
magic_batch_set_tile_shapes, magic_batch_tune).

* Every tile shape, in every sweep direction, reproduces the oracle's step3
  within the FP64 bar (1e-12 relative to each component's max), on every
  test grid; a mixed assignment (a different shape per direction) too.
* The FP32 form with a non-default shape assignment stays within its bar
  (1e-4 of the field peak) after 10 claw3 steps.
* magic_tune returns valid shapes, leaves q unchanged bit for bit, and a
  step after it equals a step with the same shapes set by hand, bit for bit.
* The y-sweep overlapped with the x- and z-sweeps on a second stream
  (the default) and the three sweeps in turn: each order gives the same
  state on every run, bit for bit, and the two agree within the FP64 bar
  after 6 claw3 steps (FP32: 1e-5 of each component's max).
* A batch with a non-default assignment, and a batch after magic_batch_tune,
  equals each grid stepped alone with the same shapes and order of the
  sweeps (in turn, the batch's default), bit for bit, and the tuner leaves
  the scenarios' states unchanged.
"""
import numpy as np
import pytest

from oracle.oracle import COracle
from tests.helpers import CASES, perturbed, rel_err

pytestmark = pytest.mark.gpu
NSHAPE = 4


@pytest.fixture(scope="module")
def oracle():
    return COracle()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_every_shape_step3_vs_oracle(case, oracle):
    from magic_master_amd.hip import MagicHIP
    (mx, my, mz), mthbc = CASES[case]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=case)
    qo, cfl_o, st = oracle.step3(g, q, aux, 2.0, p.nspecadv)
    assert st == 0
    for shapes in [[k, k, k] for k in range(NSHAPE)] + [[3, 1, 2], [2, 0, 1]]:
        print(f"[shapes] case {case} grid {mx}x{my}x{mz} shapes {shapes}", flush=True)
        d = MagicHIP(g, p, aux, q)
        try:
            d.set_tile_shapes(shapes)
            assert d.tile_shapes() == shapes
            cfl_g = d.step3(2.0)
            qg = d.get_q()
        finally:
            d.close()
        assert abs(cfl_g - cfl_o) <= 1e-12 * cfl_o
        assert rel_err(qg, qo) <= 1e-12, shapes


def test_invalid_shapes_rejected():
    from magic_master_amd.hip import MagicError, MagicHIP
    (mx, my, mz), mthbc = CASES[0]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=0)
    d = MagicHIP(g, p, aux, q)
    try:
        for bad in ([NSHAPE, 0, 0], [0, -1, 0]):
            with pytest.raises(MagicError):
                d.set_tile_shapes(bad)
        assert d.tile_shapes() == [0, 0, 0]
    finally:
        d.close()


def test_fp32_shapes_vs_oracle():
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    from magic_master_amd.report import form_error
    (mx, my, mz), mthbc = CASES[0]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=0, vel=0.05)
    O = COracle()
    qo = q.copy()
    dto = [1.0, 10.0, 0, 0, 0]
    t_o = 0.0
    for _ in range(10):
        st, t_o, dto, _, _, _ = O.claw3(g, p, qo, aux, t_o, t_o - 1.0, dto, c.cflv, [100, 0], c.method, mthbc)
        assert st == 0
        dto[0] = dto[4]
    for shapes in ([3, 2, 1], [1, 3, 2]):
        print(f"[fp32 shapes] {shapes}", flush=True)
        d = MagicHIP(g, p, aux, q, form=FORM_FP32P)
        try:
            d.set_tile_shapes(shapes)
            dtg = [1.0, 10.0, 0, 0, 0]
            t_g = 0.0
            for _ in range(10):
                t_g, dtg, _, _, _ = d.claw3(t_g, t_g - 1.0, dtg, c.cflv, [100, 0], c.method, mthbc)
                dtg[0] = dtg[4]
            qg = d.get_q()
        finally:
            d.close()
        err = form_error(g, qg, qo, aux)
        assert all(err[k]["max_rel"] <= 1e-4 for k in ("w", "T'", "rho'")), (shapes, err)


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_tune_keeps_state_and_picks_a_shape(form):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64, MagicHIP
    f = FORM_FP32P if form == "fp32" else FORM_FP64
    (mx, my, mz), mthbc = CASES[1]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=3, vel=0.05)
    print(f"[tune] {form}", flush=True)
    d = MagicHIP(g, p, aux, q, form=f)
    try:
        d.bc3(mthbc)
        q0 = d.get_q()
        shapes, ms = d.tune(2.0, reps=2)
        assert d.tile_shapes() == shapes
        assert all(0 <= s < NSHAPE for s in shapes)
        assert all(m > 0 for row in ms for m in row)
        for k in range(3):
            assert ms[k][shapes[k]] == min(ms[k])
        np.testing.assert_array_equal(d.get_q(), q0)
        d.backup()
        d.step3(2.0)
        q_tuned = d.get_q()
    finally:
        d.close()
    d = MagicHIP(g, p, aux, q0, form=f)
    try:
        d.set_tile_shapes(shapes)
        d.backup()
        d.step3(2.0)
        np.testing.assert_array_equal(d.get_q(), q_tuned)
    finally:
        d.close()


def _scenarios():
    from tests.test_gpu_batch import _scenario
    return [_scenario(k, forcing=False) for k in range(3)]


def _single(scen, shapes, nsteps, dt0):
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    out = []
    for c, p, g, aux, col, q in scen:
        d = MagicHIP(g, p, aux, q, form=FORM_FP32P)
        try:
            d.set_tile_shapes(shapes)
            d.set_sweep_overlap(False)   # the batch's default order of additions
            dtv = [dt0, c.dtv[1], 0, 0, 0]
            t = 0.0
            for _ in range(nsteps):
                t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [100, 0], c.method, c.mthbc)
                dtv[0] = dtv[4]
            out.append(d.get_q())
        finally:
            d.close()
    return out


def _batch(scen, nsteps, dt0, shapes=None, tune=False):
    from magic_master_amd.hip import MagicBatch
    b = MagicBatch([s[2] for s in scen], [s[1] for s in scen], [s[4] for s in scen], [s[5] for s in scen])
    try:
        for k, (c, p, g, aux, col, q) in enumerate(scen):
            b.start(k, 0.0, 1e300, dt0, [dt0, c.dtv[1], 0, 0, 0], c.cflv, nsteps, c.mthbc)
        if shapes is not None:
            b.set_tile_shapes(shapes)
        if tune:
            before = [b.get_q(k) for k in range(len(scen))]
            shapes, ms = b.tune(reps=2)
            print(f"  batch tuner: shapes {shapes} ms {ms}", flush=True)
            assert all(0 <= s < NSHAPE for s in shapes)
            for k in range(len(scen)):
                np.testing.assert_array_equal(b.get_q(k), before[k])
        assert b.tile_shapes() == list(shapes)
        b.run(max_attempts=10000, check_every=4)
        assert all(b.state(k)["nsteps"] == nsteps for k in range(len(scen)))
        return list(shapes), [b.get_q(k) for k in range(len(scen))]
    finally:
        b.close()


def test_batch_with_shapes_equals_single_grids():
    scen = _scenarios()
    print("[batch shapes] 3 scenarios, shapes (3, 1, 2)", flush=True)
    shapes, qb = _batch(scen, 4, 1.0, shapes=[3, 1, 2])
    for k, qs in enumerate(_single(scen, shapes, 4, 1.0)):
        np.testing.assert_array_equal(qb[k], qs)


def test_batch_tune_then_equals_single_grids():
    scen = _scenarios()
    print("[batch tune] 3 scenarios", flush=True)
    shapes, qb = _batch(scen, 4, 1.0, tune=True)
    for k, qs in enumerate(_single(scen, shapes, 4, 1.0)):
        np.testing.assert_array_equal(qb[k], qs)


@pytest.mark.parametrize("form", ["fp64", "fp32"])
@pytest.mark.parametrize("grid", ["case0", "odd"])
def test_sweep_overlap_order(form, grid):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64, MagicHIP
    if grid == "case0":
        (mx, my, mz), mthbc = CASES[0]
        c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=77, vel=0.3)
    else:
        # no tracer and 13 x 15 x 11 padded cells: meqn * cells = 10725 is odd,
        # so the merge's last elements are past its 16-byte vectors
        (mx, my, mz), mthbc = (9, 11, 7), [1, 1, 2, 2, 3, 1]
        c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=78, vel=0.3, meqn=5, nspecadv=0)
    print(f"[sweep overlap] {form} grid {mx}x{my}x{mz} meqn {g.meqn}", flush=True)

    def run(on):
        d = MagicHIP(g, p, aux, q, form=FORM_FP64 if form == "fp64" else FORM_FP32P)
        try:
            d.set_tile_shapes([2, 1, 0])
            d.set_sweep_overlap(on)
            dtv = [1.0, c.dtv[1], 0, 0, 0]
            t = 0.0
            for _ in range(6):
                t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [100, 0], c.method, mthbc)
                dtv[0] = dtv[4]
            return t, d.get_q()
        finally:
            d.close()

    t1, q1 = run(True)
    t2, q2 = run(True)
    t0, q0 = run(False)
    assert t1 == t2
    np.testing.assert_array_equal(q1, q2)
    assert abs(t1 - t0) <= 1e-12 * t0
    assert rel_err(q1, q0) <= (1e-12 if form == "fp64" else 1e-5)
