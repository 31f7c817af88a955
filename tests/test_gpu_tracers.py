"""Tracer counts other than one: meqn = 5 (no tracer), 7, 8, 9 and 10 (2, 3,
4 and 5 tracers, some advected with the tracer limiter and the rest in
continuity form), i.e. every sweep instantiation the library carries.

* FP64 form against the C oracle (itself pinned bit for bit to the compiled
  reference for these counts, tests/test_oracle_reference.py): one step3 call
  ≤ 1e-12, with the default and a pencil-major tile shape; 10 claw3 steps
  (first attempt rejected) with every dt to 1e-12 and the state ≤ 1e-9.
* FP32 form against the oracle after 10 claw3 steps: w, T', rho' within 1e-4
  of their peaks, tracers within 1e-4 (their range is [0, 1]).
* Batched driver with 0 and 2 tracers, FP32 and FP64 forms: each scenario
  equals the same grid stepped alone in the same form (sweeps in turn, the
  batch's default), bit for bit.
"""
import numpy as np
import pytest

from oracle.oracle import COracle
from tests.helpers import perturbed, rel_err

pytestmark = pytest.mark.gpu
COUNTS = [(5, 0), (7, 1), (8, 2), (9, 1), (10, 3)]
GRID = ((12, 10, 8), [2, 2, 2, 2, 3, 1])


@pytest.fixture(scope="module")
def oracle():
    return COracle()


def _case(meqn, nspec, seed, vel=0.5):
    (mx, my, mz), mthbc = GRID
    return perturbed(mx, my, mz, mthbc, seed=seed, vel=vel, meqn=meqn, nspecadv=nspec)


@pytest.mark.parametrize("meqn,nspec", COUNTS)
def test_step3_tracer_counts(meqn, nspec, oracle):
    from magic_master_amd.hip import MagicHIP
    c, p, g, aux, q = _case(meqn, nspec, seed=meqn)
    qo, cfl_o, st = oracle.step3(g, q, aux, 2.0, nspec)
    assert st == 0
    for shapes in ([0, 0, 0], [2, 1, 3]):
        print(f"[tracers step3] meqn {meqn} nspecadv {nspec} shapes {shapes}", flush=True)
        d = MagicHIP(g, p, aux, q)
        try:
            d.set_tile_shapes(shapes)
            cfl_g = d.step3(2.0)
            qg = d.get_q()
        finally:
            d.close()
        assert abs(cfl_g - cfl_o) <= 1e-12 * cfl_o
        assert rel_err(qg, qo) <= 1e-12, shapes


@pytest.mark.parametrize("meqn,nspec", COUNTS)
def test_claw3_tracer_counts(meqn, nspec, oracle):
    from magic_master_amd.hip import MagicHIP
    c, p, g, aux, q = _case(meqn, nspec, seed=meqn + 1, vel=0.05)
    mthbc = GRID[1]
    print(f"[tracers claw3] meqn {meqn} nspecadv {nspec}", flush=True)
    qo = q.copy()
    d = MagicHIP(g, p, aux, q)
    dto = [50.0, 10.0, 0, 0, 0]
    dtg = list(dto)
    t_o = t_g = 0.0
    try:
        for n in range(10):
            st, t_o, dto, _, _, _ = oracle.claw3(g, p, qo, aux, t_o, t_o - 1.0, dto, c.cflv, [100, 0],
                                                 c.method, mthbc)
            assert st == 0
            t_g, dtg, _, _, _ = d.claw3(t_g, t_g - 1.0, dtg, c.cflv, [100, 0], c.method, mthbc)
            assert abs(dtg[4] - dto[4]) <= 1e-12 * dto[4], (n, dtg[4], dto[4])
            dto[0] = dto[4]
            dtg[0] = dtg[4]
        qg = d.get_q()
    finally:
        d.close()
    assert rel_err(qg, qo) <= 1e-9


@pytest.mark.parametrize("meqn,nspec", COUNTS)
def test_fp32_tracer_counts(meqn, nspec, oracle):
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    from magic_master_amd.report import form_error
    c, p, g, aux, q = _case(meqn, nspec, seed=meqn + 2, vel=0.05)
    mthbc = GRID[1]
    print(f"[tracers fp32] meqn {meqn} nspecadv {nspec}", flush=True)
    qo = q.copy()
    d = MagicHIP(g, p, aux, q, form=FORM_FP32P)
    dto = [1.0, 10.0, 0, 0, 0]
    dtg = list(dto)
    t_o = t_g = 0.0
    try:
        for _ in range(10):
            st, t_o, dto, _, _, _ = oracle.claw3(g, p, qo, aux, t_o, t_o - 1.0, dto, c.cflv, [100, 0],
                                                 c.method, mthbc)
            assert st == 0
            t_g, dtg, _, _, _ = d.claw3(t_g, t_g - 1.0, dtg, c.cflv, [100, 0], c.method, mthbc)
            dto[0] = dto[4]
            dtg[0] = dtg[4]
        qg = d.get_q()
    finally:
        d.close()
    err = form_error(g, qg, qo, aux)
    assert all(err[k]["max_rel"] <= 1e-4 for k in ("w", "T'", "rho'")), err
    if meqn > 5:
        m = g.mbc
        it = (slice(m, m + g.mz), slice(m, m + g.my), slice(m, m + g.mx))
        assert np.abs(qg[it][..., 5:] - qo[it][..., 5:]).max() <= 1e-4


@pytest.mark.parametrize("form", ["fp32", "fp64"])
@pytest.mark.parametrize("meqn,nspec", [(5, 0), (7, 2)])
def test_batch_tracer_counts(meqn, nspec, form):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64, MagicBatch, MagicHIP
    fm = FORM_FP32P if form == "fp32" else FORM_FP64
    scen = []
    for k, (dims, mthbc) in enumerate([((12, 10, 8), [2, 2, 2, 2, 3, 1]), ((9, 11, 7), [1, 1, 1, 1, 3, 1])]):
        c, p, g, aux, q = perturbed(*dims, mthbc, seed=20 + k, vel=0.05, meqn=meqn, nspecadv=nspec)
        p.tcenter = 1e9          # forcing off: the batch and the single grid then agree bit for bit
        scen.append((c, p, g, aux, aux[:, 0, 0, :].copy(), q))
    print(f"[tracers batch] meqn {meqn} {form}", flush=True)
    b = MagicBatch([s[2] for s in scen], [s[1] for s in scen], [s[4] for s in scen], [s[5] for s in scen],
                   form=fm)
    try:
        for k, (c, p, g, aux, col, q) in enumerate(scen):
            b.start(k, 0.0, 1e300, 1.0, [1.0, c.dtv[1], 0, 0, 0], c.cflv, 4, c.mthbc)
        b.run(max_attempts=1000, check_every=4)
        qb = [b.get_q(k) for k in range(len(scen))]
        assert all(b.state(k)["nsteps"] == 4 for k in range(len(scen)))
    finally:
        b.close()
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        d = MagicHIP(g, p, aux, q, form=fm)
        d.set_sweep_overlap(False)   # the batch's default order of additions
        try:
            dtv = [1.0, c.dtv[1], 0, 0, 0]
            t = 0.0
            for _ in range(4):
                t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [100, 0], c.method, c.mthbc)
                dtv[0] = dtv[4]
            np.testing.assert_array_equal(qb[k], d.get_q())
        finally:
            d.close()
