"""The batched driver (magic_batch_*) against single grids and the oracle.

* Each scenario of a batch, stepped by the device-resident step control and
  the captured graph, equals the same grid stepped alone by magic_claw3 in
  the same form (FP32 perturbation form and FP64 form) with the same order
  of the sweeps' additions (the default overlap, and the sweeps in turn):
  the same dt of every attempt (a rejected first attempt included) and the
  same state bit for bit, with the plume forcing off.  With it on, the
  forcing's time factor is evaluated by the host's exp for a single grid
  and by the device's exp in a batch: rho', T', w within 1e-5 of their
  peaks in the FP32 form, 1e-10 in the FP64 form.
* FP64 batch against the C oracle's claw3 (oracle/magic_oracle.c, pinned to
  the compiled reference): the dt of every attempt within 1e-12 relative,
  the fields within 1e-9 of each component's max after 6 steps, as the
  single-grid parity tests.
* claw3's interval semantics: a scenario started at t with tend > t stops
  exactly at tend (dt clipped, claw3_mpi.f:741).
* A scenario's result is the same alone and in a mix of scenarios, bit for
  bit.
"""
import numpy as np
import pytest

from magic_master_amd.report import form_error
from oracle.oracle import COracle
from tests.helpers import interior, perturbed, rel_err

pytestmark = pytest.mark.gpu

SHAPES = [((12, 10, 8), [2, 2, 2, 2, 3, 1]), ((9, 11, 7), [1, 1, 1, 1, 3, 1]),
          ((16, 12, 10), [1, 1, 2, 2, 3, 1])]


def _form(name):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64
    return {"fp32": FORM_FP32P, "fp64": FORM_FP64}[name]


def _scenario(k, forcing):
    (mx, my, mz), mthbc = SHAPES[k]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=10 + k, vel=0.05)
    if not forcing:
        p.tcenter = 1e9          # precipitation below the 0.1 threshold everywhere
    col = aux[:, 0, 0, :].copy()
    return c, p, g, aux, col, q


def _single(c, p, g, aux, q, mthbc, nsteps, dt0, overlap=False, form="fp32"):
    from magic_master_amd.hip import MagicHIP
    d = MagicHIP(g, p, aux, q, form=_form(form))
    d.set_sweep_overlap(overlap)
    dtv = [dt0, c.dtv[1], 0, 0, 0]
    t = 0.0
    log = []
    try:
        for _ in range(nsteps):
            t, dtv, _, _, lg = d.claw3(t, t - 1.0, dtv, c.cflv, [100, 0], c.method, mthbc)
            log += list(zip(lg.dt, lg.accepted))
            dtv[0] = dtv[4]
        return t, dtv[0], d.get_q(), log
    finally:
        d.close()


def _batch(scen, nsteps, dt0, tend=None, overlap=False, form="fp32"):
    from magic_master_amd.hip import MagicBatch
    b = MagicBatch([s[2] for s in scen], [s[1] for s in scen], [s[4] for s in scen], [s[5] for s in scen],
                   form=_form(form))
    try:
        b.set_sweep_overlap(overlap)
        for k, (c, p, g, aux, col, q) in enumerate(scen):
            mthbc = c.mthbc
            # nsteps accepted steps (an end time that is never reached), or up to tend
            b.start(k, 0.0, 1e300 if tend is None else tend, dt0, [dt0, c.dtv[1], 0, 0, 0], c.cflv,
                    nsteps if tend is None else 100000, mthbc)
        n = b.run(max_attempts=10000, check_every=4)
        print(f"  batch of {len(scen)} ({form}): {n} attempts", flush=True)
        return [b.state(k) for k in range(len(scen))], [b.get_q(k) for k in range(len(scen))]
    finally:
        b.close()


@pytest.mark.parametrize("form", ["fp32", "fp64"])
@pytest.mark.parametrize("overlap", [True, False])
def test_batch_equals_single_grids_bitwise(overlap, form):
    scen = [_scenario(k, forcing=False) for k in range(3)]
    nsteps = 6
    print(f"[batch] {form}, 3 scenarios, 6 steps, first dt 50 s (rejected), overlap {overlap}", flush=True)
    states, qb = _batch(scen, nsteps, 50.0, overlap=overlap, form=form)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        t, dt_next, qs, log = _single(c, p, g, aux, q, c.mthbc, nsteps, 50.0, overlap=overlap, form=form)
        st = states[k]
        assert st["status"] == 0 and st["active"] == 0
        assert st["nsteps"] == nsteps
        assert st["attempts"] == len(log) and st["rejects"] == sum(not a for _, a in log)
        assert st["rejects"] >= 1
        assert st["t"] == t and st["dt"] == dt_next
        np.testing.assert_array_equal(qb[k], qs)


@pytest.mark.parametrize("form,tol", [("fp32", 1e-5), ("fp64", 1e-10)])
def test_batch_with_forcing_close_to_single_grids(form, tol):
    scen = [_scenario(k, forcing=True) for k in range(3)]
    print(f"[batch forcing] {form}, 3 scenarios, 8 steps", flush=True)
    states, qb = _batch(scen, 8, 1.0, form=form)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        t, dt_next, qs, log = _single(c, p, g, aux, q, c.mthbc, 8, 1.0, form=form)
        assert abs(states[k]["t"] - t) <= 1e-12 * t
        err = form_error(g, qb[k], qs, aux)
        assert all(e["max_rel"] <= tol for e in err.values()), err


def test_batch_fp64_vs_oracle():
    oracle = COracle()
    scen = [_scenario(k, forcing=True) for k in range(3)]
    nsteps = 6
    print("[batch oracle] fp64, 3 scenarios, 6 steps, first dt 50 s (rejected)", flush=True)
    states, qb = _batch(scen, nsteps, 50.0, form="fp64")
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        qo = q.copy()
        dto = [50.0, c.dtv[1], 0, 0, 0]
        t_o = 0.0
        attempts = 0
        for _ in range(nsteps):
            st, t_o, dto, _, _, log_o = oracle.claw3(g, p, qo, aux, t_o, t_o - 1.0, dto, c.cflv, [100, 0],
                                                     c.method, c.mthbc)
            assert st == 0
            attempts += len(log_o)
            dto[0] = dto[4]
        s = states[k]
        assert s["status"] == 0 and s["nsteps"] == nsteps and s["attempts"] == attempts
        assert abs(s["t"] - t_o) <= 1e-12 * t_o
        assert abs(s["dt"] - dto[0]) <= 1e-12 * dto[0]
        assert rel_err(qb[k], qo, interior(g)) <= 1e-9


@pytest.mark.parametrize("form", ["fp32", "fp64"])
def test_batch_interval_stops_at_tend(form):
    from magic_master_amd.hip import MagicHIP
    scen = [_scenario(k, forcing=False) for k in range(2)]
    states, qb = _batch(scen, 0, 1.0, tend=30.0, form=form)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        assert states[k]["t"] == 30.0
        d = MagicHIP(g, p, aux, q, form=_form(form))
        d.set_sweep_overlap(False)   # the batch's default order of additions
        try:
            te, dtv, _, nv, _ = d.claw3(0.0, 30.0, [1.0, c.dtv[1], 0, 0, 0], c.cflv, [100000, 0], c.method,
                                         c.mthbc)
            qs = d.get_q()
        finally:
            d.close()
        assert te == 30.0 and nv[1] == states[k]["nsteps"]
        np.testing.assert_array_equal(qb[k], qs)


@pytest.mark.parametrize("form", ["fp32", "fp64"])
def test_scenario_alone_equals_in_the_mix(form):
    scen = [_scenario(k, forcing=True) for k in range(3)]
    _, alone = _batch([scen[1]], 5, 1.0, form=form)
    _, mix = _batch(scen, 5, 1.0, form=form)
    np.testing.assert_array_equal(alone[0], mix[1])
