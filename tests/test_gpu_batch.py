"""The batched driver (magic_batch_*, FP32 form) against single grids.

* Each scenario of a batch, stepped by the device-resident step control and
  the captured graph, equals the same grid stepped alone by magic_claw3 in
  the FP32 form with the same order of the sweeps' additions (the default
  overlap, and the sweeps in turn): the same dt of every attempt (a rejected first attempt
  included) and the same state bit for bit, with the plume forcing off (with
  it on, the forcing's time factor is evaluated by the host's exp for a
  single grid and by the device's exp in a batch: rho', T', w within 1e-5
  of their peaks).
* claw3's interval semantics: a scenario started at t with tend > t stops
  exactly at tend (dt clipped, claw3_mpi.f:741).
* A scenario's result is the same alone and in a mix of scenarios, bit for
  bit.
"""
import numpy as np
import pytest

from magic_master_amd.report import form_error
from tests.helpers import perturbed

pytestmark = pytest.mark.gpu

SHAPES = [((12, 10, 8), [2, 2, 2, 2, 3, 1]), ((9, 11, 7), [1, 1, 1, 1, 3, 1]),
          ((16, 12, 10), [1, 1, 2, 2, 3, 1])]


def _scenario(k, forcing):
    (mx, my, mz), mthbc = SHAPES[k]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=10 + k, vel=0.05)
    if not forcing:
        p.tcenter = 1e9          # precipitation below the 0.1 threshold everywhere
    col = aux[:, 0, 0, :].copy()
    return c, p, g, aux, col, q


def _single(c, p, g, aux, q, mthbc, nsteps, dt0, overlap=False):
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    d = MagicHIP(g, p, aux, q, form=FORM_FP32P)
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


def _batch(scen, nsteps, dt0, tend=None, overlap=False):
    from magic_master_amd.hip import MagicBatch
    b = MagicBatch([s[2] for s in scen], [s[1] for s in scen], [s[4] for s in scen], [s[5] for s in scen])
    try:
        b.set_sweep_overlap(overlap)
        for k, (c, p, g, aux, col, q) in enumerate(scen):
            mthbc = c.mthbc
            # nsteps accepted steps (an end time that is never reached), or up to tend
            b.start(k, 0.0, 1e300 if tend is None else tend, dt0, [dt0, c.dtv[1], 0, 0, 0], c.cflv,
                    nsteps if tend is None else 100000, mthbc)
        n = b.run(max_attempts=10000, check_every=4)
        print(f"  batch of {len(scen)}: {n} attempts", flush=True)
        return [b.state(k) for k in range(len(scen))], [b.get_q(k) for k in range(len(scen))]
    finally:
        b.close()


@pytest.mark.parametrize("overlap", [True, False])
def test_batch_equals_single_grids_bitwise(overlap):
    scen = [_scenario(k, forcing=False) for k in range(3)]
    nsteps = 6
    print(f"[batch] 3 scenarios, 6 steps, first dt 50 s (rejected), overlap {overlap}", flush=True)
    states, qb = _batch(scen, nsteps, 50.0, overlap=overlap)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        t, dt_next, qs, log = _single(c, p, g, aux, q, c.mthbc, nsteps, 50.0, overlap=overlap)
        st = states[k]
        assert st["status"] == 0 and st["active"] == 0
        assert st["nsteps"] == nsteps
        assert st["attempts"] == len(log) and st["rejects"] == sum(not a for _, a in log)
        assert st["rejects"] >= 1
        assert st["t"] == t and st["dt"] == dt_next
        np.testing.assert_array_equal(qb[k], qs)


def test_batch_with_forcing_close_to_single_grids():
    scen = [_scenario(k, forcing=True) for k in range(3)]
    print("[batch forcing] 3 scenarios, 8 steps", flush=True)
    states, qb = _batch(scen, 8, 1.0)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        t, dt_next, qs, log = _single(c, p, g, aux, q, c.mthbc, 8, 1.0)
        assert abs(states[k]["t"] - t) <= 1e-12 * t
        err = form_error(g, qb[k], qs, aux)
        assert all(e["max_rel"] <= 1e-5 for e in err.values()), err


def test_batch_interval_stops_at_tend():
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    scen = [_scenario(k, forcing=False) for k in range(2)]
    states, qb = _batch(scen, 0, 1.0, tend=30.0)
    for k, (c, p, g, aux, col, q) in enumerate(scen):
        assert states[k]["t"] == 30.0
        d = MagicHIP(g, p, aux, q, form=FORM_FP32P)
        d.set_sweep_overlap(False)   # the batch's default order of additions
        try:
            te, dtv, _, nv, _ = d.claw3(0.0, 30.0, [1.0, c.dtv[1], 0, 0, 0], c.cflv, [100000, 0], c.method,
                                         c.mthbc)
            qs = d.get_q()
        finally:
            d.close()
        assert te == 30.0 and nv[1] == states[k]["nsteps"]
        np.testing.assert_array_equal(qb[k], qs)


def test_scenario_alone_equals_in_the_mix():
    scen = [_scenario(k, forcing=True) for k in range(3)]
    _, alone = _batch([scen[1]], 5, 1.0)
    _, mix = _batch(scen, 5, 1.0)
    np.testing.assert_array_equal(alone[0], mix[1])
