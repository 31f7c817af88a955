"""GPU parity on the two named workloads.

1. Plain isothermal workload (64x64x48, problem.config_plain) for 200 claw3
   calls against the compiled reference's own run (tests/golden/plain_200.npz):
   every step's next dt and CFL, and every interior cell of the state after
   steps 1, 10 and 200.  Bars: 1e-12 after one step, 1e-9 after 10 and 200
   (relative to each component's largest value; the momenta share one
   scale).
2. Full-domain workload (250x200x300, problem.config_full) at its real size,
   through a size-independent property: with one horizontal axis periodic and
   the state uniform along it, every row along that axis must equal the
   oracle's run of the same problem two cells wide.  This covers the sweeps,
   bc3 and src3 at full extents, where the oracle cannot run in seconds.
3. The same workload with the plume on from t = 0 (a fully 3-D flow), through
   the domain of dependence: a step moves information a bounded number of
   cells, so after 10 steps the cells of a 40x40x300 window around the plume
   that lie 12 or more cells inside its edges equal the oracle's run of the
   window alone (tests/helpers.plume_window; the radius is checked on the
   oracle itself by tests/test_oracle_kat.py::test_plume_window_radius).
   Outside the plume's reach (40 cells and more from it) every column of the
   full run stays equal to the corner column to round-off.
"""
import dataclasses
import functools
import os

import numpy as np
import pytest

from magic_master_amd import problem as P
from oracle.oracle import COracle
from magic_master_amd.report import form_error
from tests.helpers import WINDOW_MARGIN, plume_window, rel_err_mom, uniform_state, unshuffle_planes

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _dev(g, p, aux, q):
    from magic_master_amd.hip import MagicHIP
    return MagicHIP(g, p, aux, q)


def _dev_form(g, p, aux, q, form):
    from magic_master_amd.hip import MagicHIP
    return MagicHIP(g, p, aux, q, form=form)


def test_plain_workload_200_steps_vs_reference_run():
    z = np.load(os.path.join(GOLD, "plain_200.npz"))
    c, p, g, aux, q = P.config_plain()
    m = g.mbc
    q0 = q[m:m + g.mz, m:m + g.my, m:m + g.mx].copy()
    print(f"[plain200] grid {g.mx}x{g.my}x{g.mz}", flush=True)
    d = _dev(g, p, aux, q)
    dtv = list(c.dtv)
    cflv = list(c.cflv)
    t = 0.0
    dts, cfls = [], []
    try:
        for n in range(1, 201):
            t, dtv, cflv, _, _ = d.claw3(t, t - 1.0, dtv, cflv, [c.nv[0], 0], c.method, c.mthbc)
            dts.append(dtv[4])
            cfls.append(cflv[3])
            dtv[0] = dtv[4]
            if n in (1, 10, 200):
                qg = d.get_q()
                it = qg[m:m + g.mz, m:m + g.my, m:m + g.mx]
                ref = q0 + unshuffle_planes(z[f"step{n}_dq"], q0.shape)   # every interior cell of the reference's run
                tol = 1e-12 if n == 1 else 1e-9
                err = rel_err_mom(it, ref)
                print(f"  step {n}: field err {err:.3e} over {it[..., 0].size} cells", flush=True)
                assert err <= tol, (n, err)
    finally:
        d.close()
    dts = np.array(dts)
    cfls = np.array(cfls)
    assert abs(dts[0] - z["next_dt"][0]) <= 1e-12 * z["next_dt"][0]
    assert np.abs(dts / z["next_dt"] - 1.0).max() <= 1e-9
    assert np.abs(cfls / z["cfl"] - 1.0).max() <= 1e-9
    assert abs(t - float(z["t"])) <= 1e-9 * float(z["t"])


def _oracle_run(c, p, g, aux, q, nsteps):
    O = COracle()
    dtv = list(c.dtv)
    t = 0.0
    out = []
    for _ in range(nsteps):
        st, t, dtv, _, _, _ = O.claw3(g, p, q, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                      c.method, c.mthbc)
        assert st == 0
        out.append((t, dtv[4], q.copy()))
        dtv[0] = dtv[4]
    return out


@pytest.mark.timeout(400)        # host set-up of the 4 GB reference-layout aux array
@pytest.mark.parametrize("axis", ["y", "x"])
def test_full_domain_rows_match_narrow_oracle_run(axis):
    mx, my, mz = 250, 200, 300
    nsteps = 2
    narrow = (mx, 2, mz) if axis == "y" else (2, my, mz)
    print(f"[full-uniform-{axis}] oracle on {narrow}", flush=True)
    ref = _oracle_run(*uniform_state(*narrow, axis), nsteps)
    c, p, g, aux, q = uniform_state(mx, my, mz, axis)
    print(f"[full-uniform-{axis}] device on {mx}x{my}x{mz}", flush=True)
    d = _dev(g, p, aux, q)
    del aux, q
    m = g.mbc
    dtv = list(c.dtv)
    t = 0.0
    try:
        for n in range(nsteps):
            t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [c.nv[0], 0], c.method, c.mthbc)
            t_o, dt_o, q_o = ref[n]
            assert abs(t - t_o) <= 1e-12 * t_o
            assert abs(dtv[4] - dt_o) <= 1e-12 * dt_o
            dtv[0] = dtv[4]
            qg = d.get_q()
            tol = 1e-12 if n == 0 else 1e-11
            if axis == "y":
                row = q_o[m:m + mz, m, m:m + mx]
                err = max(rel_err_mom(qg[m:m + mz, j, m:m + mx], row) for j in range(m, m + my))
            else:
                row = q_o[m:m + mz, m:m + my, m]
                err = max(rel_err_mom(qg[m:m + mz, m:m + my, i], row) for i in range(m, m + mx))
            print(f"  step {n + 1}: worst row err {err:.3e}", flush=True)
            assert err <= tol
            del qg
    finally:
        d.close()


@functools.lru_cache(maxsize=1)
def _window_run(i0, j0, nw, mz, nsteps):
    print(f"[plume-window] oracle on {nw}x{nw}x{mz} at ({i0}, {j0})", flush=True)
    c, p, g, aux, q = plume_window(i0, j0, nw, nw, mz)
    return aux, _oracle_run(c, p, g, aux, q, nsteps)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_full_domain_plume_matches_oracle_window(form):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64
    mx, my, mz = 250, 200, 300
    nw, nsteps = 40, 10
    i0, j0 = mx // 2 - nw // 2, my // 2 - nw // 2
    auxw, ref = _window_run(i0, j0, nw, mz, nsteps)
    c, p, g, aux, q = plume_window(0, 0, mx, my, mz)
    print(f"[plume-window {form}] device on {mx}x{my}x{mz}", flush=True)
    d = _dev_form(g, p, aux, q, FORM_FP64 if form == "fp64" else FORM_FP32P)
    del q
    m = g.mbc
    r = WINDOW_MARGIN
    dtv = list(c.dtv)
    t = 0.0
    try:
        for n in range(nsteps):
            t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [c.nv[0], 0], c.method, c.mthbc)
            t_o, dt_o, q_o = ref[n]
            dtol = 1e-12 if form == "fp64" else 1e-5
            assert abs(t - t_o) <= dtol * t_o, (n, t, t_o)
            assert abs(dtv[4] - dt_o) <= dtol * dt_o, (n, dtv[4], dt_o)
            dtv[0] = dtv[4]
        qg = d.get_q()
    finally:
        d.close()
    # the window's inner n x n columns with m ghost cells around them, cut
    # from the full run and from the window run alike
    n = nw - 2 * r
    a = qg[:, j0 + r:j0 + r + n + 2 * m, i0 + r:i0 + r + n + 2 * m]
    b = q_o[:, r:r + n + 2 * m, r:r + n + 2 * m]
    auxb = auxw[:, r:r + n + 2 * m, r:r + n + 2 * m]
    inner = (slice(m, m + mz), slice(m, m + n), slice(m, m + n))
    w_peak = float(np.abs(b[inner][..., 3] / b[inner][..., 0]).max())
    err = rel_err_mom(a[inner], b[inner])
    fe = form_error(dataclasses.replace(g, mx=n, my=n), a, b, auxb)
    print(f"  t = {t:.2f} s, peak |w| {w_peak:.3e} m/s, state err {err:.3e}, "
          + " ".join(f"{k} {e['max_rel']:.2e}" for k, e in fe.items()), flush=True)
    assert w_peak > 0.1            # the plume has set the fluid moving
    if form == "fp64":
        assert err <= 1e-11, err
        for k in ("rho'", "T'", "w"):
            assert fe[k]["max_rel"] <= 1e-8, (k, fe[k])
    else:
        for k in ("rho'", "T'", "w"):
            assert fe[k]["max_rel"] <= 1e-4, (k, fe[k])
    # 40 or more cells from the plume every column equals the corner column
    # to round-off.  The momenta there are the rest state's well-balanced
    # residual (about eps * p per interface and step, summed in an order that
    # depends on the column's place in its tile), so they are measured
    # against the flow's largest momentum, as above.
    qi = qg[m:m + mz, m:m + my, m:m + mx]
    corner = qi[:, :1, :1]
    ii = np.abs(np.arange(mx) - mx // 2) >= nw
    jj = np.abs(np.arange(my) - my // 2) >= nw
    scale = np.abs(b[inner]).reshape(-1, g.meqn).max(0)
    scale[1:4] = scale[1:4].max()
    dev = np.maximum((np.abs(qi[:, jj] - corner) / scale).reshape(-1, g.meqn).max(0),
                     (np.abs(qi[:, :, ii] - corner) / scale).reshape(-1, g.meqn).max(0))
    print("  far field: largest column deviation per component "
          + " ".join(f"{v:.2e}" for v in dev), flush=True)
    assert dev.max() <= 1e-12, dev
