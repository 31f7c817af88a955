"""The FP32 form on the batched-ensemble workload at its own sizes, against
the FP64 form.

* Batch: the first 4 seeded scenarios of the ensemble workload
  (problem.ensemble: grids of 64-200 x 64-200 x 96-300 cells, plume
  amplitudes 2-60 mm/10 min) advanced together by the batched driver in the
  FP32 perturbation form to t = 420 s (about 300 steps each), every one
  against the same grid stepped alone in the FP64 form to the same time:
  |w| and |T'| within 1e-4 of their FP64 peaks, rho' within 1e-4, and the
  compensated global mass and energy within 1e-9 relative.
* Serving: requests drawn like the ensemble workload with end times in
  [300, 420] s and a 60-s snapshot cadence, served by the continuous-
  batching loop; each request's final state against the FP64 form stepped
  alone through the same snapshot intervals (claw3 stops at each snapshot
  time, as the served run does), to the same bars.
"""
import numpy as np
import pytest

from magic_master_amd import problem as P
from magic_master_amd import serve
from magic_master_amd.report import compare_ensemble, form_error

pytestmark = pytest.mark.gpu

BAR = 1e-4


def _fp64_through(g, p, col, c, t_end, cadence=None):
    """The FP64 form stepped alone from rest to t_end, stopping at every
    multiple of cadence (None: one interval)."""
    from magic_master_amd.hip import FORM_FP64, MagicHIP
    aux = np.ascontiguousarray(np.broadcast_to(col[:, None, None, :], g.shape_aux))
    d = MagicHIP(g, p, aux, P.qinit(g, aux), form=FORM_FP64)
    try:
        t, dtv = 0.0, list(c.dtv)
        while t < t_end * (1 - 1e-12):
            stop = t_end if cadence is None else min(t + cadence, t_end)
            while t < stop * (1 - 1e-12):
                t, dtv, _, _, _ = d.claw3(t, stop, dtv, c.cflv, [1 << 30, 0], c.method, c.mthbc)
                dtv[0] = dtv[4]
        return aux, d.get_q()
    finally:
        d.close()


def test_ensemble_scenarios_fp32_batch_vs_fp64():
    print("[config4 batch] 4 ensemble scenarios to t = 420 s", flush=True)
    summary, rows = compare_ensemble(nscen=4, t_end=420.0, check=4)
    print(summary, flush=True)
    for r in rows:
        print(f"  scenario {r['scenario']} cells {r['cells']} steps {r['steps']}: " +
              ", ".join(f"{k} {r[k]['max_rel']:.2e}" for k in ("w", "T'", "rho'")), flush=True)
        assert r["steps"][0] >= 200
        assert r["w"]["max_rel"] <= BAR and r["T'"]["max_rel"] <= BAR and r["rho'"]["max_rel"] <= BAR, r
        assert abs(r["mass"]["rel"]) <= 1e-9 and abs(r["energy"]["rel"]) <= 1e-9, r


def test_served_requests_fp32_vs_fp64():
    rq = serve.poisson_requests(2.0, 2.0, seed=8, t_end=(300.0, 420.0), cadence=60.0)
    assert len(rq) >= 3
    print(f"[config4 serve] {len(rq)} requests: {[(r.grid.mx, r.grid.my, r.grid.mz) for r in rq]}", flush=True)
    res = serve.Server(snapshots="none", keep_final=True).run(rq)
    print({k: v for k, v in res.items() if k != "outcomes"}, flush=True)
    for r, o in zip(sorted(rq, key=lambda r: r.arrival), res["outcomes"]):
        assert o.status == 0 and o.final_q is not None
        aux, q64 = _fp64_through(r.grid, r.prob, r.column, r.claw, r.t_end, r.cadence)
        err = form_error(r.grid, o.final_q, q64, aux)
        print(f"  request {r.rid} t_end {r.t_end:.1f}: " + ", ".join(f"{k} {v['max_rel']:.2e}"
                                                              for k, v in err.items()), flush=True)
        assert all(err[k]["max_rel"] <= BAR for k in ("w", "T'", "rho'")), err
