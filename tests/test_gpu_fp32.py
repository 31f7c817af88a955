"""The FP32 perturbation form (MAGIC_FORM_FP32P) against the FP64 answer.

Bar: the errors of w and T' stay within 1e-4 of the field's peak (rho' is
reported and held to the same bar), measured by magic_master_amd.report:
  * small grids, 10 claw3 steps, against the C oracle (FP64);
  * the full-domain workload (250x200x300) with the plume on from t = 0 and
    peak amplitudes 5, 20 and 60 mm/10 min, against the FP64 form on the GPU
    (itself pinned to the oracle), at every 60 s of model time up to 120 s;
  * the full-domain workload as stated (plume peak at 2400 s) through its
    500 steps, with the compensated global sums within 1e-9.
The dt sequences agree to the FP32 rounding of the wave speeds (1e-5).
"""
import copy

import numpy as np
import pytest

from magic_master_amd import problem as P
from magic_master_amd.report import form_error
from oracle.oracle import COracle
from tests.helpers import CASES, perturbed

pytestmark = pytest.mark.gpu
TOL = 1e-4


def _dev(g, p, aux, q, form):
    from magic_master_amd.hip import MagicHIP
    return MagicHIP(g, p, aux, q, form=form)


def _check(err, where):
    for k, e in err.items():
        print(f"  {where} {k}: max {e['max']:.3e} rms {e['rms']:.3e} peak {e['peak']:.3e} "
              f"max/peak {e['max_rel']:.2e}", flush=True)
    for k in ("w", "T'", "rho'"):
        assert err[k]["max_rel"] <= TOL, (where, k, err[k])


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp32_form_small_grids_vs_oracle(case):
    from magic_master_amd.hip import FORM_FP32P
    (mx, my, mz), mthbc = CASES[case]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=case, vel=0.05)
    print(f"[fp32 small] case {case} grid {mx}x{my}x{mz}", flush=True)
    O = COracle()
    d = _dev(g, p, aux, q, FORM_FP32P)
    qo = q.copy()
    dto = [50.0, 10.0, 0, 0, 0]
    dtg = list(dto)
    t_o = t_g = 0.0
    try:
        for n in range(10):
            st, t_o, dto, _, _, _ = O.claw3(g, p, qo, aux, t_o, t_o - 1.0, dto, c.cflv, [100, 0],
                                            c.method, mthbc)
            assert st == 0
            t_g, dtg, _, _, _ = d.claw3(t_g, t_g - 1.0, dtg, c.cflv, [100, 0], c.method, mthbc)
            assert abs(dtg[4] - dto[4]) <= 1e-5 * dto[4], (n, dtg[4], dto[4])
            dto[0] = dto[4]
            dtg[0] = dtg[4]
        qg = d.get_q()
    finally:
        d.close()
    _check(form_error(g, qg, qo, aux), f"case {case}")


@pytest.fixture(scope="module")
def full_domain():
    c, p, g, aux, q = P.config_full()
    return c, p, g, aux, q


@pytest.mark.timeout(600)
@pytest.mark.parametrize("amplitude", [5.0, 20.0, 60.0])
def test_fp32_form_full_domain_vs_fp64(full_domain, amplitude):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64
    c, p0, g, aux, q = full_domain
    p = copy.copy(p0)
    p.amplitude = amplitude
    p.tcenter = 0.0
    print(f"[fp32 full] amplitude {amplitude}", flush=True)
    runs = {}
    for form in (FORM_FP64, FORM_FP32P):
        d = _dev(g, p, aux, q, form)
        t = 0.0
        dtv = list(c.dtv)
        snaps = []
        try:
            for t_out in (60.0, 120.0):
                while t < t_out - 1e-9:
                    t, dtv, _, _, _ = d.claw3(t, t_out, dtv, c.cflv, [100000, 0], c.method, c.mthbc)
                    dtv[0] = dtv[4]
                snaps.append((t, d.get_q()))
        finally:
            d.close()
        runs[form] = snaps
        print(f"  form {form}: reached t = {t:.3f}", flush=True)
    for (t64, q64), (t32, q32) in zip(runs[FORM_FP64], runs[FORM_FP32P]):
        assert abs(t64 - t32) <= 1e-9 * t64
        _check(form_error(g, q32, q64, aux), f"amp {amplitude} t {t64:.0f}s")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("amplitude", [5.0, 20.0, 60.0])
def test_fp32_form_full_domain_500_steps(full_domain, amplitude):
    """The full-domain workload as stated (plume peak at t = 2400 s) through
    its 500 steps: both forms to t = 780 s (about 560 steps; the amplitude-5
    plume's fields leave round-off level between 720 and 780 s), then w, T'
    and rho' within 1e-4 of their FP64 peaks and the compensated global mass
    and energy within 1e-9."""
    from magic_master_amd.report import compare_forms
    c, p0, g, aux, q = full_domain
    p = copy.copy(p0)
    p.amplitude = amplitude
    print(f"[fp32 full 500 steps] amplitude {amplitude}", flush=True)
    rows = compare_forms(c, p, g, aux, q, 780.0, 780.0)
    r = rows[-1]
    print(f"  t {r['t']:.0f} s steps {r['steps']}: " +
          ", ".join(f"{k} {r[k]['max_rel']:.2e} of peak {r[k]['peak']:.2e}" for k in ("w", "T'", "rho'")) +
          f", mass {r['mass']['rel']:.1e}, energy {r['energy']['rel']:.1e}", flush=True)
    assert min(r["steps"]) >= 500 and r["steps"][0] == r["steps"][1]
    for k in ("w", "T'", "rho'"):
        assert r[k]["peak"] > 1e3 * {"w": 1e-6, "T'": 1e-6, "rho'": 1e-9}[k], (k, r[k])   # above round-off
        assert r[k]["max_rel"] <= TOL, (k, r[k])
    assert abs(r["mass"]["rel"]) <= 1e-9 and abs(r["energy"]["rel"]) <= 1e-9, r
