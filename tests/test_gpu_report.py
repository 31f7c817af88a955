"""The error report's comparisons (magic_master_amd/report.py) on small
workloads:

* compare_runs with two FP64 orders of the sweeps' additions (overlap on and
  off): the same dt sequence and fields within the FP64 bar (1e-12 of each
  field's peak) at every output time;
* compare_ensemble: a small seeded ensemble advanced by the batched driver
  in the FP32 form, each scenario against the FP64 form stepped alone to
  the same model time, within the FP32 bar (1e-4 of each field's peak, for
  the fields whose peak is above round-off level);
* layer_error: the same comparison of one scenario stepped alone, layer by
  layer.
"""
import pytest

from magic_master_amd import report
from tests.helpers import CASES, perturbed

pytestmark = pytest.mark.gpu


def test_compare_runs_two_fp64_orders():
    from magic_master_amd.hip import FORM_FP64
    (mx, my, mz), mthbc = CASES[0]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=91, vel=0.3)
    print(f"[report order] grid {mx}x{my}x{mz}", flush=True)
    rows = report.compare_runs(c, p, g, aux, q, 60.0, 20.0, (FORM_FP64, True), (FORM_FP64, False))
    assert [r["t"] for r in rows] == [20.0, 40.0, 60.0]
    for r in rows:
        assert r["steps"][0] == r["steps"][1] > 0
        for k in ("w", "T'", "rho'"):
            assert r[k]["max_rel"] <= 1e-12, (r["t"], k, r[k])


def test_compare_ensemble_small():
    print("[report ensemble] 3 scenarios of 24-32 x 24-32 x 32-40 cells to 150 s", flush=True)
    summary, rows = report.compare_ensemble(nscen=3, t_end=150.0, check=3, seed=7, nmin=(24, 24, 32),
                                            nmax=(32, 32, 40))
    assert summary["scenarios"] == 3 and summary["rejected"] >= 0 and summary["attempts"] > 0
    assert len(rows) == 3
    for r in rows:
        assert r["steps"][0] > 0 and abs(r["steps"][0] - r["steps"][1]) <= 1, r["steps"]
        for k, floor in report.PEAK_FLOOR.items():
            print(f"  scenario {r['scenario']} {k}: peak {r[k]['peak']:.3e} max/peak {r[k]['max_rel']:.2e}",
                  flush=True)
            if r[k]["peak"] >= floor:
                assert r[k]["max_rel"] <= 1e-4, (r["scenario"], k, r[k])


def test_layer_error_small():
    """layer_error: one row per interior layer, the other FP64 order within
    1e-9 of each field's peak in every layer, the FP32 form within 1e-4 of each live
    field's peak, and each field's worst cell inside the grid."""
    print("[report layers] scenario 1 of 2 (seed 7, 24-32 x 24-32 x 32-40 cells) to 150 s", flush=True)
    r = report.layer_error(1, nscen=2, seed=7, t_end=150.0, nmin=(24, 24, 32), nmax=(32, 32, 40))
    mx, my, mz = r["grid"]
    assert len(r["layers"]) == mz and [row["k"] for row in r["layers"]] == list(range(mz))
    rho = [row["rho_bar"] for row in r["layers"]]
    assert all(a > b > 0 for a, b in zip(rho, rho[1:]))       # hydrostatic column: density falls with height
    for k, floor in report.PEAK_FLOOR.items():
        kk, jj, ii = r["worst"][k]
        assert 0 <= kk < mz and 0 <= jj < my and 0 <= ii < mx
        if r["peak"][k] > 0:
            assert max(row[k]["value"] for row in r["layers"]) == pytest.approx(1.0)
        assert max(row[k]["order"] for row in r["layers"]) <= 1e-9, k
        if r["peak"][k] >= floor:
            assert max(row[k]["fp32"] for row in r["layers"]) <= 1e-4, k
