"""Error report of a reduced-precision run against the FP64 form.

For each output time the FP32 perturbation form is compared with the FP64
form on the same inputs through the fields a user of MAGIC looks at: the
density perturbation rho' = rho - rho_bar, the temperature perturbation
T' = T - T_bar with T = (gamma-1)(E - |m|^2/(2 rho)) / (rho R) (the state
equation of src3.f90:242-304), and the vertical velocity w = m_z / rho.
Errors are reported as max and RMS of the difference, absolute and
relative to the field's peak |value| in the FP64 run, next to both runs'
global mass and energy (compensated FP64 sums over the interior).
"""
from __future__ import annotations

import numpy as np

from .problem import AUX_E, AUX_GAMMA, AUX_R, AUX_RHO, Grid


def fields(g: Grid, q: np.ndarray, aux: np.ndarray) -> dict:
    """rho', T', w on the interior of a reference-layout q (FP64)."""
    m = g.mbc
    sl = (slice(m, m + g.mz), slice(m, m + g.my), slice(m, m + g.mx))
    qi, ai = q[sl], aux[sl]
    rho = qi[..., 0]
    ke = 0.5 * (qi[..., 1] ** 2 + qi[..., 2] ** 2 + qi[..., 3] ** 2) / rho
    g1 = ai[..., AUX_GAMMA - 1] - 1.0
    R = ai[..., AUX_R - 1]
    rb = ai[..., AUX_RHO - 1]
    T = g1 * (qi[..., 4] - ke) / (rho * R)
    Tb = g1 * ai[..., AUX_E - 1] / (rb * R)
    return {"rho'": rho - rb, "T'": T - Tb, "w": qi[..., 3] / rho}


def form_error(g: Grid, q: np.ndarray, q_ref: np.ndarray, aux: np.ndarray) -> dict:
    """{field: {max, rms, peak, max_rel, rms_rel}} of q against q_ref."""
    a, b = fields(g, q, aux), fields(g, q_ref, aux)
    out = {}
    for k in a:
        d = a[k] - b[k]
        peak = float(np.abs(b[k]).max())
        mx = float(np.abs(d).max())
        rms = float(np.sqrt(np.mean(d * d)))
        out[k] = {"max": mx, "rms": rms, "peak": peak,
                  "max_rel": mx / peak if peak > 0 else float("inf") if mx > 0 else 0.0,
                  "rms_rel": rms / peak if peak > 0 else float("inf") if rms > 0 else 0.0}
    return out


def sums_row(ref: np.ndarray, other: np.ndarray) -> dict:
    """The compensated FP64 global mass and energy (magic_global_sums: kg and
    J over the interior) of two runs and the relative difference of the
    second from the first."""
    out = {}
    for name, m in (("mass", 0), ("energy", 4)):
        r, o = float(ref[m]), float(other[m])
        out[name] = {"ref": r, "other": o, "rel": (o - r) / r if r != 0 else float("inf")}
    return out


# Peaks below which a field is at round-off level (fluid at rest): w in m/s,
# T' in K, rho' in kg/m^3.
PEAK_FLOOR = {"w": 1e-6, "T'": 1e-6, "rho'": 1e-9}


def compare_runs(c, p, g: Grid, aux: np.ndarray, q: np.ndarray, t_end: float, every: float, ref, other,
                 on_row=None) -> list:
    """Advance two configurations of the step, ref and other, each (form,
    sweep overlap on/off), from the same state under claw3's step control, to
    every multiple of `every` model seconds up to t_end (each claw3 call
    stops exactly at the output time, as the claw3ez frame loop does), and
    compare other against ref there.  Returns one row per output time:
    {"t", "steps" (ref, other), field: form_error entry}."""
    from .hip import MagicHIP
    cfgs = (ref, other)
    devs = []
    rows = []
    try:
        for form, overlap in cfgs:
            d = MagicHIP(g, p, aux, q, form=form)
            devs.append(d)
            d.set_sweep_overlap(overlap)
        t = [0.0, 0.0]
        dtv = [list(c.dtv), list(c.dtv)]
        steps = [0, 0]
        k = 1
        while k * every <= t_end * (1 + 1e-12):
            t_out = k * every
            for n, d in enumerate(devs):
                while t[n] < t_out * (1 - 1e-12):
                    t[n], dtv[n], _, nv, _ = d.claw3(t[n], t_out, dtv[n], c.cflv, [1 << 30, 0], c.method, c.mthbc)
                    steps[n] += nv[1]
                    dtv[n][0] = dtv[n][4]
            row = {"t": t_out, "steps": list(steps)}
            row.update(form_error(g, devs[1].get_q(), devs[0].get_q(), aux))
            row.update(sums_row(devs[0].global_sums(), devs[1].global_sums()))
            rows.append(row)
            if on_row:
                on_row(row)
            k += 1
    finally:
        for d in devs:
            d.close()
    return rows


def layer_error(scenario: int, nscen: int = 64, seed: int = 2024, t_end: float = 420.0,
                nmin=(64, 64, 96), nmax=(200, 200, 300)) -> dict:
    """Where in the column the FP32 form's error sits: scenario `scenario`
    of problem.ensemble(nscen, seed) stepped alone from rest to t_end in
    three configurations, the FP64 form with the sweep overlap (the
    reference), the FP64 form with the sweeps in turn (another order of
    additions) and the FP32 form with the sweeps in turn (the batch's
    order).  Returns {"grid", "amplitude", "peak": {field: FP64 peak},
    "worst": {field: (k, j, i) of the FP32 form's largest error},
    "layers": one row per interior layer k: {"k", "rho_bar", field:
    {"fp32", "order", "value"}}, each the layer's largest |difference| (or
    |value| of the FP64 field) over the field's peak}."""
    from . import problem as P
    from .hip import FORM_FP32P, FORM_FP64, MagicHIP
    c, p, g, col = P.ensemble(nscen, seed, nmin=nmin, nmax=nmax)[scenario]
    aux = np.ascontiguousarray(np.broadcast_to(col[:, None, None, :], g.shape_aux))
    q0 = P.qinit(g, aux)
    out = []
    for form, overlap in ((FORM_FP64, True), (FORM_FP64, False), (FORM_FP32P, False)):
        d = MagicHIP(g, p, aux, q0, form=form)
        try:
            d.set_sweep_overlap(overlap)
            t, dtv = 0.0, list(c.dtv)
            while t < t_end * (1 - 1e-12):
                t, dtv, _, _, _ = d.claw3(t, t_end, dtv, c.cflv, [1 << 30, 0], c.method, c.mthbc)
                dtv[0] = dtv[4]
            out.append(fields(g, d.get_q(), aux))
        finally:
            d.close()
    ref, order, f32 = out
    rho_bar = aux[g.mbc:g.mbc + g.mz, 0, 0, AUX_RHO - 1]
    peak, worst, per = {}, {}, {}
    for k in ref:
        peak[k] = float(np.abs(ref[k]).max())
        d32, dor = np.abs(f32[k] - ref[k]), np.abs(order[k] - ref[k])
        worst[k] = tuple(int(v) for v in np.unravel_index(np.argmax(d32), d32.shape))
        scale = peak[k] if peak[k] > 0 else 1.0
        per[k] = [a.reshape(g.mz, -1).max(1) / scale for a in (d32, dor, np.abs(ref[k]))]
    layers = [{"k": kk, "rho_bar": float(rho_bar[kk]),
               **{f: {"fp32": float(per[f][0][kk]), "order": float(per[f][1][kk]), "value": float(per[f][2][kk])}
                  for f in ref}} for kk in range(g.mz)]
    return {"grid": [g.mx, g.my, g.mz], "amplitude": p.amplitude, "peak": peak, "worst": worst, "layers": layers}


def compare_forms(c, p, g: Grid, aux: np.ndarray, q: np.ndarray, t_end: float, every: float,
                  on_row=None) -> list:
    """The FP32 perturbation form against the FP64 form (compare_runs; both
    with the default sweep overlap)."""
    from .hip import FORM_FP32P, FORM_FP64
    return compare_runs(c, p, g, aux, q, t_end, every, (FORM_FP64, True), (FORM_FP32P, True), on_row)


def compare_ensemble(nscen: int = 64, t_end: float = 400.0, check: int = 4, seed: int = 2024, shapes=None,
                     on_row=None, nmin=(64, 64, 96), nmax=(200, 200, 300)):
    """The batched ensemble workload (problem.ensemble(nscen, seed)): every
    scenario advanced by the batched driver in the FP32 form from rest to
    model time t_end (claw3 intervals, clipped there), then the first
    `check` scenarios stepped alone in the FP64 form to the same time and
    compared (nmin, nmax: the grid-size range of problem.ensemble).  Returns
    (summary of the batch run, one row per checked scenario: {"scenario",
    "cells", "steps" (FP32, FP64), field: form_error entry})."""
    import time
    from . import problem as P
    from .hip import FORM_FP64, MagicBatch, MagicHIP
    ens = P.ensemble(nscen, seed, nmin=nmin, nmax=nmax)
    b = MagicBatch([e[2] for e in ens], [e[1] for e in ens], [e[3] for e in ens], None)
    try:
        if shapes is not None:
            b.set_tile_shapes(shapes)
        for s, (c, p, g, col) in enumerate(ens):
            b.start(s, 0.0, t_end, c.dtv[0], c.dtv, c.cflv, 1 << 30, c.mthbc)
        t0 = time.perf_counter()
        attempts = b.run(max_attempts=1 << 30, check_every=8)
        wall = time.perf_counter() - t0
        states = [b.state(s) for s in range(nscen)]
        q32 = [b.get_q(s) for s in range(min(check, nscen))]
        s32 = [b.global_sums(s) for s in range(min(check, nscen))]
    finally:
        b.close()
    cell_steps = sum(e[2].mx * e[2].my * e[2].mz * st["nsteps"] for e, st in zip(ens, states))
    summary = {"scenarios": nscen, "t_end": t_end, "attempts": attempts, "wall_s": wall,
               "steps": [min(st["nsteps"] for st in states), max(st["nsteps"] for st in states)],
               "rejected": sum(st["rejects"] for st in states), "mcell_updates_per_s": cell_steps / wall / 1e6}
    rows = []
    for s in range(min(check, nscen)):
        c, p, g, col = ens[s]
        aux = np.ascontiguousarray(np.broadcast_to(col[:, None, None, :], g.shape_aux))
        d = MagicHIP(g, p, aux, P.qinit(g, aux), form=FORM_FP64)
        try:
            t, dtv, n64 = 0.0, list(c.dtv), 0
            while t < t_end * (1 - 1e-12):
                t, dtv, _, nv, _ = d.claw3(t, t_end, dtv, c.cflv, [1 << 30, 0], c.method, c.mthbc)
                n64 += nv[1]
                dtv[0] = dtv[4]
            q64 = d.get_q()
            s64 = d.global_sums()
        finally:
            d.close()
        row = {"scenario": s, "cells": [g.mx, g.my, g.mz], "steps": [states[s]["nsteps"], n64]}
        row.update(form_error(g, q32[s], q64, aux))
        row.update(sums_row(s64, s32[s]))
        rows.append(row)
        if on_row:
            on_row(row)
    return summary, rows


def main(argv=None):
    """Error report of the FP32 perturbation form on the full-domain
    workload (250x200x300), one run per plume amplitude: max and RMS of
    rho', T' and w against the FP64 form every `--every` model seconds
    (--pair order: two FP64 runs that differ only in the order of the
    sweeps' additions)."""
    import argparse
    import copy
    import json
    from . import problem as P
    ap = argparse.ArgumentParser(prog="python -m magic_master_amd.report", description=main.__doc__)
    ap.add_argument("--amplitudes", default="5,20,60", help="plume peak amplitudes, mm/10 min")
    ap.add_argument("--t-end", type=float, default=3000.0, help="model seconds")
    ap.add_argument("--every", type=float, default=60.0, help="output interval, model seconds")
    ap.add_argument("--tcenter", type=float, default=None,
                    help="plume peak time (default: the workload's 2400 s)")
    ap.add_argument("--json", default=None, help="also write every row to this JSON file")
    ap.add_argument("--pair", default="fp32", choices=("fp32", "order"),
                    help="fp32: the FP32 form against the FP64 form; order: the FP64 form with the sweeps "
                         "in turn against the FP64 form with the sweep overlap (the same arithmetic in "
                         "another order of additions: how far rounding alone moves the solution)")
    ap.add_argument("--ensemble", type=int, default=0,
                    help="instead: the batched FP32 ensemble of this many seeded scenarios to --t-end, the "
                         "first --check of them against the FP64 form stepped alone")
    ap.add_argument("--check", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2024, help="seed of the ensemble's scenarios")
    ap.add_argument("--layers", type=int, default=None, metavar="S",
                    help="instead: layer by layer, the FP32 form's error in scenario S of the ensemble "
                         "(--ensemble scenarios, default 64; --seed) stepped alone to --t-end, next to "
                         "the FP64 form's with the other order of additions")
    a = ap.parse_args(argv)
    if a.layers is not None:
        r = layer_error(a.layers, a.ensemble or 64, a.seed, a.t_end)
        mz = r["grid"][2]
        print(f"scenario {a.layers} of seed {a.seed}: grid {r['grid']}, amplitude {r['amplitude']:.2f} mm/10 min, "
              f"t = {a.t_end:g} s; peaks " + ", ".join(f"{k} {v:.4g}" for k, v in r["peak"].items())
              + "; FP32 worst at (k, j, i) " + ", ".join(f"{k} {v}" for k, v in r["worst"].items()), flush=True)
        print("    k   rho_bar   | per field: max |FP32 - FP64| / peak, max |other order - FP64| / peak, "
              "max |FP64| / peak", flush=True)
        for row in r["layers"][::max(1, mz // 24)]:
            print(f"  {row['k']:3d} {row['rho_bar']:.3e} | " + " | ".join(
                f"{k} {row[k]['fp32']:.2e} {row[k]['order']:.2e} {row[k]['value']:.2e}" for k in ("T'", "w", "rho'")),
                flush=True)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(r, f, indent=1)
        return
    from .hip import FORM_FP32P, FORM_FP64
    if a.ensemble > 0:
        def show_s(r):
            print(f"scenario {r['scenario']} cells {r['cells']} steps (fp32/fp64) {r['steps']}: " + " | ".join(
                f"{k}: peak {r[k]['peak']:.3e} max/peak {r[k]['max_rel']:.2e} rms/peak {r[k]['rms_rel']:.2e}"
                for k in ("w", "T'", "rho'")) + f" | mass {r['mass']['rel']:+.2e} energy {r['energy']['rel']:+.2e}",
                flush=True)
        summary, rows = compare_ensemble(a.ensemble, a.t_end, a.check, seed=a.seed, on_row=show_s)
        print("ensemble: " + json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in summary.items()}),
              flush=True)
        if a.json:
            with open(a.json, "w") as f:
                json.dump({"summary": summary, "rows": rows}, f, indent=1)
        return
    pair = ((FORM_FP64, True), (FORM_FP32P, True)) if a.pair == "fp32" else ((FORM_FP64, True), (FORM_FP64, False))
    c, p0, g, aux, q = P.config_full()
    out = {}
    for amp in [float(v) for v in a.amplitudes.split(",")]:
        p = copy.copy(p0)
        p.amplitude = amp
        if a.tcenter is not None:
            p.tcenter = a.tcenter
        print(f"amplitude {amp} mm/10 min, plume peak at t = {p.tcenter} s", flush=True)
        print("   t(s) steps(ref/other) | w: peak(m/s) max/peak rms/peak | T': peak(K) max/peak rms/peak | "
              "rho': peak(kg/m3) max/peak rms/peak | mass(kg) other/ref-1 | energy(J) other/ref-1", flush=True)

        def show(r):
            print(f"{r['t']:7.0f} {r['steps'][0]:5d}/{r['steps'][1]:5d} | " + " | ".join(
                f"{r[k]['peak']:.3e} {r[k]['max_rel']:.2e} {r[k]['rms_rel']:.2e}" for k in ("w", "T'", "rho'"))
                + f" | {r['mass']['ref']:.10e} {r['mass']['rel']:+.2e} | {r['energy']['ref']:.10e} "
                  f"{r['energy']['rel']:+.2e}", flush=True)
        out[amp] = compare_runs(c, p, g, aux, q, a.t_end, a.every, pair[0], pair[1], on_row=show)
        # a field whose FP64 peak is at round-off level (the fluid still at
        # rest) has no meaningful relative error; the worst case is taken over
        # the output times where every field's peak is above its floor
        live = [r for r in out[amp] if all(r[k]["peak"] >= PEAK_FLOOR[k] for k in PEAK_FLOOR)]
        if live:
            worst = {k: max(r[k]["max_rel"] for r in live) for k in PEAK_FLOOR}
            print(f"  worst max/peak from t = {live[0]['t']:.0f} s (peaks above "
                  + ", ".join(f"{k} {v:g}" for k, v in PEAK_FLOOR.items()) + "): "
                  + ", ".join(f"{k} {v:.3e}" for k, v in worst.items()), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump({str(k): v for k, v in out.items()}, f, indent=1)


if __name__ == "__main__":
    main()
