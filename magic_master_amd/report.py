"""Error report of a reduced-precision run against the FP64 form.

For each output time the FP32 perturbation form is compared with the FP64
form on the same inputs through the fields a user of MAGIC looks at: the
density perturbation rho' = rho - rho_bar, the temperature perturbation
T' = T - T_bar with T = (gamma-1)(E - |m|^2/(2 rho)) / (rho R) (the state
equation of src3.f90:242-304), and the vertical velocity w = m_z / rho.
Errors are reported as max and RMS of the difference, absolute and
relative to the field's peak |value| in the FP64 run.
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
