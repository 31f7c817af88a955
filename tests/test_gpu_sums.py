"""Compensated FP64 global sums (magic_global_sums) on the GPU.

* Against the exactly rounded sum (math.fsum) of the same values, on states
  with a dynamic range of 16 decades and random signs, where plain FP64
  summation loses digits: within a few units in the last place, in both
  forms (the FP32 form's background enters as exact per-layer products).
* Mass conservation: periodic x and y, walls at top and bottom, a blob
  away from the walls; the reference's scheme changes sum(rho) only through
  the z faces (flux differencing; nothing flows through a face at rest), so
  the FP64 form keeps the mass to round-off over the steps, and the FP32
  form's mass and energy agree with the FP64 form's to the FP32 form's
  precision.
"""
import math

import numpy as np
import pytest

from magic_master_amd import problem as P
from tests.helpers import perturbed

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _interior(g, q):
    m = g.mbc
    return q[m:m + g.mz, m:m + g.my, m:m + g.mx]


def _wide_range_state(g, aux, rng):
    """Every component random with magnitudes over 16 decades and random signs
    (the sums do not look at the physics)."""
    q = np.empty(g.shape_q)
    mag = 10.0 ** rng.uniform(-8, 8, size=g.shape_q)
    q[...] = rng.choice([-1.0, 1.0], size=g.shape_q) * mag * rng.random(g.shape_q)
    return q


@pytest.mark.parametrize("dims", [(12, 10, 8), (80, 60, 50)])
def test_fp64_sums_match_exactly_rounded_sums(dims):
    from magic_master_amd.hip import MagicHIP
    mx, my, mz = dims
    c, p, g, aux, _ = perturbed(mx, my, mz, [1, 1, 1, 1, 3, 1], seed=3)
    rng = np.random.default_rng(17)
    q = _wide_range_state(g, aux, rng)
    print(f"[sums fp64] {dims}", flush=True)
    d = MagicHIP(g, p, aux, q)
    try:
        s = d.global_sums()
    finally:
        d.close()
    vol = g.dx * g.dy * g.dz
    it = _interior(g, q).reshape(-1, g.meqn)
    for m in range(g.meqn):
        exact = math.fsum(it[:, m]) * vol
        naive = float(np.sum(it[:, m])) * vol
        err = abs(s[m] - exact)
        print(f"  m={m}: exact {exact:.17e} err {err:.2e} naive err {abs(naive - exact):.2e}", flush=True)
        assert err <= 3 * EPS * abs(exact), (m, s[m], exact)


def test_fp32_form_sums_match_exactly_rounded_sums():
    """FP32 form: the full state is the FP64 background column plus FP32
    perturbations; the device sums equal the exactly rounded sum of those
    values."""
    from magic_master_amd.hip import FORM_FP32P, MagicHIP
    c, p, g, aux, q = P.config_plain(mx=40, my=36, mz=30)
    rng = np.random.default_rng(5)
    m = g.mbc
    col = aux[:, 0, 0, :]
    # perturbations exactly representable in FP32, 1e-3 of the background
    pert = np.zeros(g.shape_q, dtype=np.float32)
    pert[..., 0] = (1e-3 * col[:, None, None, 0] * rng.standard_normal(g.shape_q[:3])).astype(np.float32)
    pert[..., 1:4] = rng.standard_normal(g.shape_q[:3] + (3,)).astype(np.float32)
    pert[..., 4] = (1e-3 * col[:, None, None, 4] * rng.standard_normal(g.shape_q[:3])).astype(np.float32)
    pert[..., 5] = rng.random(g.shape_q[:3]).astype(np.float32)
    qf = pert.astype(np.float64)
    qf[..., 0] += col[:, None, None, 0]
    qf[..., 4] += col[:, None, None, 4]
    print("[sums fp32]", flush=True)
    d = MagicHIP(g, p, aux, qf, form=FORM_FP32P)
    try:
        s = d.global_sums()
    finally:
        d.close()
    vol = g.dx * g.dy * g.dz
    nxy = g.mx * g.my
    # what the device holds: (q - background) in FP64, rounded to FP32
    dev = qf.copy()
    dev[..., 0] -= col[:, None, None, 0]
    dev[..., 4] -= col[:, None, None, 4]
    pi = _interior(g, dev.astype(np.float32)).reshape(-1, g.meqn).astype(np.float64)
    for comp in range(g.meqn):
        vals = list(pi[:, comp])
        if comp in (0, 4):
            for k in range(1, g.mz + 1):
                vals += [col[k + m - 1, comp]] * nxy
        exact = math.fsum(vals) * vol
        assert abs(s[comp] - exact) <= 3 * EPS * abs(exact), (comp, s[comp], exact)


def _blob_state(mx, my, mz):
    """Isothermal background, periodic sides, walls top and bottom, no
    forcing, no diffusion and no noise (the fluid at the walls at rest); a
    density/temperature blob and a swirl in the middle, 8 or more cells from
    the walls."""
    c, p, g, aux, q = P.config_plain(mx=mx, my=my, mz=mz, noise=0.0)
    c.mthbc = [2, 2, 2, 2, 3, 3]
    p.forcemth = 0.0
    aux[..., 10] = 0.0
    aux[..., 11] = 0.0
    K, J, I = q.shape[:3]
    k = np.arange(K)[:, None, None] - K / 2
    j = np.arange(J)[None, :, None] - J / 2
    i = np.arange(I)[None, None, :] - I / 2
    blob = np.exp(-(i ** 2 + j ** 2) / 18.0 - k ** 2 / 4.0)
    rho = q[..., 0]
    q[..., 0] = rho * (1 + 1e-2 * blob)
    q[..., 1] = 3.0 * j / 3.0 * blob * q[..., 0]
    q[..., 2] = -3.0 * i / 3.0 * blob * q[..., 0]
    q[..., 4] = aux[..., 4] * (1 + 2e-2 * blob) + 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2) / q[..., 0]
    return c, p, g, aux, q


def _run(form, nsteps):
    from magic_master_amd.hip import MagicHIP
    c, p, g, aux, q = _blob_state(24, 24, 24)
    d = MagicHIP(g, p, aux, q, form=form)
    try:
        sums = [d.global_sums()]
        dtv = list(c.dtv)
        dtv[0] = 3.0
        t = 0.0
        for _ in range(nsteps):
            t, dtv, _, _, _ = d.claw3(t, t - 1.0, dtv, c.cflv, [100, 0], c.method, c.mthbc)
            dtv[0] = dtv[4]
            sums.append(d.global_sums())
        return g, aux, np.array(sums), d.get_q()
    finally:
        d.close()


def test_mass_held_and_forms_agree():
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64
    print("[sums conservation]", flush=True)
    # 5 steps: the blob's acoustic front stays below 1e-12 of the density at
    # the walls (the oracle's run of the same state holds its mass to 7e-16)
    g, aux, s64, q64 = _run(FORM_FP64, 5)
    _, _, s32, q32 = _run(FORM_FP32P, 5)
    vol = g.dx * g.dy * g.dz
    m0 = s64[0, 0]
    drift = np.abs(s64[:, 0] - m0).max() / m0
    moved = np.abs(_interior(g, q64)[..., 0] - _interior(g, aux)[..., 0]).sum() * vol
    print(f"  FP64 mass {m0:.17e}, largest drift {drift:.2e} of it; FP32 mass drift "
          f"{np.abs(s32[:, 0] - s32[0, 0]).max() / m0:.2e}", flush=True)
    assert np.abs(_interior(g, q64)[..., 3]).max() > 1e-3 * np.abs(_interior(g, q64)[..., 1]).max()
    assert drift <= 1e-14
    # FP32 form: the same state to its own precision, the perturbation mass
    # (about `moved`) held to FP32 rounding
    assert np.abs(s32[:, 0] - s64[:, 0]).max() <= 1e-6 * moved + 1e-14 * m0
    e_pert = np.abs(_interior(g, q64)[..., 4] - _interior(g, aux)[..., 4]).sum() * vol
    assert np.abs(s32[:, 4] - s64[:, 4]).max() <= 1e-5 * e_pert + 1e-14 * abs(s64[0, 4])
