"""Known-answer and invariant checks that pin the C oracle independently of
the reference build: textbook identities of the Roe solver and transverse
splitting (LeVeque 2002, sec. 15.3), limiter consistency (Kemm 2011/2012
properties), the well-balanced hydrostatic rest state, the explicit
diffusion factor of a Fourier mode, the claw3 dt recursion, and the domain
of dependence the full-domain window test relies on.  Each check
is chosen so that a dropped term, a wrong sign or a transposed index fails."""
import math

import numpy as np
import pytest

from magic_master_amd import problem as P
from oracle.oracle import COracle
from tests.helpers import WINDOW_MARGIN, perturbed, plume_window


@pytest.fixture(scope="module")
def O():
    return COracle()


def _pencil(rng, N=24, meqn=5, gamma=1.4, speed=0.4):
    rho = 1.0 + 0.3 * rng.random(N)
    mom = speed * rng.standard_normal((N, 3)) * rho[:, None]
    T = 250 + 20 * rng.standard_normal(N)
    p = rho * 287.0 * T
    E = p / (gamma - 1.0) + 0.5 * (mom ** 2).sum(1) / rho
    q = np.zeros((N, meqn))
    q[:, 0], q[:, 1:4], q[:, 4] = rho, mom, E
    if meqn > 5:
        q[:, 5:] = rng.random((N, meqn - 5))
    return q


def _flux(q, ixyz, gamma=1.4):
    """Euler flux in direction ixyz (1, 2, 3) of conserved states (rows)."""
    mu = ixyz
    rho = q[:, 0]
    vel = q[:, 1:4] / rho[:, None]
    p = (gamma - 1.0) * (q[:, 4] - 0.5 * (q[:, 1:4] ** 2).sum(1) / rho)
    un = vel[:, mu - 1]
    f = np.zeros((q.shape[0], 5))
    f[:, 0] = q[:, mu]
    f[:, 1:4] = q[:, 1:4] * un[:, None]
    f[:, mu] += p
    f[:, 4] = (q[:, 4] + p) * un
    return f


@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_rpn3_fluctuations_sum_to_flux_jump(ixyz, O):
    """amdq + apdq = f(q_i) - f(q_{i-1}) (f-wave splitting of the flux jump,
    computed before limiting), and without gravity (aux17 = 0)."""
    rng = np.random.default_rng(ixyz)
    q = _pencil(rng)
    N = q.shape[0]
    st, fw, s, am, ap, roe = O.rpn3(ixyz, q, np.full(N, 1.4), np.zeros(N), 1.0, 1000.0, 0)
    assert st == 0
    f = _flux(q, ixyz)
    df = f[1:] - f[:-1]                       # interface i between cells i-1 and i
    tot = (am + ap)[1:]
    scale = np.abs(df).max(axis=0)
    assert (np.abs(tot - df).max(axis=0) <= 1e-12 * scale).all()


@pytest.mark.parametrize("ixyz", [1, 2, 3])
def test_rpn3_gravity_enters_momentum_and_energy(ixyz, O):
    """With aux17 = g*dz, the normal-momentum jump loses (rho_l+rho_r)/2*aux17
    and the energy jump (m_l+m_r)/2*aux17 -- on the z sweep only."""
    rng = np.random.default_rng(10 + ixyz)
    q = _pencil(rng)
    N = q.shape[0]
    vg = -9.81 * 1000 + rng.standard_normal(N)
    st, fw, s, am, ap, roe = O.rpn3(ixyz, q, np.full(N, 1.4), vg, 1.0, 1000.0, 0)
    df = (_flux(q, ixyz)[1:] - _flux(q, ixyz)[:-1])
    if ixyz == 3:
        df[:, 3] -= 0.5 * (q[1:, 0] + q[:-1, 0]) * vg[1:]
        df[:, 4] -= 0.5 * (q[1:, 3] + q[:-1, 3]) * vg[1:]
    tot = (am + ap)[1:]
    assert (np.abs(tot - df).max(axis=0) <= 1e-12 * np.abs(df).max(axis=0)).all()


def test_roe_average_property(O):
    """Roe matrix A(u, H) of the ideal gas satisfies A (q_r - q_l) = f_r - f_l
    (LeVeque 2002 eq. 15.41) for the averages the solver stores."""
    rng = np.random.default_rng(5)
    q = _pencil(rng)
    N = q.shape[0]
    st, fw, s, am, ap, roe = O.rpn3(1, q, np.full(N, 1.4), np.zeros(N), 1.0, 1000.0, 0)
    g1 = 0.4
    for i in range(1, N):
        u, v, w, H = roe[1, i], roe[2, i], roe[3, i], roe[4, i]
        q2 = u * u + v * v + w * w
        A = np.array([
            [0, 1, 0, 0, 0],
            [0.5 * g1 * q2 - u * u, (3 - 1.4) * u, -g1 * v, -g1 * w, g1],
            [-u * v, v, u, 0, 0],
            [-u * w, w, 0, u, 0],
            [u * (0.5 * g1 * q2 - H), H - g1 * u * u, -g1 * u * v, -g1 * u * w, 1.4 * u]])
        dq = q[i] - q[i - 1]
        df = _flux(q[i:i + 1], 1)[0] - _flux(q[i - 1:i], 1)[0]
        assert np.abs(A @ dq - df).max() <= 1e-9 * np.abs(df).max()
    # speeds are u - a, u, u + a with a**2 = (gamma-1)(H - q2/2)
    a = np.sqrt(g1 * (roe[4, 1:] - 0.5 * roe[0, 1:]))
    assert np.allclose(s[1:, 0], roe[1, 1:] - a, rtol=1e-13, atol=1e-12)
    assert np.allclose(s[1:, 2], roe[1, 1:] + a, rtol=1e-13, atol=1e-12)


def test_supersonic_interface_has_no_left_fluctuation(O):
    rng = np.random.default_rng(7)
    q = _pencil(rng, speed=0.0)
    q[:, 1] = 2.5 * 330.0 * q[:, 0]           # Mach 2.5 in +x
    q[:, 4] += 0.5 * q[:, 1] ** 2 / q[:, 0]
    N = q.shape[0]
    st, fw, s, am, ap, roe = O.rpn3(1, q, np.full(N, 1.4), np.zeros(N), 1.0, 1000.0, 0)
    assert (s[1:, 0] > 0).all()
    assert np.array_equal(am[1:], np.zeros_like(am[1:]))


def _jacobian(roe_i, d, gamma=1.4):
    """Euler flux Jacobian in direction d (0, 1, 2) at a Roe state (u, v, w, H)."""
    u = np.array([roe_i[1], roe_i[2], roe_i[3]])
    H = roe_i[4]
    g1 = gamma - 1.0
    q2 = (u ** 2).sum()
    A = np.zeros((5, 5))
    un = u[d]
    A[0, 1 + d] = 1.0
    for k in range(3):
        A[1 + k, 0] = -u[k] * un + (0.5 * g1 * q2 if k == d else 0.0)
        A[1 + k, 1 + d] += u[k]
        A[1 + k, 1 + k] += un
        A[1 + k, 1:4] += (-g1 * u if k == d else 0.0)
        if k == d:
            A[1 + k, 4] = g1
    A[4, 0] = un * (0.5 * g1 * q2 - H)
    A[4, 1:4] = -g1 * un * u
    A[4, 1 + d] += H
    A[4, 4] = gamma * un
    return A


@pytest.mark.parametrize("ixyz", [1, 2, 3])
@pytest.mark.parametrize("icoor", [2, 3])
def test_rpt3_splits_transverse_jacobian(ixyz, icoor, O):
    """bmasdq + bpasdq = B(roe) asdq, with B the flux Jacobian in the
    transverse direction of icoor (rpt3euX.f table of directions), and each
    part carries only waves of its speed sign."""
    rng = np.random.default_rng(100 * ixyz + icoor)
    q = _pencil(rng)
    N = q.shape[0]
    st, fw, s, am, ap, roe = O.rpn3(ixyz, q, np.full(N, 1.4), np.zeros(N), 1.0, 1000.0, 0)
    asdq = rng.standard_normal((N, 5))
    bm = np.zeros((N, 5)); bp = np.zeros((N, 5))
    import ctypes as C
    dp = C.POINTER(C.c_double)
    O.lib.orc_rpt3(ixyz, icoor, 5, 2, N - 4, roe.ctypes.data_as(dp), asdq.ctypes.data_as(dp),
                   bm.ctypes.data_as(dp), bp.ctypes.data_as(dp))
    # Roe velocities are stored in the sweep frame (m_n, m_v, m_w)
    perm = {1: [0, 1, 2, 3, 4], 2: [0, 2, 3, 1, 4], 3: [0, 3, 1, 2, 4]}[ixyz]
    # transverse global direction: y-like or z-like of the sweep
    d = {(1, 2): 1, (1, 3): 2, (2, 2): 2, (2, 3): 0, (3, 2): 0, (3, 3): 1}[(ixyz, icoor)]
    for i in range(1, N):
        r = roe[:, i]
        # roe holds (u2v2w2, u_n, u_v, u_w, H, ...) in the local frame
        vel_local = r[1:4]
        vel = np.zeros(3)
        for k in range(3):
            vel[perm[1 + k] - 1] = vel_local[k]
        B = _jacobian(np.array([0, vel[0], vel[1], vel[2], r[4]]), d)
        y = B @ asdq[i]                        # pencil arrays use global components
        got = (bm + bp)[i]
        assert np.abs(got - y).max() <= 1e-10 * max(np.abs(y).max(), 1.0)


def test_rptt3_equals_rpt3_in_y_like_direction(O):
    import ctypes as C
    dp = C.POINTER(C.c_double)
    rng = np.random.default_rng(3)
    q = _pencil(rng, meqn=6)
    N = q.shape[0]
    st, fw, s, am, ap, roe = O.rpn3(2, q, np.full(N, 1.4), np.zeros(N), 1.0, 1000.0, 1)
    x = rng.standard_normal((N, 6))
    outs = []
    for fn in (O.lib.orc_rpt3, O.lib.orc_rptt3):
        bm = np.zeros((N, 6)); bp = np.zeros((N, 6))
        fn(2, 2, 6, 2, N - 4, roe.ctypes.data_as(dp), x.ctypes.data_as(dp), bm.ctypes.data_as(dp),
           bp.ctypes.data_as(dp))
        outs.append((bm, bp))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("cfl", [0.05, 0.3, 0.7, 0.95])
def test_limiter_consistency_and_tvd(cfl, O):
    """phi(1) = 1 (second order on smooth data) and phi(r <= 0) = 0 (TVD) for
    both limiters at CFL = CFLup; CFL-Superbee stays below its bounds."""
    beta = 0.666666666666666
    assert abs(O.lib.orc_suppow(1.0, cfl, cfl) - 1.0) < 1e-15
    assert abs(O.lib.orc_cflsupbee(1.0, beta, 1.0, cfl, cfl) - 1.0) < 1e-15
    for r in (-3.0, -1.0, -1e-9, 0.0):
        assert O.lib.orc_suppow(r, cfl, cfl) == 0.0
        assert O.lib.orc_cflsupbee(1.0, beta, r, cfl, cfl) == 0.0
    for r in np.linspace(0.01, 10, 50):
        phi = O.lib.orc_cflsupbee(1.0, beta, r, cfl, cfl)
        assert phi <= 2.0 / (1.0 - cfl) + 1e-12
        assert phi <= 2.0 * r / cfl + 1e-12
        assert 0.0 <= O.lib.orc_suppow(r, cfl, cfl) <= 2.0 / (1.0 - cfl) + 1e-12


def test_hydrostatic_rest_is_preserved(O):
    """The well-balanced f-wave form keeps an atmosphere at rest at rest: one
    step3 of the isothermal column changes q only at roundoff level."""
    c, p, g, aux, q = P.config_plain(mx=6, my=5, mz=24, noise=0.0)
    O.bc3(g, q, aux, c.mthbc)
    qn, cfl, st = O.step3(g, q, aux, 5.0, p.nspecadv)
    assert st == 0
    it = (slice(2, -2),) * 3
    drho = np.abs(qn[it][..., 0] - q[it][..., 0]).max() / q[..., 0].max()
    dmom = np.abs(qn[it][..., 1:4]).max() / (q[..., 0].max() * 300.0)
    assert drho < 1e-12 and dmom < 1e-12


def test_conduction_fourier_mode_decay(O):
    """Temperature perturbation sin(2 pi i/mx) on a uniform background with one
    conduction substep, lam = dtrc*kappa/dx**2 per direction (conduct.f:262-323).
    The x pass multiplies the mode by fac = 1 - 4 lam sin^2(pi/mx).  The y and
    z passes only see the mode through ghost cells, which conduct copies once
    and never refreshes between passes (conduct.f:251-256): where a pass reads
    a stale (undecayed) ghost, the value f*Tp becomes (f + lam (1-f))*Tp.
    That happens on the y-boundary rows (j = 1, my) and, because the z-ghost
    refresh only runs i up to mz+mbc (conduct.f:86-91), on the z-boundary
    layers of the columns i > mz+mbc.  With mx > mz this pins both rules."""
    mx, my, mz = 16, 4, 4
    c = P.ClawData(mx=mx, my=my, mz=mz, xupper=16e3, yupper=4e3, zupper=4e3,
                   mthbc=[2, 2, 2, 2, 1, 1])
    g = P.Grid.from_claw(c)
    aux = np.zeros(g.shape_aux)
    rho0, T0, R, gam = 1.0, 250.0, 287.0, 1.4
    aux[..., 0] = rho0
    aux[..., 4] = rho0 * R * T0 / (gam - 1)
    aux[..., 17] = R
    aux[..., 18] = gam
    kap = 2000.0
    aux[..., 10] = 1e-9    # ndtr = 0 (no viscosity) would skip conduction too (src3.f90:317, 341)
    aux[..., 11] = kap
    q = np.array(aux[..., :6])
    A = 0.5
    ii = np.arange(g.shape_q[2]) + 1 - g.mbc
    Tp = A * np.sin(2 * math.pi * ii / mx)
    q[..., 4] = rho0 * R * (T0 + Tp)[None, None, :] / (gam - 1)
    p = P.ProbData(forcemth=0.0)
    dt = 50.0
    st, ndtr, ndtrc = O.src3(g, p, q, aux, 0.0, dt, c.mthbc)
    assert st == 0 and ndtr == 1 and ndtrc == 1
    lam = dt * kap / g.dx ** 2
    fac = 1 - 4 * lam * math.sin(math.pi / mx) ** 2
    T = (gam - 1) * q[..., 4] / (q[..., 0] * R) - T0
    for k in range(1, mz + 1):
        for j in range(1, my + 1):
            for i in range(1, mx + 1):
                f = fac
                if j in (1, my):
                    f = f + lam * (1 - f)
                if k in (1, mz) and i > mz + g.mbc:
                    f = f + lam * (1 - f)
                got = T[k + 1, j + 1, i + 1]
                want = f * A * math.sin(2 * math.pi * i / mx)
                assert abs(got - want) <= 1e-9 * A, (i, j, k, got, want)


def test_claw3_dt_recursion(O):
    """Accepted steps choose the next dt = min(dtv(2), dt*cflv(2)/cfl) and a
    step with cfl > cflv(1) is redone with that dt (claw3_mpi.f:861-919)."""
    c, p, g, aux, q = perturbed(10, 8, 8, [2, 2, 2, 2, 3, 1], seed=1, vel=0.05)
    st, t, dtv, cflv, nv, log = O.claw3(g, p, q, aux, 0.0, -1.0, [50.0, 10.0, 0, 0, 0], c.cflv,
                                        [100, 0], c.method, c.mthbc)
    assert st == 0
    assert log[0][2] is False and log[-1][2] is True
    for (dt, cfl, acc), nxt in zip(log, log[1:]):
        assert nxt[0] == min(10.0, dt * 0.9 / cfl)
    dt, cfl, _ = log[-1]
    assert dtv[4] == min(10.0, dt * 0.9 / cfl)
    assert t == log[-1][0]


def test_plume_window_radius(O):
    """After 10 claw3 steps of the plume workload, a 40-cell window and a
    64-cell window around the plume agree exactly on every cell at least
    WINDOW_MARGIN - 2 cells inside the smaller one's edges: the radius the
    full-domain GPU test against an oracle window relies on.  (The 60 km top
    keeps this quick; with the full 300 km column the two runs differ by
    1.4e-10 of the field at the window's edge and by nothing from 10 cells
    in.)"""
    mz, nsteps = 60, 10
    runs = []
    for nw in (40, 64):
        c, p, g, aux, q = plume_window(125 - nw // 2, 100 - nw // 2, nw, nw, mz)
        dtv = list(c.dtv)
        t = 0.0
        for _ in range(nsteps):
            st, t, dtv, _, _, _ = O.claw3(g, p, q, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                          c.method, c.mthbc)
            assert st == 0
            dtv[0] = dtv[4]
        runs.append((t, q[2:2 + mz, 2:2 + nw, 2:2 + nw]))
    (ta, a), (tb, b) = runs
    assert ta == tb
    b = b[:, 12:52, 12:52]
    diff = np.abs(a - b).max(axis=(0, 3))
    r = WINDOW_MARGIN - 2
    assert diff[r:40 - r, r:40 - r].max() == 0.0
    assert np.abs(a[..., 3] / a[..., 0]).max() > 0.1
