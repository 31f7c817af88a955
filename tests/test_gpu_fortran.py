"""The Fortran host on the GPU: (1) a caller that uses the drop-in claw3 with
the reference's argument list (magic_master_amd/bin/dropin_check), (1b) a
caller that keeps its own step control and calls the drop-in step3, bc3,
b4step3 and src3 (magic_master_amd/bin/step3_loop_check), and (2)
the claw3ez_hip driver reading claw3ez.data/setprob.data and writing the
reference's HDF5 frames; both against the C oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from magic_master_amd import problem as P
from oracle.oracle import COracle
from tests.helpers import WIND_CASES, interior, perturbed, rel_err

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "magic_master_amd", "bin")


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    print(r.stdout[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


def test_dropin_claw3_matches_oracle(tmp_path):
    (mx, my, mz), mthbc = (14, 10, 9), [2, 2, 1, 1, 3, 1]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=5, vel=0.05)
    p.write(tmp_path / "setprob.data")
    with open(tmp_path / "dims.txt", "w") as f:
        f.write(f"{mx} {my} {mz} {g.meqn} {g.mbc} {g.maux}\n")
        f.write(f"{g.xlower!r} {g.ylower!r} {g.zlower!r} {g.dx!r} {g.dy!r} {g.dz!r}\n")
        f.write("4\n" + " ".join(map(str, mthbc)) + "\n" + str(tmp_path / "setprob.data") + "\n")
    aux.tofile(tmp_path / "aux.bin")
    q.tofile(tmp_path / "q.bin")
    _run([os.path.join(BIN, "dropin_check"), "dims.txt", "aux.bin", "q.bin", "qout.bin"], tmp_path)
    qg = np.fromfile(tmp_path / "qout.bin").reshape(g.shape_q)
    O = COracle()
    qo = q.copy()
    dtv = [1.0, 10.0, 0, 0, 0]
    t = 0.0
    for _ in range(4):
        st, t, dtv, _, _, _ = O.claw3(g, p, qo, aux, t, t - 1.0, dtv, [0.95, 0.9, 0, 0], [10000, 0],
                                      [1, 2, 22, 0, 1, 0, g.maux], mthbc)
        assert st == 0
        dtv[0] = dtv[4]
    assert rel_err(qg, qo, interior(g)) <= 1e-10


def test_dropin_step3_loop_matches_oracle(tmp_path):
    """A caller that keeps its own claw3 step control and calls the drop-in
    step3, bc3, b4step3 and src3 with the reference's argument lists
    (magic_master_amd/bin/step3_loop_check): the same attempted steps (a
    rejected first one included) and the same state as the oracle's claw3."""
    (mx, my, mz), mthbc = (12, 10, 9), [2, 2, 1, 1, 3, 1]
    c, p, g, aux, q = perturbed(mx, my, mz, mthbc, seed=8, vel=0.05)
    p.write(tmp_path / "setprob.data")
    with open(tmp_path / "dims.txt", "w") as f:
        f.write(f"{mx} {my} {mz} {g.meqn} {g.mbc} {g.maux}\n")
        f.write(f"{g.xlower!r} {g.ylower!r} {g.zlower!r} {g.dx!r} {g.dy!r} {g.dz!r}\n")
        f.write("4 50.0\n" + " ".join(map(str, mthbc)) + "\n" + str(tmp_path / "setprob.data") + "\n")
    aux.tofile(tmp_path / "aux.bin")
    q.tofile(tmp_path / "q.bin")
    out = _run([os.path.join(BIN, "step3_loop_check"), "dims.txt", "aux.bin", "q.bin", "qout.bin"], tmp_path)
    attempts = [ln.split() for ln in out.splitlines() if len(ln.split()) == 4 and ln.split()[3] in ("T", "F")]
    assert "SRC3...  Vis. Steps" in out
    qg = np.fromfile(tmp_path / "qout.bin").reshape(g.shape_q)
    O = COracle()
    qo = q.copy()
    dtv = [50.0, 10.0, 0, 0, 0]
    t = 0.0
    log = []
    for _ in range(4):
        st, t, dtv, _, _, lg = O.claw3(g, p, qo, aux, t, t - 1.0, dtv, [0.95, 0.9, 0, 0], [10000, 0],
                                       [1, 2, 22, 0, 1, 0, g.maux], mthbc)
        assert st == 0
        log += lg
        dtv[0] = dtv[4]
    assert len(attempts) == len(log) and not log[0][2]     # the first attempt is rejected
    for a, (dt_o, cfl_o, acc_o) in zip(attempts, log):
        assert abs(float(a[1]) - dt_o) <= 1e-12 * dt_o and (a[3] == "T") == acc_o
        assert abs(float(a[2]) - cfl_o) <= 1e-12 * cfl_o
    assert rel_err(qg, qo, interior(g)) <= 1e-10


def test_claw3ez_driver_matches_oracle(tmp_path):
    c, p, g, aux, q = P.config_plain(mx=16, my=12, mz=10, noise=0.0)
    c.nout = c.nstop = c.nstepout = 3
    c.write(tmp_path / "claw3ez.data")
    p.write(tmp_path / "setprob.data")
    out = _run([os.path.join(BIN, "claw3ez_hip")], tmp_path)
    assert "Frame    1" in out
    from magic_master_amd import h5
    f = h5.in3(str(tmp_path / h5.frame_name(1)))
    t = float(f["params"][2])
    assert f["dims"] == (g.mz, g.my, g.mx, g.meqn)
    qg = f["data"]
    O = COracle()
    qo = q.copy()
    dtv = list(c.dtv)
    to = 0.0
    for _ in range(3):
        st, to, dtv, _, _, _ = O.claw3(g, p, qo, aux, to, to - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                       c.method, c.mthbc)
        assert st == 0
        dtv[0] = dtv[4]
    assert abs(t - to) <= 1e-12 * to
    assert rel_err(qg, qo[interior(g)]) <= 1e-10


@pytest.mark.parametrize("wind", [0, 3])
def test_claw3ez_driver_profile_background(tmp_path, wind):
    """loadprof = 1: the driver builds the background from profile.data (the
    reference's own file) as setaux.f does, without and with the analytic x
    wind (amp0 and amp1 with windscale1 = 1); 10 components as in the
    reference's deck.  Against the oracle started from problem.setaux_profile
    (itself equal bit for bit to the compiled reference's setaux)."""
    import shutil
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks", "profile.data")
    shutil.copy(src, tmp_path / "profile.data")
    c, p, g, aux, q = P.config_plain(mx=12, my=10, mz=20, noise=0.0, meqn=10)
    c.zlower, c.zupper = 0.0, 20e3
    c.nout = c.nstop = c.nstepout = 3
    p.loadprof, p.nspecadv = 1, 5
    for k, v in WIND_CASES[wind].items():
        setattr(p, k, v)
    print(f"[claw3ez profile] wind case {wind}: {WIND_CASES[wind]}", flush=True)
    g = P.Grid.from_claw(c)
    aux = P.setaux_profile(g, p, P.read_profile(src))
    q = P.qinit(g, aux)
    c.write(tmp_path / "claw3ez.data")
    p.write(tmp_path / "setprob.data")
    out = _run([os.path.join(BIN, "claw3ez_hip")], tmp_path)
    assert "Frame    1" in out
    from magic_master_amd import h5
    f = h5.in3(str(tmp_path / h5.frame_name(1)))
    t = float(f["params"][2])
    assert f["dims"] == (g.mz, g.my, g.mx, g.meqn)
    O = COracle()
    qo = q.copy()
    dtv = list(c.dtv)
    to = 0.0
    for _ in range(3):
        st, to, dtv, _, _, _ = O.claw3(g, p, qo, aux, to, to - 1.0, dtv, c.cflv, [c.nv[0], 0],
                                       c.method, c.mthbc)
        assert st == 0
        dtv[0] = dtv[4]
    assert abs(t - to) <= 1e-12 * to
    assert rel_err(f["data"], qo[interior(g)]) <= 1e-10
