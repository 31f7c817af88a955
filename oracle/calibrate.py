"""CPU speed of the compiled reference against the C restatement (oracle),
one host core each, on the same workload: a 250 x 16 x 300 slab of the
full-domain FP64 workload (the slab one worker of the benchmark's CPU
baseline advances), a few claw3 steps after one untimed step.  The ratio
relates the benchmark's CPU baseline (the restatement on every host core of
the GPU machine) to what the reference itself would do there.

Test infrastructure: it needs oracle/_ref/libmagic_ref.so, which is built only
where the reference's sources are mounted (python -m magic_master_amd.build).

    python oracle/calibrate.py [--steps 5]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[0] = os.path.dirname(HERE)   # the repository root, not oracle/

from magic_master_amd import problem as P  # noqa: E402
from oracle.oracle import COracle, Reference  # noqa: E402


def _run(step, c, nsteps):
    """One untimed claw3 step, then nsteps timed ones (claw3ez frame loop,
    one accepted step per call); returns seconds of the timed steps."""
    dtv = list(c.dtv)
    t = 0.0
    t0 = None
    for n in range(nsteps + 1):
        if n == 1:
            t0 = time.perf_counter()
        t, dtv = step(t, dtv)
        dtv[0] = dtv[4]
    return time.perf_counter() - t0, t


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=5)
    a = ap.parse_args(argv)
    if not Reference.available():
        sys.exit("oracle/_ref/libmagic_ref.so is not built")
    c, p, g, aux, q0 = P.config_full(mx=250, my=16, mz=300)
    cells = g.mx * g.my * g.mz
    O = COracle()
    R = Reference.get()
    R.setup(p, c.mthbc)

    qo = q0.copy()

    def step_o(t, dtv):
        st, t, dtv, _, _, _ = O.claw3(g, p, qo, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0], c.method, c.mthbc)
        if st != 0:
            raise RuntimeError(f"oracle status {st}")
        return t, dtv

    qr = q0.copy()

    def step_r(t, dtv):
        info, t, dtv, _, _ = R.claw3(g, qr, aux, t, t - 1.0, dtv, c.cflv, [c.nv[0], 0], c.method, c.mthbc)
        if info != 0:
            raise RuntimeError(f"reference info {info}")
        return t, dtv

    so, to = _run(step_o, c, a.steps)
    sr, tr = _run(step_r, c, a.steps)
    if to != tr or (qo != qr).any():
        raise RuntimeError("the restatement and the reference disagree on this workload")
    ro, rr = cells * a.steps / so / 1e6, cells * a.steps / sr / 1e6
    print(f"workload: {g.mx}x{g.my}x{g.mz} slab of the full-domain FP64 workload, {a.steps} steps, one core")
    print(f"restatement (oracle/magic_oracle.c, gcc -O2): {ro:.3f} Mcell-updates/s ({so / a.steps:.3f} s/step)")
    print(f"reference (flang -O2, oracle/_ref):           {rr:.3f} Mcell-updates/s ({sr / a.steps:.3f} s/step)")
    print(f"reference / restatement speed: {rr / ro:.3f}")


if __name__ == "__main__":
    main()
