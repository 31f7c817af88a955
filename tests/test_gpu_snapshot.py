"""Asynchronous snapshots (magic_snapshot_begin / magic_snapshot_wait).

The snapshot equals magic_get_q's download of the state it was queued on,
although steps queued after it run before it is waited for, in both forms;
its handle stays valid after the context is destroyed; the batch's
snapshot of a scenario equals that scenario's get_q."""
import numpy as np
import pytest

from tests.helpers import perturbed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_snapshot_is_the_state_it_was_queued_on(form):
    from magic_master_amd.hip import FORM_FP32P, FORM_FP64, MagicHIP, PinnedBuffer, snapshot_wait
    c, p, g, aux, q = perturbed(12, 10, 8, [1, 1, 1, 1, 3, 1], seed=3, vel=0.05)
    d = MagicHIP(g, p, aux, q, form=FORM_FP32P if form == "fp32" else FORM_FP64)
    buf = PinnedBuffer(8 * int(np.prod(g.shape_q)))
    try:
        t, dtv, _ = d.run_steps(c, 2)
        ref = d.get_q()
        h = d.snapshot_begin(buf)
        t, dtv, _ = d.run_steps(c, 3, t0=t, dt0=dtv[0])   # queued after the snapshot
        later = d.get_q()
        snapshot_wait(h)
        np.testing.assert_array_equal(buf.array(g.shape_q), ref)
        assert not np.array_equal(later, ref)
        h2 = d.snapshot_begin(buf)
    finally:
        d.close()
    snapshot_wait(h2)   # the context is gone; its last copy finished before
    np.testing.assert_array_equal(buf.array(g.shape_q), later)
    buf.close()


def test_batch_snapshot_equals_get_q():
    from magic_master_amd.hip import MagicBatch, PinnedBuffer, snapshot_wait
    c, p, g, aux, q = perturbed(12, 10, 8, [1, 1, 1, 1, 3, 1], seed=4, vel=0.05)
    b = MagicBatch([g], [p], [aux[:, 0, 0, :].copy()], [q])
    buf = PinnedBuffer(8 * int(np.prod(g.shape_q)))
    try:
        b.start(0, 0.0, 1e300, 1.0, [1.0, c.dtv[1], 0, 0, 0], c.cflv, 3, c.mthbc)
        b.run(max_attempts=100, check_every=2)
        ref = b.get_q(0)
        h = b.snapshot_begin(0, buf)
        snapshot_wait(h)
        np.testing.assert_array_equal(buf.array(g.shape_q), ref)
    finally:
        b.close()
        buf.close()
