"""The serving loop on the GPU (continuous batching of scenario requests).

Small requests arrive while others run; every request completes, its
snapshots are the reference's HDF5 frames at every cadence multiple up to its
end time (the last one exactly at the end time), the scenarios running at
once never exceed the memory budget, and each request's final state in
the mix equals the same request served alone, bit for bit."""
import numpy as np
import pytest

from magic_master_amd import h5, serve

pytestmark = pytest.mark.gpu


def _requests():
    rq = serve.poisson_requests(40.0, 0.2, seed=11, t_end=(20.0, 45.0), cadence=10.0,
                                nmin=(8, 8, 12), nmax=(16, 14, 20))
    assert len(rq) >= 3
    return rq


def test_serving_loop_completes_and_snapshots(tmp_path):
    rq = _requests()
    print(f"[serve] {len(rq)} requests", flush=True)
    budget = 2.5 * max(serve.footprint_bytes(r.grid) for r in rq)
    res = serve.Server(budget_bytes=budget, check_every=2, snapshots="hdf5", outdir=str(tmp_path),
                       keep_final=True).run(rq)
    print({k: v for k, v in res.items() if k != "outcomes"}, flush=True)
    assert res["max_footprint_bytes"] <= budget
    assert res["max_concurrent"] >= 2            # requests overlapped on the device
    for r, o in zip(sorted(rq, key=lambda r: r.arrival), res["outcomes"]):
        assert o.status == 0 and o.completed >= o.admitted >= o.arrival
        times = [t for t, _ in o.snapshots]
        expect = list(np.arange(10.0, r.t_end, 10.0)) + [r.t_end]
        assert np.allclose(times, expect, rtol=0, atol=1e-9), (times, expect)
        for t, path in o.snapshots:
            f = h5.in3(path)
            assert f["params"][2] == t and f["dims"] == (r.grid.mz, r.grid.my, r.grid.mx, r.grid.meqn)
        m = r.grid.mbc
        np.testing.assert_array_equal(f["data"], o.final_q[m:m + r.grid.mz, m:m + r.grid.my, m:m + r.grid.mx])


def test_each_result_alone_equals_in_the_mix():
    rq = _requests()
    mix = serve.Server(snapshots="none", keep_final=True).run(rq)["outcomes"]
    for r, o in zip(sorted(rq, key=lambda r: r.arrival), mix):
        r.arrival = 0.0
        alone = serve.Server(snapshots="none", keep_final=True).run([r])["outcomes"][0]
        assert alone.steps == o.steps
        np.testing.assert_array_equal(alone.final_q, o.final_q)


def test_failed_request_leaves_the_others_running():
    """One request starts with a negative density in one cell: its first
    Riemann problem there fails (MAGIC_ERHO, rpn3euX.f:140-153), it leaves the
    batch with that status, and every other request completes as it does
    without it."""
    from magic_master_amd.hip import MAGIC_ERHO
    rq = _requests()
    ref = serve.Server(snapshots="none", keep_final=True).run(_requests())["outcomes"]
    bad = sorted(rq, key=lambda r: r.arrival)[1]
    g, m = bad.grid, bad.grid.mbc
    q = np.zeros(g.shape_q)
    q[..., 0] = bad.column[:, None, None, 0]
    q[..., 4] = bad.column[:, None, None, 4]
    q[m + g.mz // 2, m + g.my // 2, m + g.mx // 2, 0] = -1.0
    bad.q0 = q
    print(f"[serve-failure] {len(rq)} requests, request {bad.rid} fails", flush=True)
    res = serve.Server(snapshots="none", keep_final=True).run(rq)
    for o, o_ref in zip(res["outcomes"], ref):
        if o.rid == bad.rid:
            assert o.status == MAGIC_ERHO and o.final_q is None
        else:
            assert o.status == 0 and o.steps == o_ref.steps
            np.testing.assert_array_equal(o.final_q, o_ref.final_q)
