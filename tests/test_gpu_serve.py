"""Serving loop on the device: requests admitted under a memory budget that
forces queueing, batched while they overlap, and each result bitwise equal
to the same request solved alone."""
import pytest

import conftest
import gamera_op as go
import serve


@pytest.mark.gpu
def test_served_results_equal_alone():
    reqs = serve.make_requests(10, rate=200.0, seed=5, steps=(6, 30),
                               cart_sizes=[(16, 16, 16), (24, 24, 16), (32, 16, 16)],
                               sph_sizes=[(16, 12, 16), (24, 12, 16)])
    budget = 3 * serve.device_bytes_estimate((32, 24, 16), 0)  # about three at a time
    print("serving", [(r.rid, r.kind, r.shape, r.steps, r.precision) for r in reqs], flush=True)
    res, wall = serve.serve(reqs, budget_bytes=budget, quantum=7, tune=True, digests=True)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    alone, _ = serve.serve_plain(reqs, digests=True, realtime=False)
    da = {r.rid: r.digest for r in alone}
    ta = {r.rid: r.time_sim for r in alone}
    for r in res:
        assert r.digest == da[r.rid], r.rid
        assert r.time_sim == ta[r.rid], r.rid
    print(serve.summarize(res, wall), flush=True)


@pytest.mark.gpu
@pytest.mark.skipif(conftest.PREPARER is None, reason="needs the preparation pool (-m gpu)")
def test_served_with_preparation_pool_equal_alone():
    reqs = serve.make_requests(12, rate=500.0, seed=8, steps=(5, 25),
                               cart_sizes=[(16, 16, 16), (24, 24, 16)],
                               sph_sizes=[(16, 12, 16)])
    budget = 4 * serve.device_bytes_estimate((24, 24, 16), 0)
    res, wall = serve.serve(reqs, budget_bytes=budget, quantum=6, digests=True,
                            preparer=conftest.PREPARER)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    alone, _ = serve.serve_plain(reqs, digests=True, realtime=False)
    da = {r.rid: r.digest for r in alone}
    for r in res:
        assert r.digest == da[r.rid], r.rid
    print(serve.summarize(res, wall), flush=True)
