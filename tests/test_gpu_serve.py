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


@pytest.mark.gpu
def test_zero_step_and_nonfinite_requests():
    """A request with no steps completes with its initial state; a request
    whose state turns non-finite is evicted alone and the others equal their
    solo runs."""
    reqs = serve.make_requests(8, rate=400.0, seed=11, steps=(5, 20),
                               cart_sizes=[(16, 16, 16), (24, 16, 16)],
                               sph_sizes=[(16, 12, 16)])
    reqs[2].steps = 0
    reqs[5].poison = True
    res, wall = serve.serve(reqs, budget_bytes=64 << 30, quantum=6, digests=True)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    alone, _ = serve.serve_plain(reqs, digests=True, realtime=False)
    da = {r.rid: r for r in alone}
    for r in res:
        assert r.status == ("nonfinite" if r.rid == 5 else "ok"), (r.rid, r.status)
        assert da[r.rid].status == r.status
        if r.rid != 5:
            assert r.digest == da[r.rid].digest, r.rid
            assert r.time_sim == da[r.rid].time_sim
    assert [r.time_sim for r in res if r.rid == 2] == [0.0]


@pytest.mark.gpu
@pytest.mark.skipif(conftest.PREPARER is None, reason="needs the preparation pool (-m gpu)")
@pytest.mark.timeout(600)
def test_served_stream_of_the_configured_mix():
    """A scaled stream of the configured serving mix: 16 requests (not
    2000) with its size ranges (Cartesian 32^3 ... 128^3, spherical shell
    64x32x64 ... 128x64x128), 70/30 split, half FP32-bulk, steps U(50, 500),
    admitted under a 6 GiB budget that forces queueing.  Each result equals
    the same request solved alone; the 32^3 FP64 requests are bitwise the
    strict oracle's."""
    from parity import exact_mismatches, run_oracle

    reqs = serve.make_requests(16, rate=50.0, seed=3)
    print("stream", [(r.rid, r.kind, r.shape, r.steps, r.precision) for r in reqs], flush=True)
    res, wall = serve.serve(reqs, budget_bytes=6 << 30, digests=True, preparer=conftest.PREPARER)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    print(serve.summarize(res, wall), flush=True)
    alone, _ = serve.serve_plain(reqs, digests=True, realtime=False)
    da = {r.rid: r for r in alone}
    for r in res:
        assert r.status == "ok" and r.digest == da[r.rid].digest, r.rid
        assert r.time_sim == da[r.rid].time_sim, r.rid
    small = [r for r in reqs if r.shape == (32, 32, 32) and r.precision == 0]
    assert small, "the draw has no 32^3 FP64 request"
    for r in small[:2]:
        P = go.Problem(r.kind, r.seed, *r.shape, full_b0=True)
        try:
            O = run_oracle(P, r.steps)
            S = go.Solver(P, 0)
            S.step(r.steps)
            prim, bf = S.state()
            S.close()
            assert serve.state_digest_of(prim, bf) == da[r.rid].digest
            for f in range(6):
                P.gas(f)[...] = prim[f]
            for f in range(3):
                P.gem(f)[...] = bf[f]
            assert exact_mismatches(P, O) == 0, r.rid
        finally:
            P.close()
