"""Failure paths and run control on the device.

* Positivity floors and gravity: the floor box (non-default floors set by
  the problem, a uniform gravity, a state that crosses both floors) is
  bitwise the oracle's, which is bitwise the compiled reference's
  (test_oracle.py).
* The entropy channel with p_S != 0: density, velocity, pressure and face B
  are bitwise the oracle's (p_S feeds nothing else); p_S itself goes
  through the device's pow, which is not glibc's, so it is held to relative
  L1 <= 1e-12 and its bitwise mismatches are reported.
* A non-finite state: gk_step returns GK_ERR_NAN, solve() returns -1, and in
  a batch only the member that holds it is flagged.
* Run control: gk_step_until ends the run where the reference's loop breaks
  (time > time_stop) or where a dump falls due (time >= t_event); batch
  members with step budgets stop on the device while the batch goes on.
"""
import numpy as np
import pytest

import gamera_op as go
from parity import (active_slices, compare_states, exact_mismatches, oracle_from_problem,
                    run_oracle)


def same_state(a, b):
    pa, ba = a.state()
    pb, bb = b.state()
    return all(np.array_equal(x, y) for x, y in zip(pa + ba, pb + bb))


@pytest.mark.gpu
@pytest.mark.parametrize("seed,ni,nj,nk,steps", [(31, 24, 20, 12, 20), (33, 32, 24, 16, 25)])
def test_floor_box_matches_oracle(seed, ni, nj, nk, steps):
    print(f"floor box seed={seed} {ni}x{nj}x{nk} steps={steps}", flush=True)
    P = go.Problem(go.FLOOR_BOX, seed, ni, nj, nk, full_b0=True)
    try:
        O = oracle_from_problem(P)
        O.prologue()
        cell, _ = active_slices(P)
        hits = 0
        for _ in range(steps):
            O.step(1)
            hits += int((O.gas[0][cell] == O.scalar("rho_floor")).sum())
        assert hits > 0  # the floor branch is taken
        P.device_open(0)
        P.device_step(steps)
        P.device_fetch()
        nbad = exact_mismatches(P, O)
        sc = P.device_scalars()
        print("floored cell-steps", hits, "bitwise mismatches", nbad, "divB", sc["divB"], flush=True)
        assert nbad == 0
        assert sc["time"] == O.scalar("time_sim") and sc["dt"] == O.scalar("dt")
        assert sc["nonfinite"] == 0.0
    finally:
        P.close()


@pytest.mark.gpu
def test_entropy_channel():
    kind, seed, ni, nj, nk, steps = go.ENTROPY_BOX, 32, 16, 16, 8, 10
    P = go.Problem(kind, seed, ni, nj, nk, full_b0=True)
    try:
        O = run_oracle(P, steps)
        P.device_open(0)
        P.device_step(steps)
        P.device_fetch()
        cell, faces = active_slices(P)
        for f in range(5):
            assert np.array_equal(P.gas(f)[cell], O.gas[f][cell]), f
        for f in range(3):
            assert np.array_equal(P.gem(f)[faces[f]], O.gem[f][faces[f]]), f
        ps_d, ps_o = P.gas(5)[cell], O.gas[5][cell]
        assert np.abs(ps_o).min() > 0.0
        rel = float(np.abs(ps_d - ps_o).sum() / np.abs(ps_o).sum())
        print("p_S rel L1", rel, "bitwise mismatches", int((ps_d != ps_o).sum()), "of", ps_o.size,
              flush=True)
        assert rel <= 1e-12
    finally:
        P.close()
    # FP32-bulk form: the entropy channel too within the form's bar
    P = go.Problem(kind, seed, 24, 24, 16, full_b0=True)
    try:
        O = run_oracle(P, 20)
        P.device_open(1)
        P.device_step(20)
        P.device_fetch()
        rel = compare_states(P, O)
        print({k: f"{v:.2e}" for k, v in rel.items()}, flush=True)
        for k in ("rho", "p", "eng", "p_S"):
            assert rel[k] <= 1e-5, (k, rel[k])
    finally:
        P.close()


def poisoned(kind, seed, n, precision=0):
    P = go.Problem(kind, seed, *n)
    P.gas(0)[4, 5, 6] = np.nan
    S = go.Solver(P, precision)
    P.close()
    return S


@pytest.mark.gpu
def test_nonfinite_state_is_reported():
    S = poisoned(go.SINUSOID, 3, (16, 16, 8))
    try:
        with pytest.raises(go.GkNonFinite):
            S.step(2)
        assert S.scalars()["nonfinite"] == 1.0
    finally:
        S.close()
    # the drop-in solve() returns -1
    P = go.Problem(go.SINUSOID, 3, 16, 16, 8)
    try:
        P.gas(0)[4, 5, 6] = np.nan
        with pytest.raises(go.GkError):
            P.solve(3)
    finally:
        P.close()


@pytest.mark.gpu
def test_nonfinite_member_flagged_alone():
    specs = [(go.SINUSOID, 4, (16, 16, 8), 0), (go.SINUSOID_BLAST, 5, (16, 12, 16), 1)]
    alone, members = [], []
    for kind, seed, n, prec in specs:
        P = go.Problem(kind, seed, *n)
        alone.append(go.Solver(P, prec))
        members.append(go.Solver(P, prec))
        P.close()
    bad = poisoned(go.SINUSOID, 6, (16, 16, 8))
    B = go.Batch()
    try:
        for S in members + [bad]:
            B.add(S)
        B.step_async(9)
        flagged = B.sync_status()
        assert flagged == [bad]
        B.remove(bad)
        B.step(3)
        for a, m in zip(alone, members):
            a.step(12)
            assert same_state(a, m)
    finally:
        B.close()
        for S in alone + members + [bad]:
            S.close()


@pytest.mark.gpu
def test_step_until_stops_where_the_reference_loop_breaks():
    P = go.Problem(go.SINUSOID_BLAST, 7, 16, 16, 16, full_b0=True)
    try:
        O = oracle_from_problem(P)
        O.prologue()
        times = []
        for _ in range(12):
            O.step(1)
            times.append(O.scalar("time_sim"))
        S = go.Solver(P, 0)
        T = go.Solver(P, 0)
        try:
            # time_stop between steps 7 and 8: the loop breaks after the first
            # step whose time exceeds it, step 8
            t_stop = 0.5 * (times[6] + times[7])
            assert S.step_until(50, t_stop=t_stop) == 8
            assert S.scalars()["time"] == times[7]
            # an event at exactly the time of step 10 ends the next run there
            assert S.step_until(50, t_event=times[9]) == 2
            # nothing due: all requested steps
            assert S.step_until(2) == 2
            T.step(12)
            assert same_state(S, T)
            assert T.scalars()["time"] == times[11]
        finally:
            S.close()
            T.close()
    finally:
        P.close()


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_budgets(graphs):
    specs = [(go.SINUSOID, 8, (16, 16, 8), 0, 5), (go.DIPOLE_SHELL, 9, (16, 12, 16), 1, 12),
             (go.SINUSOID_SPHERE, 10, (16, 8, 8), 0, None), (go.SINUSOID, 11, (8, 8, 8), 1, 0)]
    alone, members = [], []
    for kind, seed, n, prec, _ in specs:
        P = go.Problem(kind, seed, *n)
        alone.append(go.Solver(P, prec))
        members.append(go.Solver(P, prec))
        P.close()
    B = go.Batch()
    B.set_graphs(graphs)
    try:
        for S, spec in zip(members, specs):
            B.add(S)
            if spec[4] is not None:
                B.set_budget(S, spec[4])
        B.step(8)
        assert [B.budget(S) for S in members] == [0, 4, -1, 0]
        B.step(8)
        assert [B.budget(S) for S in members] == [0, 0, -1, 0]
        for a, m, spec in zip(alone, members, specs):
            n = 16 if spec[4] is None else spec[4]
            a.step(n)
            assert a.scalars()["time"] == m.scalars()["time"], spec
            assert same_state(a, m), spec
    finally:
        B.close()
        for S in alone + members:
            S.close()
