"""Batched stepping and the launch-shape autotuner on the device.

A grid stepped inside a batch (one launch per stage for all grids, HIP
graphs) must give bitwise the results it gets stepped alone, whatever else is
in the batch, whether graphs are used, and when grids join or leave between
steps.  Autotuning must not change results either.
"""
import numpy as np
import pytest

import gamera_op as go
from parity import exact_mismatches, run_oracle

MIX = [
    # kind, seed, ni, nj, nk, precision
    (go.SINUSOID, 11, 16, 16, 8, 0),
    (go.SINUSOID_BLAST, 12, 24, 16, 16, 1),
    (go.DIPOLE_SHELL, 13, 16, 12, 16, 0),
    (go.SINUSOID_SPHERE, 14, 16, 8, 8, 1),
    (go.SINUSOID_BLAST, 15, 32, 24, 16, 0),
    # k extents past one 64-wide tile: staged E_i tiles and the leftover
    # k columns (k faces, E_i edges) in both forms
    (go.SINUSOID_BLAST, 16, 12, 8, 66, 0),
    (go.SINUSOID, 17, 8, 12, 65, 1),
    # pole ring-averaged grids (full sphere FP64, polar cap FP32-bulk)
    (go.POLE_SPHERE, 18, 8, 12, 32, 0),
    (go.POLE_CAP, 19, 8, 10, 64, 1),
    # background fields read per point, in both forms (their own launch
    # groups)
    (go.VARYING_B0_BOX, 20, 16, 12, 16, 1),
    (go.DIPOLE_B0_SHELL, 21, 12, 8, 12, 0),
]


def same(a, b):
    pa, ba = a.state()
    pb, bb = b.state()
    return all(np.array_equal(x, y) for x, y in zip(pa + ba, pb + bb))


def make_pairs():
    pairs = []
    for kind, seed, ni, nj, nk, prec in MIX:
        P = go.Problem(kind, seed, ni, nj, nk)
        pairs.append((go.Solver(P, prec), go.Solver(P, prec)))
        P.close()
    return pairs


@pytest.mark.gpu
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_alone(graphs):
    print(f"batch vs alone, graphs={graphs}", flush=True)
    pairs = make_pairs()
    B = go.Batch()
    B.set_graphs(graphs)
    try:
        for alone, inb in pairs:
            B.add(inb)
        B.step(7)           # phases 0..6: a 6-step graph and a single-step one
        late = pairs[1][1]
        B.remove(late)      # leaves; stepped alone for a while, then rejoins
        late.step(3)
        B.step(3)
        B.add(late)
        B.step(8)
        for alone, inb in pairs:
            alone.step(18)
        for q, (alone, inb) in enumerate(pairs):
            sa, sb = alone.scalars(), inb.scalars()
            print(q, "time", sa["time"], sb["time"], flush=True)
            assert sa["time"] == sb["time"] and sa["dt"] == sb["dt"]
            assert same(alone, inb), q
        if graphs:
            # every FP64 grid of the mix is bitwise the strict oracle's after
            # the same 18 steps (the batch joined and left around it)
            for q, (kind, seed, ni, nj, nk, prec) in enumerate(MIX):
                if prec != 0:
                    continue
                P = go.Problem(kind, seed, ni, nj, nk, full_b0=True)
                try:
                    O = run_oracle(P, 18)
                    prim, bf = pairs[q][1].state()
                    for f in range(6):
                        P.gas(f)[...] = prim[f]
                    for f in range(3):
                        P.gem(f)[...] = bf[f]
                    nbad = exact_mismatches(P, O)
                    print(q, "vs oracle: bitwise mismatches", nbad, flush=True)
                    assert nbad == 0, q
                finally:
                    P.close()
    finally:
        B.close()
        for a, b in pairs:
            a.close()
            b.close()


@pytest.mark.gpu
def test_autotune_keeps_results():
    P = go.Problem(go.SINUSOID_BLAST, 21, 32, 32, 32)
    A, Bs = go.Solver(P, 0), go.Solver(P, 0)
    P.close()
    try:
        best = A.autotune(2)
        print("autotuned chunk", best, flush=True)
        # a batched solver's launch shapes belong to the batch's tables
        B = go.Batch()
        try:
            B.add(Bs)
            with pytest.raises(go.GkError):
                Bs.set_tuning(go.TUNE_CHUNK, 16)
            B.remove(Bs)
        finally:
            B.close()
        assert best in (8, 16, 32, 64, 128, 256)
        assert A.get_tuning(go.TUNE_CHUNK) == best
        A.step(10)
        Bs.step(10)
        assert same(A, Bs)
    finally:
        A.close()
        Bs.close()


@pytest.mark.gpu
def test_timing_controls():
    """Per-stage timing: every stage, then one stage only (the others read
    zero); batch run timing accumulates the device time of synced runs."""
    P = go.Problem(go.SINUSOID_BLAST, 31, 32, 24, 16)
    S = go.Solver(P, 0)
    T = go.Solver(P, 1)
    P.close()
    B = None
    try:
        S.enable_timing(True)
        S.step(3)
        kt = S.kernel_times()
        print("all stages", kt, flush=True)
        assert kt["k_cell"] > 0.0 and kt["k_sweep_i"] > 0.0 and kt["k_ghost"] > 0.0
        S.time_stage("k_cell")
        S.step(3)
        kt = S.kernel_times()
        assert kt["k_cell"] > 0.0
        assert all(v == 0.0 for k, v in kt.items() if k != "k_cell"), kt
        S.enable_timing(False)
        S.step(1)
        B = go.Batch()
        B.add(S)
        B.add(T)
        B.run_timing(True)
        assert B.run_ms() == 0.0
        B.step(6)
        ms = B.run_ms()
        assert ms > 0.0
        B.step_async(2)
        B.sync()
        assert B.run_ms() > ms
        B.run_timing(True)  # resets the total
        assert B.run_ms() == 0.0
    finally:
        if B is not None:
            B.close()
        S.close()
        T.close()
