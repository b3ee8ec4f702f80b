"""Ghost layers of the state read back from the device.

The reference's arrays carry 4 ghost layers per side, filled at the end of
every step by the periodic exchange (src/setup_mpi.c:62-253,
src/boundary_exchange.c:7-201), and its dumps write them.  On a periodic box
every ghost value must therefore equal its periodic image exactly, for the
FP64 form, for the FP32-bulk form (whose steps fill only the FP32 copies'
ghosts, the FP64 ones being filled when the state is read back), and for a
grid stepped inside a batch.
"""
import numpy as np
import pytest

import gamera_op as go

NG = 4


def periodic_src(n, face):
    """Source index along one axis of length n + 9 (n active cells): low
    ghosts take the last active layers, high ghosts the first; a face axis
    keeps its upper boundary face n + 4 (the slot n + 8 of a cell axis is
    never written and maps to itself)."""
    N = n + 9
    src = np.arange(N)
    lo = np.arange(NG)
    src[lo] = lo + n
    if face:
        hi = np.arange(n + NG + 1, N)
    else:
        hi = np.arange(n + NG, n + 2 * NG)
    src[hi] = hi - n
    return src


def check_periodic(prim, bface, n):
    for f, a in enumerate(prim):
        s = [periodic_src(n[d], False) for d in range(3)]
        assert np.array_equal(a, a[np.ix_(*s)]), f"primitive {f}"
    for f, a in enumerate(bface):
        s = [periodic_src(n[d], d == f) for d in range(3)]
        assert np.array_equal(a, a[np.ix_(*s)]), f"face B {f}"


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_ghosts_are_periodic_images(precision):
    n = (16, 12, 8)
    print(f"ghosts precision={precision} {n}", flush=True)
    P = go.Problem(go.SINUSOID_BLAST, 21, *n)
    S = go.Solver(P, precision)
    P.close()
    try:
        S.step(7)
        prim, bf = S.state()
        assert np.isfinite(prim[0]).all()
        check_periodic(prim, bf, n)
    finally:
        S.close()


@pytest.mark.gpu
def test_ghosts_after_batched_steps():
    n = (12, 16, 8)
    print(f"ghosts batched {n}", flush=True)
    solvers = []
    for prec, seed in ((0, 22), (1, 23)):
        P = go.Problem(go.SINUSOID, seed, *n)
        solvers.append(go.Solver(P, prec))
        P.close()
    B = go.Batch()
    try:
        for S in solvers:
            B.add(S)
        B.step(9)
        for S in solvers:
            prim, bf = S.state()
            check_periodic(prim, bf, n)
    finally:
        B.close()
        for S in solvers:
            S.close()
