"""gk_get_state_noalloc (the read-back that allocates nothing) against
gk_get_state, through the C ABI.

Two calls in a row on different solvers, the second on a smaller grid (and
the FP32-bulk form, whose read-back first fills the FP64 ghost layers): every
array, both time levels, must equal gk_get_state's bit for bit.  A grid whose
field does not fit the scratch array is refused with GK_ERR_ARG.
"""
import ctypes

import numpy as np
import pytest

import gamera_op as go

GK_ERR_ARG = -1


def read(op, S, fn):
    shape = S.shape
    arrs = [np.full(shape, np.nan) for _ in range(18)]
    ptrs = [(ctypes.c_void_p * n)(*[a.ctypes.data for a in arrs[o:o + n]])
            for o, n in ((0, 6), (6, 3), (9, 6), (15, 3))]
    rc = fn(S.handle, ptrs[0], ptrs[1], ptrs[2], ptrs[3])
    return rc, arrs


def solver(kind, seed, n, precision, steps):
    P = go.Problem(kind, seed, *n)
    try:
        S = go.Solver(P, precision)
    finally:
        P.close()
    S.step(steps)
    return S


@pytest.mark.gpu
def test_noalloc_read_back_equals_get_state():
    _, op = go._load()
    fn = op.gk_get_state_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_void_p)] * 4
    ref = op.gk_get_state
    cases = [(go.SINUSOID_BLAST, 5, (40, 32, 24), 0, 7),   # first call
             (go.DIPOLE_SHELL, 6, (16, 12, 16), 1, 5)]     # then a smaller grid, FP32-bulk
    for kind, seed, n, prec, steps in cases:
        print(f"noalloc read-back kind={kind} {n} precision={prec}", flush=True)
        S = solver(kind, seed, n, prec, steps)
        try:
            rc0, want = read(op, S, ref)
            rc1, got = read(op, S, fn)
            assert rc0 == 0 and rc1 == 0, (rc0, rc1)
            for q, (a, b) in enumerate(zip(want, got)):
                assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), q
            assert np.isfinite(got[0]).all()
        finally:
            S.close()
    # a field of more points than the scratch holds: refused, nothing copied
    S = solver(go.SINUSOID, 7, (153, 153, 153), 0, 0)
    try:
        rc, arrs = read(op, S, fn)
        assert rc == GK_ERR_ARG
        assert np.isnan(arrs[0]).all()
    finally:
        S.close()
