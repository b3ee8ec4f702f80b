"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

The oracle is the checker: tests load it, hand it the same geometry and
initial state as the device solver and compare.  Arrays use the reference
layout: field-major [f][i][j][k], NI = ni + 9 points per axis.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

NF_GAS, NF_GEM, NF_GEO, NF_REC = 82, 97, 28, 42
GO_GAS, GO_GEM, GO_GEO, GO_REC = 0, 1, 2, 3
COORD_NAMES = ["x1", "x2", "x3", "x1c", "x2c", "x3c", "x1ctr", "x2ctr",
               "x3ctr", "dx1", "dx2", "dx3"]
BC_PERIODIC, BC_SYMMETRIC, BC_NONE, BC_POLE = 0, 1, 2, 3

_libs = {}


def load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run make -C oracle)")
    lib = ctypes.CDLL(path)
    lib.go_create.restype = ctypes.c_void_p
    lib.go_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int), ctypes.c_double]
    lib.go_destroy.argtypes = [ctypes.c_void_p]
    lib.go_array.restype = ctypes.POINTER(ctypes.c_double)
    lib.go_array.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.go_npoints.restype = ctypes.c_int64
    lib.go_npoints.argtypes = [ctypes.c_void_p]
    lib.go_prologue.argtypes = [ctypes.c_void_p]
    lib.go_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.go_scalar.restype = ctypes.c_double
    lib.go_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.go_debug_fluxes.restype = ctypes.POINTER(ctypes.c_double)
    lib.go_debug_fluxes.argtypes = [ctypes.c_void_p]
    D7 = ctypes.c_double * 7
    lib.go_pdmu7.restype = ctypes.c_double
    lib.go_pdmu7.argtypes = [D7, D7, D7, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double]
    lib.go_face_flux.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 5 + [
        ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.go_set_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    lib.go_set_ring.restype = ctypes.c_int
    lib.go_set_ring.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    _libs[name] = lib
    return lib


class Oracle:
    """One oracle solver instance (reference layout, single rank)."""

    def __init__(self, ni, nj, nk, coord=0, bc=(0,) * 6, CA=10.0, fast=False):
        self.lib = load(fast)
        arr = (ctypes.c_int * 6)(*bc)
        self.h = self.lib.go_create(ni, nj, nk, coord, arr, CA)
        if not self.h:
            raise MemoryError("go_create failed")
        self.shape = (ni + 9, nj + 9, nk + 9)
        self.n3 = int(self.lib.go_npoints(self.h))

    def _view(self, which, nf):
        p = self.lib.go_array(self.h, which)
        return np.ctypeslib.as_array(p, shape=(nf,) + self.shape)

    @property
    def gas(self):
        return self._view(GO_GAS, NF_GAS)

    @property
    def gem(self):
        return self._view(GO_GEM, NF_GEM)

    @property
    def geo(self):
        return self._view(GO_GEO, NF_GEO)

    @property
    def rec(self):
        return self._view(GO_REC, NF_REC)

    def coord(self, name):
        p = self.lib.go_array(self.h, 4 + COORD_NAMES.index(name))
        return np.ctypeslib.as_array(p, shape=self.shape)

    def debug_fluxes(self):
        p = self.lib.go_debug_fluxes(self.h)
        return np.ctypeslib.as_array(p, shape=(45,) + self.shape)

    def set_ring(self, on, pole_lo, pole_hi):
        """Turn pole ring averaging on (after the geometry is loaded); returns
        the number of averaged rows per pole."""
        n = self.lib.go_set_ring(self.h, int(on), int(pole_lo), int(pole_hi))
        if n < 0:
            raise ValueError("ring averaging needs nk in 32, 64, 128, 256")
        return n

    def prologue(self):
        self.lib.go_prologue(self.h)

    def step(self, n=1):
        self.lib.go_step(self.h, n)

    SCALARS = ["dt", "dt0", "time_sim", "divB_max", "rho_floor", "p_floor", "G01", "G02", "G03"]

    def scalar(self, name):
        return self.lib.go_scalar(self.h, self.SCALARS.index(name))

    def set_scalar(self, name, v):
        q = name if isinstance(name, int) else self.SCALARS.index(name)
        self.lib.go_set_scalar(self.h, q, v)

    def close(self):
        if self.h:
            self.lib.go_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def ring_filter(q):
    """FilterBaseMode on a copy of q: returns (remainder, m = 0, 1 part)."""
    lib = load()
    q = np.ascontiguousarray(q, dtype=np.float64).copy()
    base = np.zeros_like(q)
    P = ctypes.POINTER(ctypes.c_double)
    lib.go_ring_filter.argtypes = [P, ctypes.c_int, P]
    lib.go_ring_filter(q.ctypes.data_as(P), len(q), base.ctypes.data_as(P))
    return q, base


def ring_block_avg(q, nch):
    """RingBlockAvg on a copy of q with nch blocks."""
    lib = load()
    q = np.ascontiguousarray(q, dtype=np.float64).copy()
    P = ctypes.POINTER(ctypes.c_double)
    lib.go_ring_block_avg.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.go_ring_block_avg(q.ctypes.data_as(P), len(q), nch)
    return q


def pdmu7(fh, fp, w, faceim1, facei, vol, dx):
    lib = load()
    D7 = ctypes.c_double * 7
    return lib.go_pdmu7(D7(*fh), D7(*fp), D7(*w), faceim1, facei, vol, dx)


def face_flux(dir_, L, R, bl, br, b0, Bn, CA):
    lib = load()

    def arr(v):
        return (ctypes.c_double * len(v))(*v)
    out = (ctypes.c_double * 9)()
    lib.go_face_flux(dir_, arr(L), arr(R), arr(bl), arr(br), arr(b0), Bn, CA, out)
    return np.array(out[:])
