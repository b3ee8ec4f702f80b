"""Python bindings for the GAMERA-OP MI355X solver (ctypes over the C ABI).

Two shared libraries are loaded from this directory:
  libgamera_op.so    the HIP device solver (include/gamera_op.h)
  libgamera_host.so  the C host facade: reference-compatible globals,
                     geometry, problem plugins, initialize_solver/solve/
                     finalize_solver (reference src/solver.h:17-68)

The host facade is global state, one problem at a time, as in the reference.
Loading fails loudly when the libraries are missing: there is no CPU fallback
on the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_OP = os.path.join(HERE, "libgamera_op.so")
LIB_HOST = os.path.join(HERE, "libgamera_host.so")

# seeded problem generators (gamera-op_amd/problems/gk_seeded.c)
SINUSOID, SINUSOID_BLAST, DIPOLE_SHELL, SINUSOID_SPHERE, POLE_SPHERE, POLE_CAP = 0, 1, 2, 3, 4, 5
WALL_BOX, INFLOW_BOX, ENTROPY_BOX, FLOOR_BOX = 6, 7, 8, 9
VARYING_B0_BOX, DIPOLE_B0_SHELL = 10, 11

# gk_status codes (include/gamera_op.h)
GK_OK, GK_ERR_ARG, GK_ERR_DEVICE, GK_ERR_NOMEM, GK_ERR_STATE, GK_ERR_NAN = 0, -1, -2, -3, -4, -5

NF_GAS, NF_GEM, NF_GEO, NF_REC = 82, 97, 28, 42
COORD_NAMES = ["x1", "x2", "x3", "x1c", "x2c", "x3c", "x1ctr", "x2ctr", "x3ctr",
               "dx1", "dx2", "dx3"]

_host = None
_op = None


def _load():
    global _host, _op
    if _host is not None:
        return _host, _op
    for p in (LIB_OP, LIB_HOST):
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is not built; run __graft_entry__.build()")
    _op = ctypes.CDLL(LIB_OP, mode=ctypes.RTLD_GLOBAL)
    h = ctypes.CDLL(LIB_HOST, mode=ctypes.RTLD_GLOBAL)
    h.gkh_setup.restype = ctypes.c_int
    h.gkh_setup.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, ctypes.c_int]
    h.gkh_ptr.restype = ctypes.POINTER(ctypes.c_double)
    h.gkh_ptr.argtypes = [ctypes.c_int, ctypes.c_int]
    h.gkh_sizes.argtypes = [ctypes.POINTER(ctypes.c_int)]
    h.gkh_bc_table.argtypes = [ctypes.POINTER(ctypes.c_int)]
    h.gkh_CA.restype = ctypes.c_double
    h.gkh_b0_mode.restype = ctypes.c_int
    h.gkh_device_open.restype = ctypes.c_int
    h.gkh_device_open.argtypes = [ctypes.c_int]
    h.gkh_device_fetch.restype = ctypes.c_int
    h.gkh_device.restype = ctypes.c_void_p
    h.solve.restype = ctypes.c_int
    h.solve.argtypes = [ctypes.c_int]
    h.finalize_solver.restype = ctypes.c_int
    _op.gk_step.restype = ctypes.c_int
    _op.gk_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    _op.gk_get_scalars.restype = ctypes.c_int
    _op.gk_get_scalars.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    _op.gk_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    _op.gk_time_stage.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    _op.gk_kernel_times.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p),
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    _op.gk_debug_fetch.restype = ctypes.c_int
    _op.gk_debug_fetch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double)]
    _op.gk_device_bytes.restype = ctypes.c_size_t
    _op.gk_device_bytes.argtypes = [ctypes.c_void_p]
    _op.gk_set_tuning.restype = ctypes.c_int
    _op.gk_set_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    _op.gk_get_tuning.restype = ctypes.c_int
    _op.gk_get_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int]
    vp = ctypes.c_void_p
    h.gkh_desc_blob.restype = ctypes.c_long
    h.gkh_desc_blob.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    h.gkh_solver_from_blob.restype = ctypes.c_int
    h.gkh_solver_from_blob.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_void_p)]
    h.gkh_main.restype = ctypes.c_int
    h.gkh_main.argtypes = [ctypes.c_char_p, ctypes.c_int]
    h.gkh_reseed.restype = ctypes.c_int
    h.gkh_reseed.argtypes = [ctypes.c_uint64]
    h.gkh_device_create.restype = ctypes.c_int
    h.gkh_device_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    _op.gk_batch_budget.restype = ctypes.c_longlong
    _op.gk_batch_budget.argtypes = [vp, vp]
    for name, args in [("gk_autotune", [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
                       ("gk_step_until", [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_int)]),
                       ("gk_batch_set_budget", [vp, vp, ctypes.c_longlong]),
                       ("gk_batch_status", [vp, vp]),
                       ("gk_destroy", [vp]),
                       ("gk_get_state", [vp] + [ctypes.POINTER(ctypes.c_void_p)] * 4),
                       ("gk_batch_create", [ctypes.POINTER(vp)]),
                       ("gk_batch_destroy", [vp]),
                       ("gk_batch_add", [vp, vp]),
                       ("gk_batch_remove", [vp, vp]),
                       ("gk_batch_size", [vp]),
                       ("gk_batch_step", [vp, ctypes.c_int]),
                       ("gk_batch_step_async", [vp, ctypes.c_int]),
                       ("gk_batch_sync", [vp]),
                       ("gk_batch_set_graphs", [vp, ctypes.c_int]),
                       ("gk_batch_side_stream", [vp, ctypes.POINTER(vp)]),
                       ("gk_batch_run_timing", [vp, ctypes.c_int]),
                       ("gk_batch_run_ms", [vp, ctypes.POINTER(ctypes.c_double)]),
                       ("gk_batch_enable_timing", [vp, ctypes.c_int]),
                       ("gk_batch_kernel_times", [vp, ctypes.POINTER(ctypes.c_char_p),
                                                  ctypes.POINTER(ctypes.c_double),
                                                  ctypes.POINTER(ctypes.c_int)])]:
        f = getattr(_op, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    _op.gk_copy_calibrate.restype = ctypes.c_int
    _op.gk_copy_calibrate.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    _host = h
    return _host, _op


def exported_symbols():
    """Names of the C-ABI entry points declared in include/gamera_op.h."""
    return ["gk_create", "gk_destroy", "gk_set_state", "gk_prologue", "gk_step",
            "gk_get_state", "gk_get_scalars", "gk_set_time", "gk_enable_timing",
            "gk_kernel_times", "gk_time_stage", "gk_device_bytes"]


def run_main(config_file, nt):
    """The reference main() sequence on one GPU in the current directory:
    config file, restart from this rank's last dump if there is one, problem
    set-up otherwise, then nt steps with HDF5 dumps at the reference cadence.
    The host arrays stay readable through HostArrays until finalize()."""
    h, _ = _load()
    rc = h.gkh_main(config_file.encode(), nt)
    if rc != 0:
        raise GkError(f"gkh_main failed ({rc})")
    return HostArrays()


def finalize():
    h, _ = _load()
    h.finalize_solver()


class HostArrays:
    """Views of the host facade's current arrays (reference layout)."""

    def __init__(self):
        self.h, self.op = _load()
        sz = (ctypes.c_int * 6)()
        self.h.gkh_sizes(sz)
        self.n = (sz[0], sz[1], sz[2])
        self.shape = (sz[3], sz[4], sz[5])

    def _view(self, which, f):
        p = self.h.gkh_ptr(which, f)
        if not p:
            raise GkError(f"array {which}/{f} is not stored on the host")
        return np.ctypeslib.as_array(p, shape=self.shape)

    def gas(self, f):
        return self._view(0, f)

    def gem(self, f):
        return self._view(1, f)

    def scalar(self, name):
        return ctypes.c_double.in_dll(self.h, name).value

    def int(self, name):
        return ctypes.c_int.in_dll(self.h, name).value


def copy_calibrate(nbytes, reps=20):
    """Average duration (ms) of a plain device copy of nbytes (read nbytes,
    write nbytes)."""
    _, op = _load()
    ms = ctypes.c_double(0.0)
    rc = op.gk_copy_calibrate(nbytes, reps, ctypes.byref(ms))
    if rc != 0:
        raise GkError(f"gk_copy_calibrate failed ({rc})")
    return ms.value


TUNE_CHUNK = 0
TUNE_EMF_CHUNK = 1
TUNE_EMF_STAGE = 2


class GkError(RuntimeError):
    pass


class GkNonFinite(GkError):
    """The device state holds a non-finite value (GK_ERR_NAN)."""


class Problem:
    """Host facade state for one seeded problem (global, one at a time)."""

    def __init__(self, kind, seed, ni, nj, nk, coord=None, full_b0=False):
        self.h, self.op = _load()
        if coord is None:
            coord = 2 if kind in (DIPOLE_SHELL, POLE_SPHERE, POLE_CAP, DIPOLE_B0_SHELL) else 0
        rc = self.h.gkh_setup(kind, seed, ni, nj, nk, coord, 1 if full_b0 else 0)
        if rc != 0:
            raise GkError(f"gkh_setup failed ({rc})")
        sz = (ctypes.c_int * 6)()
        self.h.gkh_sizes(sz)
        self.n = (sz[0], sz[1], sz[2])
        self.shape = (sz[3], sz[4], sz[5])
        self.kind, self.seed, self.coord = kind, seed, coord
        # the host always computes the reference's background-field arrays
        # (full_b0 is accepted for compatibility and has no effect)
        self.full_b0 = True
        self.CA = self.h.gkh_CA()

    def _view(self, which, f):
        p = self.h.gkh_ptr(which, f)
        if not p:
            raise GkError(f"array {which}/{f} is not stored on the host")
        return np.ctypeslib.as_array(p, shape=self.shape)

    def reseed(self, seed):
        """Same grid, new seeded initial state (skips the geometry set-up)."""
        rc = self.h.gkh_reseed(seed)
        if rc != 0:
            raise GkError(f"gkh_reseed failed ({rc})")
        self.seed = seed

    def gas(self, f):
        return self._view(0, f)

    def gem(self, f):
        return self._view(1, f)

    def desc_blob(self):
        """The problem's grid description (geometry, weights, BCs, B0) as
        bytes; with the initial state it is all a solver needs."""
        n = self.h.gkh_desc_blob(None, 0)
        if n < 0:
            raise GkError(f"gkh_desc_blob failed ({n})")
        buf = ctypes.create_string_buffer(n)
        m = self.h.gkh_desc_blob(buf, n)
        if m != n:
            raise GkError(f"gkh_desc_blob failed ({m})")
        return buf.raw

    def geo(self, f):
        return self._view(2, f)

    def rec(self, f):
        return self._view(3, f)

    def coord_array(self, name):
        return self._view(4 + COORD_NAMES.index(name), 0)

    def bc_table(self):
        t = (ctypes.c_int * 12)()
        self.h.gkh_bc_table(t)
        return np.array(t[:], dtype=int).reshape(2, 6)

    def b0_mode(self):
        return self.h.gkh_b0_mode()

    # --- device path -----------------------------------------------------
    def device_open(self, precision=0):
        rc = self.h.gkh_device_open(precision)
        if rc != 0:
            raise GkError(f"gkh_device_open failed ({rc})")
        self.dev = self.h.gkh_device()

    def device_step(self, n):
        _check(self.op.gk_step(self.dev, n), "gk_step")

    def device_fetch(self):
        rc = self.h.gkh_device_fetch()
        if rc != 0:
            raise GkError(f"gkh_device_fetch failed ({rc})")

    def device_scalars(self):
        out = (ctypes.c_double * 6)()
        rc = self.op.gk_get_scalars(self.dev, out)
        if rc != 0:
            raise GkError(f"gk_get_scalars failed ({rc})")
        return dict(dt=out[0], dt0=out[1], time=out[2], divB=out[3], next_dt=out[4],
                    nonfinite=out[5])

    def enable_timing(self, on=True):
        self.op.gk_enable_timing(self.dev, 1 if on else 0)

    def time_stage(self, name):
        """Time only the launches of one stage (gk_time_stage)."""
        _check(self.op.gk_time_stage(self.dev, name.encode() if name else None), "gk_time_stage")

    def kernel_times(self):
        n = ctypes.c_int(32)
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_double * 32)()
        self.op.gk_kernel_times(self.dev, names, ms, ctypes.byref(n))
        return {names[q].decode(): ms[q] for q in range(min(n.value, 32))}

    def debug_fetch(self, which, field):
        out = np.zeros(self.shape)
        rc = self.op.gk_debug_fetch(self.dev, which, field,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc != 0:
            raise GkError(f"gk_debug_fetch failed ({rc})")
        return out

    def set_tuning(self, param, value):
        rc = self.op.gk_set_tuning(self.dev, param, value)
        if rc != 0:
            raise GkError(f"gk_set_tuning({param}, {value}) failed ({rc})")

    def get_tuning(self, param):
        return self.op.gk_get_tuning(self.dev, param)

    def device_bytes(self):
        return int(self.op.gk_device_bytes(self.dev))

    def solve(self, nt):
        rc = self.h.solve(nt)
        if rc != 0:
            raise GkError(f"solve failed ({rc})")

    def close(self):
        self.h.finalize_solver()


def _check(rc, what):
    if rc == GK_ERR_NAN:
        raise GkNonFinite(f"{what}: non-finite state")
    if rc != 0:
        raise GkError(f"{what} failed ({rc})")


class Solver:
    """A standalone device solver for the problem the host facade currently
    holds (created after Problem(...)); it keeps its own device state, so the
    next Problem can be set up while this one runs."""

    def __init__(self, problem, precision=0):
        self.h, self.op = _load()
        p = ctypes.c_void_p()
        _check(self.h.gkh_device_create(precision, ctypes.byref(p)), "gkh_device_create")
        self.handle = p
        self.n, self.shape = problem.n, problem.shape
        self.precision = precision

    @classmethod
    def from_parts(cls, blob, prim, bface, precision=0, stream=None):
        """Solver from a descriptor blob (Problem.desc_blob) and an initial
        state: 6 primitive and 3 face-B arrays in the reference layout."""
        self = cls.__new__(cls)
        self.h, self.op = _load()
        shape = prim[0].shape
        for a in list(prim) + list(bface):
            if a.shape != shape or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise GkError("state arrays must be C-contiguous float64 of one shape")
        P = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in prim])
        B = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in bface])
        p = ctypes.c_void_p()
        _check(self.h.gkh_solver_from_blob(blob, len(blob), precision, P, B, stream,
                                           ctypes.byref(p)), "gkh_solver_from_blob")
        self.handle = p
        self.shape = shape
        self.n = tuple(x - 9 for x in shape)
        self.precision = precision
        return self

    def step(self, n):
        _check(self.op.gk_step(self.handle, n), "gk_step")

    def step_until(self, n, t_stop=float("inf"), t_event=float("inf")):
        """Up to n steps, ending after the first step whose time exceeds
        t_stop or reaches t_event; returns the steps taken."""
        done = ctypes.c_int(0)
        _check(self.op.gk_step_until(self.handle, n, t_stop, t_event, ctypes.byref(done)),
               "gk_step_until")
        return done.value

    def scalars(self):
        out = (ctypes.c_double * 6)()
        _check(self.op.gk_get_scalars(self.handle, out), "gk_get_scalars")
        return dict(dt=out[0], dt0=out[1], time=out[2], divB=out[3], next_dt=out[4],
                    nonfinite=out[5])

    def state(self):
        """Primitives (6) and face B (3) in the reference layout."""
        prim = [np.zeros(self.shape) for _ in range(6)]
        bf = [np.zeros(self.shape) for _ in range(3)]
        P = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in prim])
        B = (ctypes.c_void_p * 3)(*[a.ctypes.data for a in bf])
        _check(self.op.gk_get_state(self.handle, P, B, None, None), "gk_get_state")
        return prim, bf

    def device_bytes(self):
        return int(self.op.gk_device_bytes(self.handle))

    def enable_timing(self, on=True):
        _check(self.op.gk_enable_timing(self.handle, 1 if on else 0), "gk_enable_timing")

    def time_stage(self, name):
        """Time only the launches of one stage (gk_time_stage; None: all)."""
        _check(self.op.gk_time_stage(self.handle, name.encode() if name else None), "gk_time_stage")

    def kernel_times(self):
        """Average ms per launch of each stage kernel over the last step() call
        (with timing enabled)."""
        n = ctypes.c_int(32)
        names = (ctypes.c_char_p * 32)()
        ms = (ctypes.c_double * 32)()
        _check(self.op.gk_kernel_times(self.handle, names, ms, ctypes.byref(n)), "gk_kernel_times")
        return {names[q].decode(): ms[q] for q in range(min(n.value, 32))}

    def autotune(self, trial_steps=3):
        """Time the candidate launch shapes (trial_steps = 0: apply a cached
        pick only; returns 0 when there is none)."""
        best = ctypes.c_int(0)
        _check(self.op.gk_autotune(self.handle, trial_steps, ctypes.byref(best)), "gk_autotune")
        return best.value

    def set_tuning(self, param, value):
        _check(self.op.gk_set_tuning(self.handle, param, value), "gk_set_tuning")

    def get_tuning(self, param):
        return self.op.gk_get_tuning(self.handle, param)

    def close(self):
        if self.handle:
            _check(self.op.gk_destroy(self.handle), "gk_destroy")
            self.handle = None


class Batch:
    """Steps many Solvers together (one launch per stage for all of them,
    captured in HIP graphs)."""

    def __init__(self):
        self.h, self.op = _load()
        p = ctypes.c_void_p()
        _check(self.op.gk_batch_create(ctypes.byref(p)), "gk_batch_create")
        self.handle = p
        self.members = []

    def add(self, solver):
        _check(self.op.gk_batch_add(self.handle, solver.handle), "gk_batch_add")
        self.members.append(solver)

    def remove(self, solver):
        _check(self.op.gk_batch_remove(self.handle, solver.handle), "gk_batch_remove")
        self.members.remove(solver)

    def step(self, n):
        _check(self.op.gk_batch_step(self.handle, n), "gk_batch_step")

    def step_async(self, n):
        _check(self.op.gk_batch_step_async(self.handle, n), "gk_batch_step_async")

    def sync(self):
        _check(self.op.gk_batch_sync(self.handle), "gk_batch_sync")

    def sync_status(self):
        """Wait for the launched steps; returns the members whose state
        holds a non-finite value (an empty list when there are none)."""
        rc = self.op.gk_batch_sync(self.handle)
        if rc == GK_ERR_NAN:
            return [m for m in self.members
                    if self.op.gk_batch_status(self.handle, m.handle) == GK_ERR_NAN]
        _check(rc, "gk_batch_sync")
        return []

    def set_budget(self, solver, steps):
        """The member takes `steps` more steps (counted from the steps
        already launched), then stops on the device while the batch goes on."""
        _check(self.op.gk_batch_set_budget(self.handle, solver.handle, steps),
               "gk_batch_set_budget")

    def budget(self, solver):
        return int(self.op.gk_batch_budget(self.handle, solver.handle))

    def set_graphs(self, on):
        _check(self.op.gk_batch_set_graphs(self.handle, 1 if on else 0), "gk_batch_set_graphs")

    def enable_timing(self, on=True):
        _check(self.op.gk_batch_enable_timing(self.handle, 1 if on else 0), "gk_batch_enable_timing")

    def side_stream(self):
        """The batch's side stream (gk_batch_side_stream) as a pointer value,
        for solvers whose set-up and read-back should run beside its steps."""
        st = ctypes.c_void_p()
        _check(self.op.gk_batch_side_stream(self.handle, ctypes.byref(st)), "gk_batch_side_stream")
        return st

    def run_timing(self, on=True):
        """Time each run on the device (gk_batch_run_timing); resets the total."""
        _check(self.op.gk_batch_run_timing(self.handle, 1 if on else 0), "gk_batch_run_timing")

    def run_ms(self):
        """Device milliseconds of the completed runs since run timing was turned on."""
        ms = ctypes.c_double()
        _check(self.op.gk_batch_run_ms(self.handle, ctypes.byref(ms)), "gk_batch_run_ms")
        return ms.value

    def kernel_times(self):
        n = ctypes.c_int(64)
        names = (ctypes.c_char_p * 64)()
        ms = (ctypes.c_double * 64)()
        self.op.gk_batch_kernel_times(self.handle, names, ms, ctypes.byref(n))
        return {names[q].decode(): ms[q] for q in range(min(n.value, 64))}

    def close(self):
        if self.handle:
            for m in list(self.members):
                self.remove(m)
            _check(self.op.gk_batch_destroy(self.handle), "gk_batch_destroy")
            self.handle = None
