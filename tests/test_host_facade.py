"""The C host facade (reference-compatible globals, geometry, weights,
background field, problem plugin, boundary library) against the compiled
reference: every array bit for bit."""
import ctypes
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import gamera_op as go
from parity import active_slices, oracle_from_problem

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLD, "ref_hashes.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _cases():
    for key in HASHES["geometry"]:
        k, s, dims = key.split("_")
        ni, nj, nk = map(int, dims.split("x"))
        yield int(k[1:]), int(s[1:]), ni, nj, nk, key


@pytest.mark.parametrize("kind,seed,ni,nj,nk,key", list(_cases()))
def test_geometry_weights_b0_bitwise(kind, seed, ni, nj, nk, key):
    h = HASHES["geometry"][key]
    P = go.Problem(kind, seed, ni, nj, nk, full_b0=True)
    try:
        bad = []
        for f in range(28):
            if sha(P.geo(f)) != h[f"geo{f}"]:
                bad.append(f"geo{f}")
        for f in range(42):
            if sha(P.rec(f)) != h[f"rec{f}"]:
                bad.append(f"rec{f}")
        for f in range(52, 91):
            if sha(P.gem(f) + 0.0) != h[f"gem{f}"]:  # signed zeros normalised
                bad.append(f"gem{f}")
        for name in go.COORD_NAMES:
            if sha(P.coord_array(name)) != h[name]:
                bad.append(name)
        assert not bad, bad
        # the solve() prologue (ghost fill, BCs, derived variables, dt0)
        O = oracle_from_problem(P)
        O.prologue()
        cell, faces = active_slices(P)
        for f in range(6):
            assert sha(O.gas[f][cell]) == h[f"init_prim{f}"], f
        for f in range(3):
            assert sha(O.gem[f][faces[f]]) == h[f"init_B{f}"], f
        assert O.scalar("dt0") == h["dt0"]
    finally:
        P.close()


def test_boundary_kinds_recorded_from_plugin():
    P = go.Problem(go.DIPOLE_SHELL, 1, 12, 10, 12)
    try:
        bc = P.bc_table()  # 1 = symmetric, 0 = not requested (periodic axis)
        assert bc.tolist() == [[1, 1, 1, 1, 0, 0], [1, 1, 1, 1, 0, 0]]
    finally:
        P.close()
    P = go.Problem(go.POLE_SPHERE, 1, 12, 12, 32)
    try:  # 6 = pole boundary on both theta sides
        assert P.bc_table().tolist() == [[1, 1, 6, 6, 0, 0], [1, 1, 6, 6, 0, 0]]
    finally:
        P.close()
    P = go.Problem(go.WALL_BOX, 1, 12, 10, 8)
    try:  # 2 = reflective (gem: recorded as none, 5), 4 = extrapolated
        assert P.bc_table().tolist() == [[2, 2, 4, 4, 0, 0], [5, 5, 4, 4, 0, 0]]
    finally:
        P.close()
    P = go.Problem(go.INFLOW_BOX, 1, 12, 10, 8)
    try:  # 3 = fixed, 4 = extrapolated, 1 = symmetric
        assert P.bc_table().tolist() == [[3, 4, 1, 1, 0, 0], [3, 4, 1, 1, 0, 0]]
    finally:
        P.close()
    P = go.Problem(go.SINUSOID, 1, 12, 10, 8)
    try:
        assert P.bc_table().tolist() == [[0] * 6, [0] * 6]
    finally:
        P.close()


def test_background_field_mode_for_device():
    """The device takes B0 as constants only where the reference's arrays
    hold one value bit for bit over the points it reads (mode 1), and per
    point otherwise (mode 2): a uniform field on widths 1/16 is the first,
    on widths 1/12 (face quadratures that differ in the last bits) or a
    field varying in space the second; no field is mode 0."""
    for kind, n, mode in [(go.SINUSOID_BLAST, (16, 16, 16), 1),
                          (go.SINUSOID_BLAST, (12, 12, 12), 2),
                          (go.SINUSOID_BLAST, (32, 24, 16), 2),
                          (go.VARYING_B0_BOX, (16, 16, 16), 2),
                          (go.DIPOLE_B0_SHELL, (16, 12, 16), 2),
                          (go.SINUSOID, (12, 12, 12), 0)]:
        P = go.Problem(kind, 3, *n)
        try:
            assert P.b0_mode() == mode, (kind, n)
        finally:
            P.close()


def test_read_config_yaml_schema():
    """config.yaml keys of the reference (src/config.c:150-163); unknown keys
    are ignored."""
    h, _ = go._load()
    h.read_config.restype = ctypes.c_int
    h.read_config.argtypes = [ctypes.c_char_p]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "config.yaml")
        with open(p, "w") as f:
            f.write("# comment\nni_global: 64\nnj_global: 32\nnk_global: 16\nproc_dims_i: 2\n"
                    "proc_dims_j: 1\nproc_dims_k: 1\nNt: 7\ntime_stop: 2.5\n"
                    "output_interval: 0.5\nx_Norm: 1.0\nu_Norm: 2.0\nrho_Norm: 3.0\nseed: 42\n")
        assert h.read_config(p.encode()) == 0
    sizes = (ctypes.c_int * 6)()
    h.gkh_sizes(sizes)
    assert list(sizes) == [32, 32, 16, 41, 41, 25]
    assert ctypes.c_int.in_dll(h, "Nt").value == 7
    assert ctypes.c_double.in_dll(h, "time_stop").value == 2.5
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "config.yaml")
        with open(p, "w") as f:
            f.write("ni_global: 30\nnj_global: 32\nnk_global: 16\nproc_dims_i: 4\n"
                    "proc_dims_j: 1\nproc_dims_k: 1\n")
        assert h.read_config(p.encode()) != 0  # 30 not divisible by 4


def test_ring_average_needs_a_supported_resolution():
    """The reference's ring table covers nk = 32, 64, 128, 256 (it exits
    otherwise); the facade refuses to describe such a grid for the device."""
    P = go.Problem(go.POLE_SPHERE, 1, 8, 12, 48)
    try:
        with pytest.raises(go.GkError):
            P.desc_blob()
    finally:
        P.close()
    P = go.Problem(go.POLE_SPHERE, 1, 8, 12, 32)
    try:
        assert len(P.desc_blob()) > 0
    finally:
        P.close()


def _stored(P, which):
    out = {}
    for f in range(64):
        try:
            out[f] = (P.gas(f) if which == 0 else P.gem(f)).copy()
        except go.GkError:
            continue
        except Exception:
            break
    return out


@pytest.mark.parametrize("kind,shape", [(go.SINUSOID_SPHERE, (24, 20, 16)),
                                        (go.DIPOLE_SHELL, (16, 12, 16)),
                                        (go.VARYING_B0_BOX, (16, 12, 8))])
def test_reseed_equals_fresh_setup(kind, shape):
    """gkh_reseed on a grid that was set up (and whose state arrays were
    then overwritten, as a device read-back does) holds, array for array and
    bit for bit, what a fresh set-up with the new seed holds."""
    P = go.Problem(kind, 5, *shape)
    try:
        for f in range(6):
            P.gas(f)[...] = np.pi  # a stale state the reseed must clear
        P.gem(0)[...] = -1.0
        P.reseed(9)
        got = (_stored(P, 0), _stored(P, 1))
    finally:
        P.close()
    Q = go.Problem(kind, 9, *shape)
    try:
        want = (_stored(Q, 0), _stored(Q, 1))
    finally:
        Q.close()
    for g, w in zip(got, want):
        assert sorted(g) == sorted(w)
        assert len(g) > 6
        for f in w:
            assert np.array_equal(g[f].view(np.uint64), w[f].view(np.uint64)), f
