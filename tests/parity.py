"""Helpers that run the CPU oracle on a host-facade problem and compare
device results with it (test infrastructure)."""
import numpy as np

from oracle_lib import Oracle, COORD_NAMES

GAS_NAMES = ["rho", "v1", "v2", "v3", "p", "p_S"]
B_NAMES = ["Bi", "Bj", "Bk"]


def oracle_from_problem(P, fast=False):
    """An oracle loaded with the problem's geometry, weights, B0 arrays and
    initial primitive state (before the prologue)."""
    ni, nj, nk = P.n
    bc = P.bc_table()
    kinds = []
    periodic = [P.coord == 0, P.coord == 0, True] if P.kind not in (2, 4, 5, 6, 7, 11) else [False, False, True]
    # recorded gas kinds (host) -> oracle kinds: symmetric, reflective,
    # fixed, extrapolated, pole; anything else leaves the ghosts as they are
    to_oracle = {1: 1, 2: 4, 3: 6, 4: 5, 6: 3}
    for dim in range(3):
        for side in range(2):
            if periodic[dim]:
                kinds.append(0)
            else:
                kinds.append(to_oracle.get(bc[0][2 * dim + side], 2))
    O = Oracle(ni, nj, nk, coord=P.coord, bc=tuple(kinds), CA=P.CA, fast=fast)
    for f in range(28):
        O.geo[f] = P.geo(f)
    for f in range(42):
        O.rec[f] = P.rec(f)
    for name in COORD_NAMES:
        O.coord(name)[:] = P.coord_array(name)
    for f in range(52, 91):
        O.gem[f] = P.gem(f)
    for f in range(6):
        O.gas[f] = P.gas(f)
    for f in range(3):
        O.gem[f] = P.gem(f)
    # the problem's floors and gravity (globals the plugin sets)
    import ctypes
    h = P.h
    for q, name in ((4, "rho_floor"), (5, "p_floor"), (6, "G01"), (7, "G02"), (8, "G03")):
        O.set_scalar(q, ctypes.c_double.in_dll(h, name).value)
    if P.kind == 4:  # full sphere: ring averaging at both poles
        O.set_ring(1, 1, 1)
    if P.kind == 5:  # northern hemisphere: the pole at theta = 0 only
        O.set_ring(1, 1, 0)
    return O


def run_oracle(P, steps, fast=False):
    O = oracle_from_problem(P, fast=fast)
    O.prologue()
    O.step(steps)
    return O


def active_slices(P):
    ni, nj, nk = P.n
    a, b, c = slice(4, 4 + ni), slice(4, 4 + nj), slice(4, 4 + nk)
    faces = [(slice(4, 5 + ni), b, c), (a, slice(4, 5 + nj), c), (a, b, slice(4, 5 + nk))]
    return (a, b, c), faces


def rel_l1(x, y):
    den = np.abs(y).sum()
    num = np.abs(x - y).sum()
    if den == 0.0:
        return float(num)
    return float(num / den)


def conserved(g):
    rho, v1, v2, v3, p = g[0], g[1], g[2], g[3], g[4]
    return {"rho": rho, "rhov1": rho * v1, "rhov2": rho * v2, "rhov3": rho * v3,
            "eng": 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3) + p / (5.0 / 3.0 - 1.0)}


def compare_states(P, O):
    """Relative L1 of the device state (already fetched into the host arrays)
    against the oracle, per conserved variable and face-B component.  Fields
    that are identically zero in the oracle report the absolute L1 sum."""
    cell, faces = active_slices(P)
    gd = [P.gas(f)[cell] for f in range(6)]
    go = [O.gas[f][cell] for f in range(6)]
    cd, co = conserved(gd), conserved(go)
    out = {k: rel_l1(cd[k], co[k]) for k in cd}
    out["p"] = rel_l1(gd[4], go[4])
    out["p_S"] = rel_l1(gd[5], go[5])
    for f in range(3):
        out[B_NAMES[f]] = rel_l1(P.gem(f)[faces[f]], O.gem[f][faces[f]])
    return out


def exact_mismatches(P, O):
    """Number of active values (primitives and face B) that differ bitwise."""
    cell, faces = active_slices(P)
    n = 0
    for f in range(6):
        n += int(np.count_nonzero(P.gas(f)[cell] != O.gas[f][cell]))
    for f in range(3):
        n += int(np.count_nonzero(P.gem(f)[faces[f]] != O.gem[f][faces[f]]))
    return n
