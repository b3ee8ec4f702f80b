"""Generate the golden fixtures from the compiled reference solver.

Run where oracle/_ref is built (needs the reference sources; see
oracle/Makefile):  python tests/golden/make_fixtures.py

Writes
  ref_states.npz     final active state (primitives, face B) and scalars of a
                     few small seeded runs
  ref_hashes.json    SHA-256 of the reference's geometry, weight and
                     background-field arrays and of its post-prologue state,
                     per case; and of the config-0 run (128x128x8, 200 steps)
The fixtures are data only (inputs are regenerated from the seeds by the
problem plugin); nothing here is reference source.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from refrun import run_reference  # noqa: E402

# kind, seed, ni, nj, nk, steps
STATE_CASES = [
    (0, 1, 16, 16, 8, 10),
    (0, 2, 24, 20, 12, 30),
    (1, 3, 16, 16, 16, 20),
    (2, 4, 16, 12, 16, 20),
    (3, 5, 24, 16, 8, 40),
    # full sphere with pole ring averaging (2 and 4 averaged rows per pole)
    (4, 8, 12, 16, 32, 12),
    (4, 9, 8, 12, 64, 8),
    # northern hemisphere, pole boundary at theta = 0 only, fixed ghosts
    # elsewhere (the boundary layout of the reference's field-loop problem)
    (5, 13, 12, 12, 32, 10),
    (5, 14, 8, 10, 64, 6),
    # reflective walls in x, extrapolated boundaries in y, periodic z
    (6, 15, 16, 12, 8, 12),
    # fixed boundary at x = 0, extrapolated at x = 1, symmetric in y
    (7, 22, 16, 12, 8, 12),
    # entropy channel carried (p_S = p at the start)
    (8, 32, 16, 16, 8, 10),
    # positivity floors crossed (both branches), non-default floors, gravity
    (9, 31, 24, 20, 12, 20),
    # uniform background field whose face quadratures differ in the last
    # bits from face to face (widths 1/32, 1/24, 1/16)
    (1, 15, 32, 24, 16, 12),
    # background field varying in space: sinusoidal box, dipole shell
    (10, 41, 16, 12, 16, 12),
    (11, 42, 16, 12, 16, 10),
]
GEOM_CASES = [(0, 1, 16, 16, 8), (1, 3, 16, 16, 16), (2, 4, 16, 12, 16), (0, 7, 40, 24, 12),
              (4, 8, 12, 16, 32), (10, 41, 16, 12, 16), (11, 42, 16, 12, 16)]
CONFIG0 = (0, 1, 128, 128, 8, 200)

COORDS = ["x1", "x2", "x3", "x1c", "x2c", "x3c", "x1ctr", "x2ctr", "x3ctr", "dx1", "dx2", "dx3"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def active_state(r, ni, nj, nk):
    sh = (ni + 9, nj + 9, nk + 9)
    g = r["gas"].reshape((18,) + sh)
    m = r["gem"].reshape((18,) + sh)
    a, b, c = slice(4, 4 + ni), slice(4, 4 + nj), slice(4, 4 + nk)
    prim = [g[f][a, b, c] for f in range(6)]
    faces = [m[0][4:5 + ni, b, c], m[1][a, 4:5 + nj, c], m[2][a, b, 4:5 + nk]]
    return prim, faces


def main():
    states = {}
    for kind, seed, ni, nj, nk, steps in STATE_CASES:
        print("state case", kind, seed, ni, nj, nk, steps, flush=True)
        r = run_reference(kind, seed, ni, nj, nk, steps)
        prim, faces = active_state(r, ni, nj, nk)
        key = f"k{kind}_s{seed}_{ni}x{nj}x{nk}_n{steps}"
        for f in range(6):
            states[f"{key}/prim{f}"] = prim[f]
        for f in range(3):
            states[f"{key}/B{f}"] = faces[f]
        states[f"{key}/scalars"] = np.array([r["dt"], r["dt0"], r["time_sim"], r["divB_max"]])
    np.savez_compressed(os.path.join(HERE, "ref_states.npz"), **states)

    hashes = {"geometry": {}, "config0": {}}
    for kind, seed, ni, nj, nk in GEOM_CASES:
        print("geometry case", kind, seed, ni, nj, nk, flush=True)
        r = run_reference(kind, seed, ni, nj, nk, 0, geometry=True)
        sh = (ni + 9, nj + 9, nk + 9)
        h = {}
        geo = r["geo"].reshape((28,) + sh)
        rec = r["rec"].reshape((42,) + sh)
        b0 = r["B0"].reshape((39,) + sh)
        for f in range(28):
            h[f"geo{f}"] = sha(geo[f])
        for f in range(42):
            h[f"rec{f}"] = sha(rec[f])
        for f in range(39):
            # +0.0 normalises signed zeros: without a background field the
            # reference integrates the zero function and may store -0.0
            h[f"gem{52 + f}"] = sha(b0[f] + 0.0)
        for name in COORDS:
            h[name] = sha(r[name].reshape(sh))
        prim, faces = active_state(r, ni, nj, nk)
        for f in range(6):
            h[f"init_prim{f}"] = sha(prim[f])
        for f in range(3):
            h[f"init_B{f}"] = sha(faces[f])
        h["dt0"] = float(r["dt0"])
        hashes["geometry"][f"k{kind}_s{seed}_{ni}x{nj}x{nk}"] = h

    kind, seed, ni, nj, nk, steps = CONFIG0
    print("config 0", CONFIG0, flush=True)
    r = run_reference(kind, seed, ni, nj, nk, steps)
    prim, faces = active_state(r, ni, nj, nk)
    c0 = {"case": list(CONFIG0)}
    for f in range(6):
        c0[f"prim{f}"] = sha(prim[f])
        c0[f"prim{f}_l1"] = float(np.abs(prim[f]).sum())
    for f in range(3):
        c0[f"B{f}"] = sha(faces[f])
        c0[f"B{f}_l1"] = float(np.abs(faces[f]).sum())
    c0.update(dt=float(r["dt"]), dt0=float(r["dt0"]), time_sim=float(r["time_sim"]),
              divB_max=float(r["divB_max"]))
    hashes["config0"] = c0
    with open(os.path.join(HERE, "ref_hashes.json"), "w") as fp:
        json.dump(hashes, fp, indent=1, sort_keys=True)
    print("done")


if __name__ == "__main__":
    main()
