"""Config 0 on the device: 128x128x8 sinusoid, 200 AB2 steps at CFL 0.3,
FP64, bit for bit against the SHA-256 of the reference CPU run (golden
fixture written by tests/golden/make_fixtures.py)."""
import hashlib
import json
import os

import numpy as np
import pytest

import gamera_op as go
from parity import active_slices

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HASHES = json.load(open(os.path.join(GOLD, "ref_hashes.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.gpu
def test_config0_device_matches_reference_hash():
    h = HASHES["config0"]
    kind, seed, ni, nj, nk, steps = h["case"]
    print(f"config0 {ni}x{nj}x{nk} steps={steps}", flush=True)
    P = go.Problem(kind, seed, ni, nj, nk)
    try:
        P.device_open(0)
        P.device_step(steps)
        P.device_fetch()
        sc = P.device_scalars()
        cell, faces = active_slices(P)
        l1 = {}
        for f in range(6):
            l1[f"prim{f}"] = abs(float(np.abs(P.gas(f)[cell]).sum()) - h[f"prim{f}_l1"]) / max(h[f"prim{f}_l1"], 1e-300)
        for f in range(3):
            l1[f"B{f}"] = abs(float(np.abs(P.gem(f)[faces[f]]).sum()) - h[f"B{f}_l1"]) / h[f"B{f}_l1"]
        print("relative L1-norm differences", l1, "time", sc["time"], flush=True)
        for f in range(6):
            assert sha(P.gas(f)[cell]) == h[f"prim{f}"], f
        for f in range(3):
            assert sha(P.gem(f)[faces[f]]) == h[f"B{f}"], f
        assert sc["time"] == h["time_sim"]
        assert sc["dt"] == h["dt"]
        assert sc["divB"] < 1e-12
    finally:
        P.close()
