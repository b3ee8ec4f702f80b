"""Dump and restart through the reference's main() sequence on the device:
N1 steps with HDF5 dumps, then a new start in the same directory restarts
from the last dump and runs N2 more steps.  The result must equal a
straight N1 + N2 step run bit for bit (both time levels travel in the file,
as in the reference)."""
import os

import numpy as np
import pytest

import gamera_op as go

H5LIB = os.path.join(os.path.dirname(go.LIB_HOST), "libgamera_h5.so")

YAML = """ni_global: {ni}
nj_global: {nj}
nk_global: {nk}
proc_dims_i: 1
proc_dims_j: 1
proc_dims_k: 1
Nt: {nt}
time_stop: 1.0e30
output_interval: 1.0e30
x_Norm: 1.0
u_Norm: 1.0
rho_Norm: 1.0
"""


def run(d, kind, seed, n, nt, monkeypatch):
    monkeypatch.chdir(d)
    monkeypatch.setenv("GK_PROBLEM", str(kind))
    monkeypatch.setenv("GK_SEED", str(seed))
    (d / "c.yaml").write_text(YAML.format(ni=n[0], nj=n[1], nk=n[2], nt=nt))
    H = go.run_main("c.yaml", nt)
    out = [H.gas(f).copy() for f in range(6)] + [H.gem(f).copy() for f in range(3)]
    t = H.scalar("time_sim")
    go.finalize()
    return out, t


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(H5LIB), reason="HDF5 layer not built")
@pytest.mark.parametrize("kind,seed,n", [(1, 8, (16, 16, 16)), (2, 9, (16, 12, 16))])
def test_restart_continues_bitwise(tmp_path, monkeypatch, kind, seed, n):
    print(f"restart kind={kind} seed={seed} {n}", flush=True)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run(a, kind, seed, n, 6, monkeypatch)
    assert sorted(f for f in os.listdir(a) if f.endswith(".h5")) == [
        "mhd_00-00-00_000000.h5", "mhd_00-00-00_000001.h5"]
    cont, tc = run(a, kind, seed, n, 5, monkeypatch)   # restarts from _000001
    straight, ts = run(b, kind, seed, n, 11, monkeypatch)
    assert tc == ts
    for q, (x, y) in enumerate(zip(cont, straight)):
        assert np.array_equal(x, y), q
