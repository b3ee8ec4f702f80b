"""HDF5 dumps and restart of the host facade against the reference's own
files (CPU; needs the compiled reference under oracle/_ref and the HDF5
layer libgamera_h5.so).

The reference run leaves its per-rank dumps (initial state and final state)
in a scratch directory.  The facade's main() sequence started there must
restart from the last dump and hold exactly the reference's final arrays;
a dump the facade writes from them must hold the same datasets (h5diff)."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import gamera_op as go

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB = os.path.join(os.path.dirname(go.LIB_HOST), "libgamera_h5.so")
H5DIFF = "/opt/conda/bin/h5diff"

try:
    from refrun import ref_available, run_reference
except ImportError:  # pragma: no cover
    ref_available = None

needs = pytest.mark.skipif(
    ref_available is None or not ref_available() or not os.path.exists(H5LIB),
    reason="needs oracle/_ref (reference sources) and the HDF5 layer")


@needs
@pytest.mark.parametrize("kind,seed,n,steps", [(0, 3, (16, 16, 8), 5),
                                               (4, 12, (12, 16, 32), 4)])
def test_restart_reads_reference_dump_and_rewrites_it(tmp_path, monkeypatch, kind, seed, n,
                                                      steps):
    """Cartesian, and a full sphere with pole ring averaging (whose dumps
    also carry the ring face areas, /geo_ring)."""
    rec = run_reference(kind, seed, *n, steps, workdir=str(tmp_path))
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".h5"))
    assert files == ["mhd_00-00-00_000000.h5", "mhd_00-00-00_000001.h5"], files
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("GK_PROBLEM", str(kind))
    monkeypatch.setenv("GK_SEED", str(seed))
    H = go.run_main("c.yaml", 0)  # restart only, no steps
    try:
        assert ctypes.c_bool.in_dll(go._load()[0], "read_restart").value
        shape = H.shape
        g = rec["gas"].reshape((18,) + shape)
        m = rec["gem"].reshape((18,) + shape)
        for f in range(18):
            assert np.array_equal(H.gas(f), g[f]), ("gas", f)
            assert np.array_equal(H.gem(f), m[f]), ("gem", f)
        assert H.scalar("time_sim") == rec["time_sim"]
        assert H.scalar("dt0") == rec["dt0"]
        assert H.int("hdf_seq_num") == 1
        # write it back under another name and compare every dataset
        h5 = ctypes.CDLL(H5LIB)
        assert h5.dump_data_hdf5(b"ours") == 0
        ours = "ours_00-00-00_000002.h5"
        assert os.path.exists(ours)
        if os.path.exists(H5DIFF):
            for ds in ["/gas", "/gem", "/x1", "/x2", "/x3", "/time_sim", "/dt0", "/config",
                       "/problem_config", "/norm_config", "/proc_coords", "/x1min_global",
                       "/x1max_global", "/x2min_global", "/x2max_global", "/x3min_global",
                       "/x3max_global"]:
                r = subprocess.run([H5DIFF, "mhd_00-00-00_000001.h5", ours, ds, ds],
                                   capture_output=True, text=True)
                assert r.returncode == 0, (ds, r.stdout[-2000:])
            # geometry datasets of the first dump (dump_extra_data)
            assert h5.dump_extra_data(b"ours") == 0
            extra = ["/geo", "/rec"] + (["/geo_ring"] if kind == 4 else [])
            for ds in extra:
                r = subprocess.run([H5DIFF, "mhd_00-00-00_000000.h5", ours, ds, ds],
                                   capture_output=True, text=True)
                assert r.returncode == 0, (ds, r.stdout[-2000:])
    finally:
        go.finalize()


@needs
def test_reference_restarts_from_facade_dump(tmp_path, monkeypatch):
    """A dump written by the facade is a valid restart file for the
    reference: continuing there equals the reference's straight run."""
    kind, seed, n = 0, 6, (16, 16, 8)
    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    run_reference(kind, seed, *n, 4, workdir=str(a))
    straight = run_reference(kind, seed, *n, 7, workdir=str(b))
    # facade: restart from the reference's step-4 dump, rewrite it as the
    # only dump of a fresh directory, and let the reference continue 3 steps
    monkeypatch.chdir(a)
    monkeypatch.setenv("GK_PROBLEM", str(kind))
    monkeypatch.setenv("GK_SEED", str(seed))
    go.run_main("c.yaml", 0)
    c = tmp_path / "c"
    c.mkdir()
    try:
        h5 = ctypes.CDLL(H5LIB)
        monkeypatch.chdir(c)
        ctypes.c_int.in_dll(go._load()[0], "hdf_seq_num").value = 0
        assert h5.dump_data_hdf5(b"mhd") == 0
    finally:
        go.finalize()
    cont = run_reference(kind, seed, *n, 3, workdir=str(c))
    for key in ("gas", "gem"):
        s = straight[key].reshape((18, -1))
        t = cont[key].reshape((18, -1))
        for f in (0, 1, 2, 3, 4, 5) if key == "gas" else (0, 1, 2):
            assert np.array_equal(s[f], t[f]), (key, f)
    assert cont["time_sim"] == straight["time_sim"]
