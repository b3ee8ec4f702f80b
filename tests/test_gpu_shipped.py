"""The reference's own main program and problem files, run through the host
facade on the GPU, against the same program built from the reference's
sources (needs oracle/_ref: facade_main_<P> and ref_main_<P>, built by
oracle/Makefile target "shipped" where the reference sources are present).

For each shipped problem file both programs run with the same config.yaml
in directories of their own: the reference on the host cores, the drop-in
build on the GPU.  The config sets a dump interval of about a third of the
run and a time_stop inside it, so the run covers the initial dump, dumps
at the reference's cadence and the time_stop break.  Both must leave the
same dump files, and every file must hold bitwise the same state:

* /gas fields 0-5 (rho, v1..3, p, p_S) and 12-17 (the previous time level
  that a restart continues from), ghosts included;
* /gem fields 0-5 (face B, cell-centred B) and 9-11, 15-17 (their previous
  levels);
  each over the points of the grid: the arrays' upper slot on each axis
  (index n + 8) is a face index only along a face field's own axis and is
  left out elsewhere (FieldLoop_spherical's initial state writes it, the
  reference carries it through its array copies, and nothing computes it);
* /time_sim, /dt0, /config, /problem_config, /proc_coords;
* in the first file the geometry and weight datasets (background-field
  datasets with signed zeros normalised: a zero field's quadratures leave
  -0.0 where the facade writes +0.0).

Not compared: the half-step fields gas 6-11 and gem 6-8, 12-14, transient
values inside a step that the device never stores (the facade writes
zeros), and /log_seq_num.
"""
import glob
import os

import numpy as np
import pytest

import conftest
import h5read
from test_shipped_problems import SHIPPED, built, write_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
NT = 40

GAS = list(range(0, 6)) + list(range(12, 18))
GEM = list(range(0, 6)) + list(range(9, 12)) + list(range(15, 18))
SCALARS = ["/time_sim", "/dt0", "/config", "/problem_config", "/proc_coords"]
EXTRA = ["/geo", "/rec", "/x1", "/x2", "/x3", "/x1c", "/x2c", "/x3c", "/x1ctr", "/x2ctr",
         "/x3ctr", "/dx1", "/dx2", "/dx3"]
EXTRA_B0 = ["/B0_1", "/B0_2", "/B0_3", "/B0_i", "/B0_j", "/B0_k"]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def region(f, face_axis=None):
    """Points of a dumped field that belong to the grid: every index but the
    last on each axis (the upper slot, n + 8, is a face index only along a
    face field's own axis; for cell-centred fields it holds no cell, nothing
    computes it, and the reference leaves there whatever its array copies
    carried)."""
    sl = [slice(0, -1)] * 3
    if face_axis is not None:
        sl[face_axis] = slice(None)
    return f[tuple(sl)]


def where(x, y, n=4):
    """First few points whose bits differ, with both values."""
    d = np.argwhere(bits(x) != bits(y))
    return [(tuple(int(v) for v in p), float(x[tuple(p)]), float(y[tuple(p)])) for p in d[:n]]


def run(exe, d, threads=16):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    rc, out, err = conftest.RUNNER.run([os.path.join(REF, exe), "c.yaml"], cwd=d, env=env,
                                       timeout=300)
    assert rc == 0, (exe, rc, out, err)
    logs = sorted(glob.glob(os.path.join(d, "LOG_*.log")))
    text = open(logs[-1]).read() if logs else ""
    return text


@pytest.mark.gpu
@pytest.mark.timeout(600)
@pytest.mark.parametrize("name,n", SHIPPED)
def test_reference_main_through_facade_equals_reference(tmp_path, name, n):
    if conftest.RUNNER is None:
        pytest.skip("needs the program runner (-m gpu)")
    if not built(f"facade_main_{name}", f"ref_main_{name}") or not h5read.available():
        pytest.skip("needs oracle/_ref (reference sources) and the HDF5 runtime")
    print(f"shipped problem {name} {n}", flush=True)
    # the end time of NT steps, from the reference
    probe = tmp_path / "probe"
    probe.mkdir()
    write_config(str(probe), n, nt=NT)
    run(f"ref_main_{name}", str(probe))
    last = sorted(glob.glob(str(probe / "mhd_*.h5")))[-1]
    t_end = float(h5read.read(last, "/time_sim")[0])
    assert t_end > 0.0
    # dumps at about a third of the run, time_stop at 80 % of it
    dirs = {}
    for who in ("ref", "facade"):
        d = tmp_path / who
        d.mkdir()
        write_config(str(d), n, nt=NT, time_stop=0.8 * t_end, output_interval=t_end / 3.3)
        dirs[who] = str(d)
    run(f"ref_main_{name}", dirs["ref"])
    log = run(f"facade_main_{name}", dirs["facade"])
    assert "ERROR" not in log, log[-3000:]
    files = {w: sorted(os.path.basename(f) for f in glob.glob(os.path.join(d, "mhd_*.h5")))
             for w, d in dirs.items()}
    print("dumps", files["ref"], flush=True)
    assert files["ref"] == files["facade"]
    assert len(files["ref"]) >= 3  # initial, at least one interval, final
    bad = []
    for q, fname in enumerate(files["ref"]):
        R = os.path.join(dirs["ref"], fname)
        F = os.path.join(dirs["facade"], fname)
        gr, gf = h5read.read(R, "/gas")[0], h5read.read(F, "/gas")[0]
        for f in GAS:
            x, y = region(gr[f]), region(gf[f])
            if not np.array_equal(bits(x), bits(y)):
                bad.append((fname, "gas", f, int((bits(x) != bits(y)).sum()), where(x, y)))
        mr, mf = h5read.read(R, "/gem"), h5read.read(F, "/gem")
        for f in GEM:
            axis = f % 3 if f in (0, 1, 2, 9, 10, 11) else None  # face B, its previous level
            x, y = region(mr[f], axis), region(mf[f], axis)
            if not np.array_equal(bits(x), bits(y)):
                bad.append((fname, "gem", f, int((bits(x) != bits(y)).sum()), where(x, y)))
        for ds in SCALARS:
            if not np.array_equal(bits(h5read.read(R, ds)), bits(h5read.read(F, ds))):
                bad.append((fname, ds))
        if q == 0:
            for ds in EXTRA:
                if not np.array_equal(bits(h5read.read(R, ds)), bits(h5read.read(F, ds))):
                    bad.append((fname, ds))
            for ds in EXTRA_B0:
                if not np.array_equal(bits(h5read.read(R, ds) + 0.0),
                                      bits(h5read.read(F, ds) + 0.0)):
                    bad.append((fname, ds))
        print(fname, "time", float(h5read.read(R, "/time_sim")[0]), flush=True)
    for b in bad:
        print("mismatch", b, flush=True)
    assert not bad, bad[:4]
    # the run ended at the time_stop break, before NT steps
    t_last = float(h5read.read(os.path.join(dirs["ref"], files["ref"][-1]), "/time_sim")[0])
    assert 0.8 * t_end < t_last < t_end
