"""The reference's own problem files and main program against the host facade
(CPU; needs oracle/_ref built from the reference sources, so it runs only
where they are present).

The five shipped problem files (src/problem_test_OT_cartesion.c,
problem_FieldLoop_cartesian.c, problem_FieldLoop_spherical.c,
problem_test_Blast_cartesion.c, problem_test_Blast_spherical.c) and the
reference's src/main.c are copied out of the reference tree and compiled,
unmodified, against gamera-op_amd/host/include and linked with
libgamera_host.so (oracle/Makefile, target "shipped": facade_main_<P>).
That they build at all is the first check.  The second runs the set-up half
of main() (config file, solver set-up, weights, the problem's initial state,
background field, ring-average set-up) through the facade
(facade_setup_<P>) and through the compiled reference (ref_setup_<P>), and
requires the host arrays to be identical: the active primitives and face B
the problem file writes, every geometry and reconstruction-weight field,
the face coordinates and the background-field arrays.  The GPU half, the
whole main() run compared dump by dump, is tests/test_gpu_shipped.py.
"""
import os
import subprocess

import numpy as np
import pytest

from refbin import read_records

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

# problem (as in -DPROBLEM_<name>), active cells ni nj nk
SHIPPED = [
    ("OT_cartesian", (32, 32, 8)),
    ("FieldLoop_cartesian", (16, 16, 16)),
    ("FieldLoop_spherical", (16, 16, 32)),
    ("Blast_cartesian", (16, 16, 16)),
    ("Blast_spherical", (16, 12, 16)),
]

YAML = """ni_global: {ni}
nj_global: {nj}
nk_global: {nk}
proc_dims_i: 1
proc_dims_j: 1
proc_dims_k: 1
Nt: {nt}
time_stop: {time_stop}
output_interval: {output_interval}
x_Norm: 1.0
u_Norm: 1.0
rho_Norm: 1.0
"""


def write_config(d, n, nt=0, time_stop=1.0e30, output_interval=1.0e30):
    with open(os.path.join(d, "c.yaml"), "w") as f:
        f.write(YAML.format(ni=n[0], nj=n[1], nk=n[2], nt=nt, time_stop=time_stop,
                            output_interval=output_interval))


def built(*names):
    return all(os.path.exists(os.path.join(REF, x)) for x in names)


def run(exe, args, cwd, threads=4):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([os.path.join(REF, exe)] + args, cwd=cwd, env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, (exe, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r


@pytest.mark.parametrize("name,n", SHIPPED)
def test_reference_main_and_problem_build_against_facade(name, n):
    if not built(f"facade_main_{name}"):
        pytest.skip("needs oracle/_ref (reference sources)")
    exe = os.path.join(REF, f"facade_main_{name}")
    # linked with the facade, not with any reference solver object
    syms = subprocess.run(["nm", "-D", "--defined-only", exe], capture_output=True,
                          text=True).stdout
    for own in ("set_problem_config", "problem_init", "boundary_conditions"):
        assert own in subprocess.run(["nm", exe], capture_output=True, text=True).stdout, own
    for facade_only in ("solve", "initialize_solver", "getweights"):
        assert f" T {facade_only}\n" not in syms, facade_only
    needed = subprocess.run(["readelf", "-d", exe], capture_output=True, text=True).stdout
    assert "libgamera_host.so" in needed


@pytest.mark.parametrize("name,n", SHIPPED)
def test_setup_through_facade_equals_reference(tmp_path, name, n):
    if not built(f"facade_setup_{name}", f"ref_setup_{name}"):
        pytest.skip("needs oracle/_ref (reference sources)")
    ref_dir, fac_dir = tmp_path / "ref", tmp_path / "facade"
    ref_dir.mkdir()
    fac_dir.mkdir()
    write_config(str(ref_dir), n)
    write_config(str(fac_dir), n)
    # the reference: its set-up, then solve(0) (the prologue fills ghosts and
    # derived fields only); geometry records on
    run(f"ref_setup_{name}", ["c.yaml", "out.bin", "1"], str(ref_dir))
    run(f"facade_setup_{name}", ["c.yaml", "out.bin"], str(fac_dir))
    R = read_records(str(ref_dir / "out.bin"))
    F = read_records(str(fac_dir / "out.bin"))
    ni, nj, nk = n
    shape = (ni + 9, nj + 9, nk + 9)
    n3 = shape[0] * shape[1] * shape[2]
    assert (int(R["ni"]), int(R["nj"]), int(R["nk"])) == n
    assert (int(F["ni"]), int(F["nj"]), int(F["nk"])) == n
    assert R["CA"] == F["CA"]
    cell = (slice(4, 4 + ni), slice(4, 4 + nj), slice(4, 4 + nk))
    faces = [(slice(4, 5 + ni), cell[1], cell[2]), (cell[0], slice(4, 5 + nj), cell[2]),
             (cell[0], cell[1], slice(4, 5 + nk))]
    gas = R["gas"].reshape((-1,) + shape)
    gem = R["gem"].reshape((-1,) + shape)

    def same(a, b):
        # bit patterns: the weight arrays hold NaN outside the faces the
        # reference computes them for, and those must match too
        return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                              np.ascontiguousarray(b).view(np.uint64))

    bad = []
    for f in range(6):
        if not same(F[f"gas{f}"].reshape(shape)[cell], gas[f][cell]):
            bad.append(f"gas{f}")
    for f in range(3):
        if not same(F[f"gem{f}"].reshape(shape)[faces[f]], gem[f][faces[f]]):
            bad.append(f"gem{f}")
    geo = R["geo"].reshape((-1,) + shape)
    rec = R["rec"].reshape((-1,) + shape)
    for f in range(geo.shape[0]):
        if not same(F[f"geo{f}"].reshape(shape), geo[f]):
            bad.append(f"geo{f}")
    for f in range(rec.shape[0]):
        if not same(F[f"rec{f}"].reshape(shape), rec[f]):
            bad.append(f"rec{f}")
    for x in ("x1", "x2", "x3"):
        if not same(F[x], R[x]):
            bad.append(x)
    # background-field arrays: for a zero field the reference's quadratures
    # leave -0.0 where the facade, which skips them, holds +0.0; the values
    # only ever enter sums, so the comparison is of values with zeros
    # normalised
    b0 = R["B0"].reshape((-1,) + shape)
    for q in range(b0.shape[0]):
        if not same(F[f"gem{52 + q}"].reshape(shape) + 0.0, b0[q] + 0.0):
            bad.append(f"gem{52 + q}")
    assert F["gas0"].size == n3
    assert not bad, bad
    # a non-trivial initial state (the problem file ran)
    assert np.abs(gem[0][faces[0]]).max() + np.abs(gem[1][faces[1]]).max() > 0.0
