"""Run the compiled reference solver (oracle/_ref) on a seeded problem.

Only usable where oracle/_ref was built (this needs the reference sources and
the image's MPICH/HDF5); fixtures generated with it are committed under
tests/golden/ so that nothing else needs the reference at run time.
"""
import os
import subprocess
import tempfile

from refbin import read_records

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

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


def ref_available(spherical=False):
    exe = os.path.join(REF_DIR, "ref_driver_sph" if spherical else "ref_driver_cart")
    return os.path.exists(exe)


def run_reference(kind, seed, ni, nj, nk, nt, geometry=False, threads=8, workdir=None):
    """Run the reference for nt steps; returns the record dict.  With workdir
    the run happens there and its files (HDF5 dumps, logs) are kept."""
    spherical = kind in (2, 4, 5, 11)
    exe = os.path.join(REF_DIR, "ref_driver_sph" if spherical else "ref_driver_cart")
    with tempfile.TemporaryDirectory(prefix="gkref") as tmp:
        d = workdir or tmp
        with open(os.path.join(d, "c.yaml"), "w") as f:
            f.write(YAML.format(ni=ni, nj=nj, nk=nk, nt=nt))
        env = dict(os.environ, GK_PROBLEM=str(kind), GK_SEED=str(seed),
                   OMP_NUM_THREADS=str(threads))
        out = os.path.join(d, "out.bin")
        r = subprocess.run([exe, "c.yaml", out, "1" if geometry else "0"], cwd=d,
                           env=env, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed ({r.returncode}): {r.stdout}\n{r.stderr}")
        return read_records(out)
