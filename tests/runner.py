"""A process that starts programs for GPU tests.

A process that holds a GPU context must not start other programs itself, so
GPU tests hand the command to a worker of this pool, which is spawned fresh
before the test process first touches the GPU (tests/conftest.py)."""
import multiprocessing as mp
import subprocess


def _run(cmd, cwd, env, timeout):
    r = subprocess.run(cmd, cwd=cwd, env=env, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout[-4000:], r.stderr[-4000:]


class Runner:
    def __init__(self):
        self.pool = mp.get_context("spawn").Pool(1)
        self.pool.apply(_run, (["true"], None, None, 60))  # started now

    def run(self, cmd, cwd=None, env=None, timeout=600):
        """(returncode, stdout tail, stderr tail) of cmd run in the worker."""
        return self.pool.apply(_run, (cmd, cwd, env, timeout))

    def close(self):
        self.pool.close()
        self.pool.join()
