import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "tests"), os.path.join(ROOT, "tests", "golden"),
          os.path.join(ROOT, "gamera-op_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


PREPARER = None
RUNNER = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP device solver)")
    # The serving tests prepare requests in worker processes.  Start them now,
    # before any test touches the GPU: workers are fresh processes and must
    # never be started from a process that holds a GPU context.
    expr = getattr(config.option, "markexpr", "") or ""
    if "gpu" in expr and "not gpu" not in expr:
        global PREPARER, RUNNER
        import serve
        PREPARER = serve.Preparer(workers=4)
        # programs that GPU tests start (reference runs, drop-in builds)
        import runner
        RUNNER = runner.Runner()


def pytest_unconfigure(config):
    global PREPARER, RUNNER
    if PREPARER is not None:
        PREPARER.close()
        PREPARER = None
    if RUNNER is not None:
        RUNNER.close()
        RUNNER = None
