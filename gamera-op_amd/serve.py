"""Serving loop for many independent MHD solves on one GPU.

Requests (a seeded problem, a grid size, a step count and a precision tag)
arrive over time.  They are admitted while the device-memory budget and the
slot count allow, their state stays on the device, and every admitted solve
advances in the same batched steps (one launch per stage for all of them):
continuous batching, where requests join and leave the batch between step
quanta.  Each member carries its own step budget on the device, so a quantum
has a fixed length: a member whose steps run out inside it stops there while
the others go on, and leaves at the end of the quantum.  The host set-up of
newly admitted requests runs while the GPU executes the current quantum.  A
solve's results do not depend on what else is in the batch (batched stepping
runs each grid's own arithmetic).  A request whose state turns non-finite is
evicted at the end of the quantum in which it does, alone.

The plain loop (`serve_plain`) is the batching-off baseline: one request at a
time, set up, stepped to completion and retired before the next starts.
"""
import collections
import hashlib
import multiprocessing as mp
import os
import sys
import time
from dataclasses import dataclass, field
from multiprocessing import shared_memory

import numpy as np

import gamera_op as go

CART_SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128)]
SPH_SIZES = [(64, 32, 64), (96, 48, 96), (128, 64, 128)]


@dataclass
class Request:
    rid: int
    arrival: float       # seconds after the start of the stream
    kind: int
    seed: int
    shape: tuple
    steps: int
    precision: int       # 0 FP64, 1 FP32-bulk
    poison: bool = False  # test hook: one initial density is NaN

    @property
    def cells(self):
        return self.shape[0] * self.shape[1] * self.shape[2]


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    cell_updates: int
    time_sim: float
    digest: str = ""
    status: str = "ok"   # "ok", or "nonfinite" (evicted)
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def make_requests(n, rate, seed=0, frac_cart=0.7, steps=(50, 500), cart_sizes=CART_SIZES,
                  sph_sizes=SPH_SIZES, frac_fp32=0.5):
    """Seeded Poisson arrivals at `rate` per second with the configured mix."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
    out = []
    for q in range(n):
        if rng.random() < frac_cart:
            kind, shape = go.SINUSOID_SPHERE, cart_sizes[rng.integers(len(cart_sizes))]
        else:
            kind, shape = go.DIPOLE_SHELL, sph_sizes[rng.integers(len(sph_sizes))]
        out.append(Request(q, float(t[q]), kind, int(rng.integers(1, 2**31)), tuple(shape),
                           int(rng.integers(steps[0], steps[1] + 1)),
                           1 if rng.random() < frac_fp32 else 0))
    return out


def device_bytes_estimate(shape, precision):
    """Device bytes gk_create allocates for a grid (state, history,
    intermediates; the per-axis tables are negligible)."""
    N = [n + 9 for n in shape]
    ka = 32 if precision == 1 else 16  # row pitch in whole 128-byte lines
    kp = (N[2] + ka - 1) // ka * ka
    fs = N[0] * N[1] * kp
    esz = 4 if precision == 1 else 8
    # FP64 state X (2 x 6), face and cell B (3 x 3 each); intermediates F, V, E;
    # the FP32-bulk form adds FP32 copies of X, B and cell B
    copies = 4 * fs * (12 + 9 + 9) if precision == 1 else 0
    return 8 * fs * (12 + 9 + 9) + esz * fs * (27 + 6 + 3) + copies


def state_digest_of(prim, bf):
    """SHA-256 of a state: 6 primitive and 3 face-B arrays (reference layout)."""
    h = hashlib.sha256()
    for a in list(prim) + list(bf):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def state_digest(solver):
    prim, bf = solver.state()
    return state_digest_of(prim, bf)


POISON_CELL = (4, 4, 4)  # first active cell


class _Setup:
    """Host set-up of one request at a time (the host facade holds one
    problem); consecutive requests of the same kind and grid reuse the grid,
    geometry and weights and only redo the seeded initial state."""

    def __init__(self):
        self.P = None
        self.key = None

    def solver(self, r, tune):
        key = (r.kind, r.shape)
        if self.P is not None and self.key == key:
            self.P.reseed(r.seed)
        else:
            if self.P is not None:
                self.P.close()
            self.P = go.Problem(r.kind, r.seed, *r.shape)
            self.key = key
        if r.poison:
            self.P.gas(0)[POISON_CELL] = np.nan
        S = go.Solver(self.P, r.precision)
        if r.poison:
            self.key = None  # the host state no longer matches the seed
        if tune:
            S.autotune(2)
        return S

    def close(self):
        if self.P is not None:
            self.P.close()
            self.P = None


# --------------------------------------------------------------------------
# Request preparation in worker processes.  The host facade holds one problem
# per process, so problem generation (grid, weights, seeded initial state)
# runs in a pool of processes; each hands back the grid descriptor blob and
# the initial state through shared memory, and the server only creates the
# device solver from them.
_W = {"P": None, "key": None}


def _worker_init(paths):
    # one thread per worker: the pool, not OpenMP, provides the parallelism
    os.environ["OMP_NUM_THREADS"] = "1"
    for p in paths:
        if p not in sys.path:
            sys.path.insert(0, p)
    # load the host libraries and run one tiny set-up now, so that the first
    # request a worker prepares does not pay for it
    import gamera_op as gw
    gw.Problem(gw.SINUSOID, 1, 8, 8, 8).close()


def _worker_prepare(kind, seed, shape):
    import gamera_op as gw
    key = (kind, tuple(shape))
    if _W["P"] is not None and _W["key"] == key:
        _W["P"].reseed(seed)
    else:
        if _W["P"] is not None:
            _W["P"].close()
        _W["P"] = gw.Problem(kind, seed, *shape)
        _W["key"] = key
    P = _W["P"]
    blob = P.desc_blob()
    fields = [P.gas(f) for f in range(6)] + [P.gem(f) for f in range(3)]
    nb = fields[0].nbytes
    shm = shared_memory.SharedMemory(create=True, size=9 * nb)
    for q, a in enumerate(fields):
        np.ndarray(a.shape, dtype=np.float64, buffer=shm.buf, offset=q * nb)[...] = a
    name = shm.name
    shm.close()
    return name, blob, fields[0].shape


class Preparer:
    """Pool of request-preparation processes.  Create it before the process
    touches the GPU (worker processes are started fresh, never forked from a
    process with a GPU context)."""

    def __init__(self, workers=None):
        n = workers or max(1, min(16, (os.cpu_count() or 2) - 1))
        ctx = mp.get_context("spawn")
        here = os.path.dirname(os.path.abspath(__file__))
        self.pool = ctx.Pool(n, initializer=_worker_init, initargs=([here],))
        self.workers = n

    def submit(self, r):
        return self.pool.apply_async(_worker_prepare, (r.kind, r.seed, r.shape))

    @staticmethod
    def solver(result, precision, poison=False, stream=None):
        name, blob, shape = result
        shm = shared_memory.SharedMemory(name=name)
        try:
            nb = int(np.prod(shape)) * 8
            arrs = [np.ndarray(shape, dtype=np.float64, buffer=shm.buf, offset=q * nb)
                    for q in range(9)]
            if poison:
                arrs[0][POISON_CELL] = np.nan
            S = go.Solver.from_parts(blob, arrs[:6], arrs[6:], precision, stream)
            del arrs
        finally:
            shm.close()
            shm.unlink()
        return S

    def close(self):
        self.pool.close()
        self.pool.join()


def tune_shapes(requests, trial_steps=2):
    """Autotune every (problem kind, grid, precision) of a request mix once,
    before serving: the picks are cached in the library per shape, so the
    serving loop only applies them and never times trial steps while a
    quantum runs (contended timings would be noise).  Returns seconds spent."""
    t0 = time.perf_counter()
    seen = set()
    for r in requests:
        key = (r.kind, tuple(r.shape), r.precision)
        if key in seen:
            continue
        seen.add(key)
        P = go.Problem(r.kind, r.seed, *r.shape)
        try:
            S = go.Solver(P, r.precision)
            S.autotune(trial_steps)
            S.close()
        finally:
            P.close()
    return time.perf_counter() - t0


def serve(requests, budget_bytes=64 << 30, max_slots=64, quantum=25, tune=True, digests=False,
          realtime=True, log=None, preparer=None, lookahead=None):
    """Run the request stream through the continuous-batching loop.  With a
    Preparer, request set-up runs in its worker processes (up to `lookahead`
    requests ahead); otherwise in this process, one request per quantum.
    With tune, admitted grids take the launch shapes cached for their shape
    (tune_shapes fills the cache; a shape without a pick keeps the defaults).
    Returns (results, wall seconds)."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    setup = _Setup() if preparer is None else None
    ahead = lookahead or (2 * preparer.workers if preparer is not None else 0)
    B = go.Batch()
    B.run_timing(True)  # device time of the quanta: the GPU-busy share
    side = B.side_stream()
    queue = collections.deque()      # arrived, not yet being prepared
    preparing = collections.deque()  # (request, async result) in arrival order
    active = []  # [request, solver, steps_left, admitted_time, bytes]
    pending = []
    leaving = []  # [member, status]: out of the batch, results still to read
    results = []
    used = 0
    nxt = 0
    in_flight = 0
    stats = collections.Counter()
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0

    def admit(r, S):
        nonlocal used
        if tune:
            S.autotune(0)  # the cached pick for this shape, if any; times nothing
        b = S.device_bytes()
        used += b
        a = [r, S, r.steps, now(), b]
        if r.steps <= 0:  # nothing to step: done at admission
            leaving.append((a, "ok"))
        else:
            pending.append(a)

    def retire(a, status):
        nonlocal used
        r, S = a[0], a[1]
        tr = time.perf_counter()
        sc = S.scalars()
        stats["retire_scalars_s"] += time.perf_counter() - tr
        done_steps = r.steps - max(a[2], 0)
        res = Result(r.rid, r.arrival, a[3], now(), done_steps, r.cells * done_steps,
                     sc["time"], status=status)
        if digests:
            res.digest = state_digest(S)
        results.append(res)
        tr = time.perf_counter()
        S.close()
        stats["retire_close_s"] += time.perf_counter() - tr
        used -= a[4]
        if log:
            log(f"{status} rid={r.rid} latency={res.latency:.3f}s active={len(active)}")

    try:
        while nxt < len(reqs) or queue or preparing or active or pending or leaving:
            tn = now()
            while nxt < len(reqs) and (reqs[nxt].arrival <= tn or not realtime):
                queue.append(reqs[nxt])
                nxt += 1
            if preparer is not None:
                while queue and len(preparing) < ahead:
                    r = queue.popleft()
                    preparing.append((r, preparer.submit(r)))
                # admit prepared requests in arrival order while memory and
                # slots allow (the device upload overlaps the quantum in flight)
                while preparing and preparing[0][1].ready() and \
                        len(active) + len(pending) < max_slots:
                    r = preparing[0][0]
                    need = device_bytes_estimate(r.shape, r.precision)
                    if used + need > budget_bytes and (active or pending):
                        break
                    _, res = preparing.popleft()
                    ta = time.perf_counter()
                    # set-up on the batch's side stream: it runs beside the
                    # quantum in flight instead of waiting behind it
                    admit(r, Preparer.solver(res.get(), r.precision, r.poison, side))
                    stats["admit_s"] += time.perf_counter() - ta
            else:
                # host set-up in this process: one request per pass while the
                # GPU is busy, as many as fit when it is idle
                while queue and len(active) + len(pending) < max_slots:
                    r = queue[0]
                    need = device_bytes_estimate(r.shape, r.precision)
                    if used + need > budget_bytes and (active or pending):
                        break
                    queue.popleft()
                    ta = time.perf_counter()
                    admit(r, setup.solver(r, False))
                    stats["admit_s"] += time.perf_counter() - ta
                    if in_flight:
                        break
            if in_flight:
                ta = time.perf_counter()
                bad = {id(S) for S in B.sync_status()}
                stats["gpu_wait_s"] += time.perf_counter() - ta
                in_flight = 0
                # members whose budget ran out, or whose state turned
                # non-finite, leave the batch before the next quantum
                for a in list(active):
                    a[2] = B.budget(a[1])
                    status = "nonfinite" if id(a[1]) in bad else ("ok" if a[2] == 0 else None)
                    if status:
                        B.remove(a[1])
                        active.remove(a)
                        leaving.append((a, status))
            for p in pending:
                B.add(p[1])
                B.set_budget(p[1], p[2])
                active.append(p)
            pending = []
            if active:
                ta = time.perf_counter()
                B.step_async(quantum)
                stats["launch_s"] += time.perf_counter() - ta
                stats["quanta"] += 1
                in_flight = quantum
            # read back the requests that left while the next quantum runs
            ta = time.perf_counter()
            for a, status in leaving:
                retire(a, status)
            if leaving:
                stats["retire_s"] += time.perf_counter() - ta
            leaving = []
            if in_flight:
                continue
            if preparing:
                ta = time.perf_counter()
                preparing[0][1].wait(0.05)
                stats["prep_wait_s"] += time.perf_counter() - ta
            elif nxt < len(reqs) and not queue:
                time.sleep(max(0.0, min(reqs[nxt].arrival - now(), 0.05)))
        wall = now()
        stats["gpu_busy_s"] = B.run_ms() / 1e3
    finally:
        B.close()
        for a in active + pending:
            a[1].close()
        for r, res in preparing:  # release prepared but unused state
            try:
                name = res.get()[0]
                shm = shared_memory.SharedMemory(name=name)
                shm.close()
                shm.unlink()
            except Exception:
                pass
        if setup is not None:
            setup.close()
    serve.last_stats = dict(stats)
    return results, wall


def serve_plain(requests, tune=True, digests=False, realtime=True, log=None):
    """Batching off: one request at a time, in arrival order.  With tune, each
    grid takes the launch shape cached for its shape (timed now if none is)."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    setup = _Setup()
    results = []
    t0 = time.perf_counter()
    try:
        for r in reqs:
            if realtime:
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            adm = time.perf_counter() - t0
            S = setup.solver(r, tune)
            status = "ok"
            try:
                S.step(max(r.steps, 0))
            except go.GkNonFinite:
                status = "nonfinite"
            sc = S.scalars()
            res = Result(r.rid, r.arrival, adm, time.perf_counter() - t0, max(r.steps, 0),
                         r.cells * max(r.steps, 0), sc["time"], status=status)
            if digests:
                res.digest = state_digest(S)
            results.append(res)
            S.close()
            if log:
                log(f"done rid={r.rid} latency={res.latency:.3f}s")
    finally:
        setup.close()
    return results, time.perf_counter() - t0


def summarize(results, wall):
    lat = np.array([r.latency for r in results])
    cu = sum(r.cell_updates for r in results)
    return {"requests": len(results), "wall_s": wall, "cell_updates": cu,
            "nonfinite": sum(1 for r in results if r.status != "ok"),
            "cell_updates_per_s": cu / wall if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "latency_max_s": float(lat.max()) if len(lat) else None}


def main(argv=None):
    """Command line: serve a seeded request stream (the configured mix) on one
    GPU and print one JSON line with throughput, latency percentiles and the
    loop's time breakdown; --plain also runs the batching-off loop on the same
    requests."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=5.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--budget-gib", type=int, default=64)
    ap.add_argument("--workers", type=int, default=0, help="preparation processes (0: auto)")
    ap.add_argument("--plain", action="store_true", help="also run the plain loop")
    ap.add_argument("--no-serve", action="store_true", help="run only the plain loop")
    ap.add_argument("--progress", type=int, default=50,
                    help="print a line to stderr every N finished requests")
    args = ap.parse_args(argv)

    reqs = make_requests(args.requests, args.rate, seed=args.seed)
    out = {"requests": len(reqs), "arrival_rate_per_s": args.rate, "seed": args.seed,
           "budget_GiB": args.budget_gib}
    done = [0]

    def progress(msg):
        done[0] += 1
        if args.progress and done[0] % args.progress == 0:
            print(f"serve: {done[0]} of {len(reqs)} finished ({msg})", file=sys.stderr, flush=True)

    if not args.no_serve:
        preparer = Preparer(args.workers or None)  # before this process touches the GPU
        try:
            tuning = tune_shapes(reqs)
            res, wall = serve(reqs, budget_bytes=args.budget_gib << 30, preparer=preparer,
                              log=progress)
        finally:
            preparer.close()
        out["served"] = summarize(res, wall)
        out["served"]["tuning_s_before_stream"] = round(tuning, 3)
        out["served"]["loop_seconds"] = {k: round(v, 3) if isinstance(v, float) else v
                                         for k, v in serve.last_stats.items()}
        out["served"]["gpu_busy_share"] = serve.last_stats.get("gpu_busy_s", 0.0) / wall
    if args.plain or args.no_serve:
        done[0] = 0
        res, wall = serve_plain(reqs, log=progress)
        out["plain_loop"] = summarize(res, wall)
    if "served" in out and "plain_loop" in out:
        out["throughput_ratio"] = (out["served"]["cell_updates_per_s"] /
                                   out["plain_loop"]["cell_updates_per_s"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
