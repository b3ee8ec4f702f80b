#!/usr/bin/env python3
"""Throughput of the per-step ideal-MHD update on one MI355X.

Workload (default): Cartesian periodic box [0,1]^3 at 256^3 cells, seeded
sinusoidal velocity and edge-potential field, uniform |B0| = 1 background in
a seeded direction and an over-pressure sphere (R = 0.1, p ratio 100); AB2
steps at CFL 0.3, FP64 arithmetic.  One "step" is one full AB2 step over the
grid (all stage kernels, dt reduction included); inputs are resident in HBM
when the timed region starts.

Prints ONE JSON line with the measured cell-updates/s, the roofline of the
dominant kernel (HIP events on the solver's stream over the timed region) and
a CPU baseline timed on this host's cores on bounded samples: the compiled
reference when oracle/_ref holds it, with the CPU restatement beside it.
`host_buffers` gives the rate of the same steps with the device set-up,
the upload of the host state and the read-back included (the facade's
solve(nt) path, state in host arrays); it is never `value`.
Further legs: the FP32-bulk form, a full sphere with pole ring averaging, a
batched ensemble and a served request stream (each can be switched off).

  python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu] [--no-fp32]
                  [--no-ring | --ring-only] [--no-ensemble] [--no-serve]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "gamera-op_amd")
HBM_PEAK_GBS = 8000.0
GAMMA = 5.0 / 3.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_alg_bytes(name, n):
    """Compulsory HBM bytes of one launch of a stage kernel on an n^3 grid
    (each distinct input value read once, each output written once, 8 B per
    value).  Counts follow the kernel's own inputs and outputs."""
    N = n + 9
    full_faces = (n + 1) * n * n
    if name.startswith("k_sweep_"):
        # full faces: X^n, X^{n-1} (6+6), cell B at n, n-1 (3+3), face-normal
        # B (1) in; 9 fluxes + 2 face velocities out.  The transverse
        # extension (8 extra rows) only forms the 2 face velocities: 2+2 in,
        # 2 out.
        nf = n + 1
        if name == "k_sweep_k":
            # whole 64-face tiles; a remainder under 32 faces per row runs in
            # the direct sweep (k_krem_k)
            rem = nf % 64
            if nf >= 64 and 0 < rem < 32:
                nf -= rem
        full = nf * n * n
        ext_faces = nf * (N - 1 - n) * n
        return 8 * (full * 30 + ext_faces * 6)
    if name.startswith("k_emf_"):
        # per edge: 2 face velocities, 2 face-B (n, n-1 each), density at n, n-1
        # in, 1 EMF out
        return 8 * n * (n + 1) * (n + 1) * 9
    if name == "k_ct":
        # 3 face-B in, 3 EMF in, 3 face-B out
        return 8 * 9 * full_faces
    if name == "k_cell":
        # 27 face fluxes, X^n (6), cell B^n (3), face B^{n+1} (3) in; X^{n+1} (6)
        # and cell B^{n+1} (3) out
        return 8 * n ** 3 * 48
    if name == "k_ghost":
        return 8 * 2 * 9 * (N ** 3 - n ** 3)
    if name == "k_mags":
        # cell B of the points outside the active cells: 3 face fields in, 3 out
        return 8 * 6 * ((N - 1) ** 3 - n ** 3)
    return 0


BREAKDOWN_STEPS = 10


def run_form(args, precision):
    """Time one form of the step on the default workload; returns wall time,
    per-kernel HIP-event averages, device scalars, the final active state,
    set-up/read-back times, the dominant stage and its HIP-event average over
    the timed steps.  The per-stage breakdown comes from its own
    BREAKDOWN_STEPS steps before the timed ones, with events around every
    launch (about 0.12 ms per FP64 step at 256^3); the timed steps carry
    events around the dominant stage's launches only."""
    sys.path.insert(0, PKG)
    import numpy as np
    import gamera_op as go

    n = args.n
    form = "FP64" if precision == 0 else "FP32-bulk"
    log(f"bench: {form}: host setup {n}^3 (sinusoid + uniform B0 + sphere, seed {args.seed})")
    t0 = time.time()
    P = go.Problem(go.SINUSOID_BLAST, args.seed, n, n, n)
    log(f"bench: host setup {time.time() - t0:.1f} s; device open")
    t0 = time.perf_counter()
    P.device_open(precision)  # device tables, state upload from host arrays, prologue
    t_open = time.perf_counter() - t0
    if args.chunk:
        P.set_tuning(go.TUNE_CHUNK, args.chunk)
    log(f"bench: device holds {P.device_bytes() / 2**30:.2f} GiB; warmup {args.warmup} steps")
    P.device_step(args.warmup)
    P.enable_timing(True)
    P.device_step(BREAKDOWN_STEPS)
    kt = P.kernel_times()  # average ms per launch, every stage
    # dominant kernel: largest time per step (each kernel launches once per step)
    dom = max(kt, key=lambda k: kt[k])
    P.time_stage(dom)
    log(f"bench: {form}: timing {args.steps} steps (events on {dom})")
    t0 = time.perf_counter()
    P.device_step(args.steps)  # returns after the solver's stream has drained
    el = time.perf_counter() - t0
    dom_ms = P.kernel_times()[dom]  # its average over the timed steps
    P.enable_timing(False)
    sc = P.device_scalars()
    log(f"bench: {form}: {el * 1e3 / args.steps:.3f} ms/step; scalars {sc}")
    if sc["nonfinite"]:
        raise RuntimeError("non-finite values in the solution")
    t0 = time.perf_counter()
    P.device_fetch()  # both time levels back into the host arrays
    t_fetch = time.perf_counter() - t0
    a, b, c = slice(4, 4 + n), slice(4, 4 + n), slice(4, 4 + n)
    state = {"rho": P.gas(0)[a, b, c].copy(), "p": P.gas(4)[a, b, c].copy(),
             "rhov1": (P.gas(0)[a, b, c] * P.gas(1)[a, b, c]),
             "Bi": P.gem(0)[4:5 + n, b, c].copy(), "Bj": P.gem(1)[a, 4:5 + n, c].copy(),
             "Bk": P.gem(2)[a, b, 4:5 + n].copy()}
    P.close()
    return el, kt, sc, state, (t_open, t_fetch), dom, dom_ms


def rel_l1(x, y):
    import numpy as np
    return float(np.abs(x - y).sum() / np.abs(y).sum())


def run_gpu(args):
    sys.path.insert(0, PKG)
    import gamera_op as go

    el, kt, sc, st64, io, dom, dom_ms = run_form(args, 0)
    fp32 = None
    if not args.no_fp32:
        el32, kt32, sc32, st32, _, _, _ = run_form(args, 1)
        cells = args.n ** 3
        fp32 = {"value": cells * args.steps / el32, "unit": "cell-updates/s",
                "ms_per_step": el32 * 1e3 / args.steps, "speedup_vs_fp64": el / el32,
                "rel_l1_vs_fp64": {k: rel_l1(st32[k], st64[k]) for k in st64},
                "max_divB": sc32["divB"],
                "kernel_ms": {k: round(v, 4) for k, v in kt32.items()},
                "note": "FP64 device path is bitwise equal to the oracle, so this is "
                        "the error against the FP64 oracle at the same step count"}
    log("bench: copy calibration")
    cbytes = 2 * 8 * (args.n + 9) ** 3 * 6  # read + write of six fields
    cms = go.copy_calibrate(cbytes // 2, 20)
    return el, kt, sc, cbytes / (cms * 1e-3) / 1e9, fp32, io, dom, dom_ms


RING_KERNELS = ("k_pole_e", "k_pole_b", "k_ring_rows", "k_ring_curl", "k_ring_fix",
                "k_ring_mags", "k_ring_dt")


def time_sphere(kind, dims, seed, warmup, steps, precision=0):
    """Steps of a spherical problem; wall time and kernel averages."""
    import gamera_op as go
    P = go.Problem(kind, seed, *dims)
    P.device_open(precision)
    P.device_step(warmup)
    # per-kernel averages from their own steps (events around every launch),
    # then the timed steps without events
    P.enable_timing(True)
    P.device_step(min(steps, BREAKDOWN_STEPS))
    kt = P.kernel_times()
    P.enable_timing(False)
    t0 = time.perf_counter()
    P.device_step(steps)
    el = time.perf_counter() - t0
    sc = P.device_scalars()
    P.close()
    if sc["nonfinite"]:
        raise RuntimeError("non-finite values in the spherical run")
    return el, kt, sc


def run_solve_path(args):
    """The facade's solve(nt) on the reference's CPU-runnable case (128x128x8
    sinusoid, FP64, 200 steps) against gk_step(nt) on the same device state:
    solve() runs the reference's loop with time_stop and the status-line
    cadence checked on the device, ending a run of steps only where one of
    them falls due.  Two configurations: time_stop finite beyond the run's
    end with no status cadence, and the same with a status line every tenth
    of an output interval equal to the run's length (about 11 host round
    trips).  solve() also reads both time levels back at the end; that
    read-back is timed on its own and reported beside the totals."""
    sys.path.insert(0, PKG)
    import ctypes
    import gamera_op as go
    nt = 200
    P = go.Problem(go.SINUSOID, args.seed, 128, 128, 8)
    h = P.h
    t_stop = ctypes.c_double.in_dll(h, "time_stop")
    t_out = ctypes.c_double.in_dll(h, "output_interval")
    log_seq = ctypes.c_int.in_dll(h, "log_seq_num")  # a fresh program starts it at 0

    def fresh():
        P.reseed(args.seed)
        P.device_open(0)

    res = {"step": [], "fetch": [], "solve_far": [], "solve_status": []}
    t_end = None
    for rep in range(4):  # the first round is warm-up
        fresh()
        t0 = time.perf_counter()
        P.device_step(nt)
        t1 = time.perf_counter()
        P.device_fetch()
        t2 = time.perf_counter()
        t_end = P.device_scalars()["time"]
        for key, stop, out in (("solve_far", 2.0 * t_end, 1e300),
                               ("solve_status", 2.0 * t_end, t_end)):
            fresh()
            t_stop.value, t_out.value, log_seq.value = stop, out, 0
            t3 = time.perf_counter()
            P.solve(nt)
            t4 = time.perf_counter()
            t_stop.value, t_out.value = 1e300, 1e300
            if P.device_scalars()["time"] != t_end:
                raise RuntimeError("solve() and gk_step() ended at different times")
            if rep:
                res[key].append(t4 - t3)
        if rep:
            res["step"].append(t1 - t0)
            res["fetch"].append(t2 - t1)
    P.close()
    t_stop.value, t_out.value = 1e300, 1e300
    med = {k: sorted(v)[len(v) // 2] * 1e3 for k, v in res.items()}
    out = {"case": "128x128x8 sinusoid, FP64, 200 steps", "gk_step_ms": med["step"],
           "readback_ms": med["fetch"], "solve_far_stop_ms": med["solve_far"],
           "solve_status_lines_ms": med["solve_status"],
           "ratio_far": med["solve_far"] / (med["step"] + med["fetch"]),
           "ratio_status": med["solve_status"] / (med["step"] + med["fetch"]),
           "note": "ratios are solve(200) over gk_step(200) plus the read-back of both "
                   "time levels that solve() ends with; medians of 3 after a warm-up"}
    log(f"bench: solve path {out}")
    return out


def run_ring(args):
    """Full sphere (theta in [0, pi], log r in [1, 10], phi periodic) with
    pole ring averaging at both poles, FP64, against the same grid of the
    spherical shell problem (theta in [pi/8, 7pi/8], no poles): the ring
    kernels' share of the step and the step rate; and the full sphere in the
    FP32-bulk form."""
    sys.path.insert(0, PKG)
    import gamera_op as go
    dims = tuple(args.ring_dims)
    cells = dims[0] * dims[1] * dims[2]
    log(f"bench: full sphere {dims} with pole ring averaging, {args.ring_steps} steps")
    el, kt, sc = time_sphere(go.POLE_SPHERE, dims, args.seed, args.warmup, args.ring_steps)
    log(f"bench: full sphere {el * 1e3 / args.ring_steps:.3f} ms/step")
    el_s, kt_s, _ = time_sphere(go.DIPOLE_SHELL, dims, args.seed, args.warmup, args.ring_steps)
    log(f"bench: shell without poles {el_s * 1e3 / args.ring_steps:.3f} ms/step")
    el_32, _, _ = time_sphere(go.POLE_SPHERE, dims, args.seed, args.warmup, args.ring_steps, 1)
    log(f"bench: full sphere, FP32-bulk {el_32 * 1e3 / args.ring_steps:.3f} ms/step")
    ring_ms = {k: round(v, 4) for k, v in kt.items() if k in RING_KERNELS}
    rows = {32: 2, 64: 4, 128: 8, 256: 16}.get(dims[2])
    return {"grid": list(dims), "rows_per_pole": rows, "steps": args.ring_steps,
            "ms_per_step": el * 1e3 / args.ring_steps,
            "cell_updates_per_s": cells * args.ring_steps / el,
            "ring_kernel_ms": ring_ms,
            "ring_share_of_kernel_time": sum(ring_ms.values()) / sum(kt.values()),
            "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
            "shell_no_poles_ms_per_step": el_s * 1e3 / args.ring_steps,
            "fp32_bulk_ms_per_step": el_32 * 1e3 / args.ring_steps,
            "final_scalars": {"time": sc["time"], "dt": sc["dt"], "divB": sc["divB"]}}


def run_ensemble(args):
    """Ensemble of independent Cartesian periodic grids (sizes drawn by seed
    from ENSEMBLE_SIZES; the sinusoid generator with its own seed, half of
    them with the over-pressure sphere), FP32-bulk: all grids advanced in
    batched steps (one launch per stage for the whole ensemble, HIP graphs,
    autotuned march lengths) against the same grids stepped one after the
    other."""
    sys.path.insert(0, PKG)
    import gamera_op as go
    from ensemble import ENSEMBLE_SIZES, ensemble_specs

    specs = ensemble_specs(args.ens_instances, 1000 + args.seed)
    log(f"bench: ensemble of {len(specs)} grids, set-up")
    solvers = []
    cells = 0
    picks = {}
    for kind, sd, shape in sorted(specs, key=lambda t: (t[0], t[2])):
        P = go.Problem(kind, sd, *shape)
        S = go.Solver(P, 1)
        P.close()
        S.autotune(2)
        picks[shape] = (S.get_tuning(go.TUNE_CHUNK), S.get_tuning(go.TUNE_EMF_CHUNK))
        solvers.append(S)
        cells += shape[0] * shape[1] * shape[2]
    log(f"bench: ensemble launch shapes (faces, edges per thread) {picks}")
    steps = args.ens_steps
    B = go.Batch()
    for S in solvers:
        B.add(S)
    B.step(6)  # warm-up, graph capture
    log(f"bench: ensemble batched, {steps} steps")
    t0 = time.perf_counter()
    B.step(steps)
    el_b = time.perf_counter() - t0
    # per-stage launch times of the batch (timed launches, no graphs): where
    # a batched step spends its time, and the launches' sum against the
    # step's wall time (the rest is gaps between launches)
    B.enable_timing(True)
    B.set_graphs(False)
    B.step(12)
    ens_kt = B.kernel_times()
    B.enable_timing(False)
    B.close()
    log(f"bench: ensemble one grid at a time, {steps} steps")
    t0 = time.perf_counter()
    for S in solvers:
        S.step(steps)
    el_u = time.perf_counter() - t0
    for S in solvers:
        S.close()
    return {"instances": len(specs), "cells": cells, "steps": steps, "precision": "fp32-bulk",
            "batched_cell_updates_per_s": cells * steps / el_b,
            "one_at_a_time_cell_updates_per_s": cells * steps / el_u,
            "speedup": el_u / el_b, "batched_s": el_b, "one_at_a_time_s": el_u,
            "batched_ms_per_step": el_b * 1e3 / steps,
            "batched_kernel_ms_per_step": {k: round(v, 4) for k, v in ens_kt.items()},
            "batched_kernel_sum_ms": round(sum(ens_kt.values()), 4),
            "sizes": "drawn from " + str(ENSEMBLE_SIZES)}


def run_ensemble_sizes(args):
    """Same-size batches of each ENSEMBLE_SIZES shape (about 8.4M cells each,
    half sinusoid, half sinusoid + sphere, FP32-bulk, autotuned) and one
    256^3 grid stepped as a one-member batch: the batched rate each size
    reaches on its own, which bounds the mixed ensemble's rate."""
    sys.path.insert(0, PKG)
    import gamera_op as go
    from ensemble import ENSEMBLE_SIZES
    out = []
    for shape in ENSEMBLE_SIZES + [(256, 256, 256)]:
        cells1 = shape[0] * shape[1] * shape[2]
        n = 1 if shape == (256, 256, 256) else max(1, (8 << 20) // cells1)
        solvers = []
        for q in range(n):
            P = go.Problem(go.SINUSOID_SPHERE if q % 2 else go.SINUSOID, 100 + q, *shape)
            S = go.Solver(P, 1)
            P.close()
            S.autotune(2)
            solvers.append(S)
        B = go.Batch()
        for S in solvers:
            B.add(S)
        B.step(6)
        steps = 60
        t0 = time.perf_counter()
        B.step(steps)
        el = time.perf_counter() - t0
        B.close()
        for S in solvers:
            S.close()
        row = {"shape": list(shape), "grids": n, "cells": n * cells1,
               "ms_per_step": el * 1e3 / steps, "cell_updates_per_s": n * cells1 * steps / el}
        log(f"bench: same-size batch {row}")
        out.append(row)
    return out


def run_serving(args, preparer):
    """Served stream (scaled): seeded Poisson arrivals, 70% Cartesian
    (sinusoid + sphere, 32^3..128^3) / 30% spherical shell (dipole + blob),
    steps U(50, 500), half FP32-bulk, admitted under a fixed device-memory
    budget; continuous batching against the batching-off plain loop on the
    same requests."""
    sys.path.insert(0, PKG)
    import serve

    reqs = serve.make_requests(args.serve_requests, args.serve_rate, seed=args.seed)
    log(f"bench: serving {len(reqs)} requests at {args.serve_rate}/s, budget "
        f"{args.serve_budget_gib} GiB")
    # launch shapes of the mix's grid shapes, tuned before the stream starts
    tuning = serve.tune_shapes(reqs)
    log(f"bench: tuned the mix's shapes in {tuning:.1f} s")
    res, wall = serve.serve(reqs, budget_bytes=args.serve_budget_gib << 30, preparer=preparer)
    srv = serve.summarize(res, wall)
    srv["tuning_s_before_stream"] = round(tuning, 3)
    srv["loop_seconds"] = {k: round(v, 3) for k, v in serve.serve.last_stats.items()}
    # device time of the batch's quanta over the stream's wall time (the
    # members' set-up and read-back on the side stream are not counted)
    srv["gpu_busy_share"] = serve.serve.last_stats.get("gpu_busy_s", 0.0) / wall
    log(f"bench: served {srv}")
    log("bench: plain loop on the same requests")
    res_p, wall_p = serve.serve_plain(reqs)
    pl = serve.summarize(res_p, wall_p)
    log(f"bench: plain {pl}")
    return {"requests": len(reqs), "arrival_rate_per_s": args.serve_rate,
            "budget_GiB": args.serve_budget_gib,
            "preparation_workers": preparer.workers if preparer else 0,
            "served": srv, "plain_loop": pl,
            "throughput_ratio": srv["cell_updates_per_s"] / pl["cell_updates_per_s"]}


def host_info():
    """The host CPU the baselines ran on: logical CPUs, CPUs this process may
    use, and lscpu's model, sockets, cores per socket and threads per core."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
        if "sockets" in info and "cores_per_socket" in info:
            info["physical_cores"] = int(info["sockets"]) * int(info["cores_per_socket"])
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


REF_BUILDS = {
    # oracle/Makefile: the reference's own sources where they lie
    "release": ("ref_driver_cart_release",
                "gcc -O3 -march=native -mtune=native -ffast-math -ftree-vectorize "
                "-fomit-frame-pointer -fopenmp (the reference's Release flags, CMakeLists.txt:45-55)"),
    "o2": ("ref_driver_cart", "gcc -O2 -ffp-contract=off -fopenmp (the fixture build)"),
}


def run_reference_cpu(args, threads, build):
    """One build of the compiled reference on a bounded sample of the same
    generator: runs of 1 and 1 + K steps, whose difference cancels set-up and
    output.  None when the binary is absent or does not run."""
    import subprocess
    import tempfile
    name, flags = REF_BUILDS[build]
    exe = os.path.join(ROOT, "oracle", "_ref", name)
    if not os.path.exists(exe):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from refrun import YAML
    n, k = args.ref_n, args.ref_steps
    log(f"bench: compiled reference ({build}) {n}^3, 1 and {1 + k} steps, {threads} threads")
    wall = {}
    try:
        for nt in (1, 1 + k):
            with tempfile.TemporaryDirectory(prefix="gkrefb") as d:
                with open(os.path.join(d, "c.yaml"), "w") as f:
                    f.write(YAML.format(ni=n, nj=n, nk=n, nt=nt))
                env = dict(os.environ, GK_PROBLEM="1", GK_SEED=str(args.seed),
                           OMP_NUM_THREADS=str(threads))
                t0 = time.perf_counter()
                r = subprocess.run([exe, "c.yaml", os.path.join(d, "out.bin"), "0"], cwd=d, env=env,
                                   capture_output=True, text=True, timeout=600)
                wall[nt] = time.perf_counter() - t0
                if r.returncode != 0:
                    log(f"bench: compiled reference ({build}) failed ({r.returncode})")
                    return None
    except (OSError, subprocess.SubprocessError) as e:
        log(f"bench: compiled reference ({build}) did not run ({e})")
        return None
    per_step = (wall[1 + k] - wall[1]) / k
    log(f"bench: compiled reference ({build}) {per_step:.3f} s/step")
    return {"value": n ** 3 / per_step, "unit": "cell-updates/s", "cores": threads,
            "kind": "reference", "build": flags,
            "sample": f"{n}^3 cells of the same generator (sinusoid + uniform B0 + over-pressure "
                      f"sphere), the compiled reference run for 1 and {1 + k} AB2 steps; value "
                      f"from the difference of the two wall times"}


def run_cpu(args):
    """CPU baselines on bounded samples of the same workload: the compiled
    reference when it is present (the reported baseline), and the CPU
    restatement (port), reported beside it."""
    threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, PKG)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import gamera_op as go
    from parity import oracle_from_problem

    n, steps = args.cpu_n, args.cpu_steps
    log(f"bench: CPU sample {n}^3, {steps} steps, {threads} threads")
    P = go.Problem(go.SINUSOID_BLAST, args.seed, n, n, n, full_b0=True)
    O = oracle_from_problem(P, fast=True)
    P.close()
    O.prologue()
    O.step(1)  # warm-up
    t0 = time.perf_counter()
    O.step(steps)
    el = time.perf_counter() - t0
    log(f"bench: CPU {el:.2f} s")
    port = {"value": n ** 3 * steps / el, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n}^3 cells of the same generator, {steps} AB2 steps after 1 warm-up, "
                      f"oracle/liboracle_fast.so (-O3 -march=native, OpenMP)"}
    host = host_info()
    refs = {} if args.no_ref_cpu else {b: run_reference_cpu(args, threads, b) for b in REF_BUILDS}
    refs = {b: r for b, r in refs.items() if r is not None}
    if not refs:
        port["host"] = host
        return port
    # the reported baseline is the faster build of the reference on these cores
    best = max(refs, key=lambda b: refs[b]["value"])
    out = dict(refs[best])
    out["builds"] = {b: {"value": r["value"], "build": r["build"]} for b, r in refs.items()}
    out["host"] = host
    out["port"] = port
    out["note"] = ("the reference keeps 261 full-grid FP64 fields and runs several loops of "
                   "the step serially (the magnetic stress has no OpenMP), so it scales poorly "
                   "with threads; the port fuses the step into a few parallel passes")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk comparison run")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--ref-n", type=int, default=64, help="compiled-reference CPU sample size")
    ap.add_argument("--ref-steps", type=int, default=5)
    ap.add_argument("--no-ref-cpu", action="store_true",
                    help="time only the CPU restatement, not the compiled reference")
    ap.add_argument("--chunk", type=int, default=0, help="faces per thread in the i/j sweeps")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-instances", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=300)
    ap.add_argument("--no-ring", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the solve(nt) run-control leg")
    ap.add_argument("--ring-only", action="store_true",
                    help="run only the full-sphere ring-average leg (profiling)")
    ap.add_argument("--ring-dims", type=int, nargs=3, default=[128, 128, 256])
    ap.add_argument("--ring-steps", type=int, default=50)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--only", choices=["serve", "ensemble", "solve", "sizes"], default=None,
                    help="run one leg alone (profiling)")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-budget-gib", type=int, default=64)
    ap.add_argument("--serve-workers", type=int, default=0,
                    help="request-preparation processes (default: host cores - 1, at most 16)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver runs a single-GPU domain "
              "(multi-GPU sharding is not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.ring_only:
        print(json.dumps({"pole_ring": run_ring(args)}), flush=True)
        return
    if args.only in ("serve", "ensemble", "solve", "sizes"):  # one leg alone
        sys.path.insert(0, PKG)
        if args.only == "sizes":
            out = {"ensemble_sizes": run_ensemble_sizes(args)}
        elif args.only == "solve":
            out = {"solve_path": run_solve_path(args)}
        elif args.only == "serve":
            import serve
            preparer = serve.Preparer(args.serve_workers or None)
            out = {"serving": run_serving(args, preparer)}
            preparer.close()
        else:
            out = {"ensemble": run_ensemble(args)}
        print(json.dumps(out), flush=True)
        return
    preparer = None
    if not args.no_serve:
        # request-preparation processes start before this process touches the GPU
        sys.path.insert(0, PKG)
        import serve
        preparer = serve.Preparer(args.serve_workers or None)

    el, kt, sc, copy_gbs, fp32, io, dom, dom_ms = run_gpu(args)
    n = args.n
    cells = n ** 3
    value = cells * args.steps / el
    ms_step = el * 1e3 / args.steps
    alg = kernel_alg_bytes(dom, n)
    achieved = alg / (dom_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"{dom}_{n}cube_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fp:
            traffic = json.load(fp).get("bytes_per_launch")
    ring = None if args.no_ring else run_ring(args)
    solve_path = None if args.no_solve else run_solve_path(args)
    ens = None if args.no_ensemble else run_ensemble(args)
    srv = None if args.no_serve else run_serving(args, preparer)
    if preparer is not None:
        preparer.close()
    cpu = None if args.no_cpu else run_cpu(args)
    step_ms_kernels = sum(kt.values())
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"cartesian periodic {n}^3, sinusoid + uniform B0 |B|=1 + "
                               f"over-pressure sphere, AB2 CFL 0.3, FP64",
                   "cells": cells, "seed": args.seed, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": dom_ms},
        "step_roofline": {"alg_bytes_per_cell_update": 216,
                          "achieved_GBs": 216 * value / 1e9,
                          "frac": 216 * value / 1e9 / HBM_PEAK_GBS,
                          "kernel_ms_per_step": step_ms_kernels},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
        "kernel_ms_note": f"HIP-event averages over {BREAKDOWN_STEPS} steps run before the "
                          "timed ones with events around every launch; the timed steps "
                          "carry events around the dominant stage only",
        "copy_kernel_GBs": copy_gbs,
        # the facade's solve(nt) path: state handed over in host arrays
        "host_buffers": {"value": cells * args.steps / (io[0] + el + io[1]),
                         "unit": "cell-updates/s", "steps": args.steps,
                         "open_upload_prologue_s": io[0], "readback_s": io[1],
                         "note": "timed steps plus device set-up, upload of the host "
                                 "state, prologue and read-back of both time levels"},
        "final_scalars": {"time": sc["time"], "dt": sc["dt"], "divB": sc["divB"]},
        "fp32_bulk": fp32,
        "pole_ring": ring,
        "solve_path": solve_path,
        "ensemble": ens,
        "serving": srv,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
