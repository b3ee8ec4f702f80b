"""The two single-grid workloads at their full sizes on the device.

* Cartesian periodic box, 256^3 cells (sinusoid + uniform B0 |B| = 1 +
  over-pressure sphere), 100 AB2 steps: the throughput workload.
* Spherical shell, 192x96x192 cells (log-spaced r in [1, 10], theta in
  [pi/8, 7pi/8], periodic phi; dipole field + blob), 200 AB2 steps.

At these sizes the FP64 device state must still equal the CPU oracle bit for
bit after the first steps (the oracle takes seconds per step here).  Over the
whole run the checks are properties that hold at any size:

* max |div B| at FP64 roundoff (constrained transport);
* in the periodic box, total mass sum(rho V) is conserved, and so is the net
  face-B flux through every i, j and k plane (the edge fields telescope in a
  periodic box), both to summation roundoff;
* the FP32-bulk form against the FP64 form at the same step count: relative
  L1 <= 1e-5 for density, pressure, total energy and face B (as a vector),
  momentum (as a vector, near zero at the start) <= 1e-4, as in the
  small-grid FP32-bulk tests.
"""
import numpy as np
import pytest

import gamera_op as go
from parity import run_oracle, compare_states, exact_mismatches, active_slices, conserved

VOL_CENTER = 15  # geometry field index of the cell volume (src/config.h layout)


def fetch_active(P):
    """Copies of the active primitives and face B (host arrays, after a fetch)."""
    cell, faces = active_slices(P)
    prim = [P.gas(f)[cell].copy() for f in range(6)]
    bf = [P.gem(f)[faces[f]].copy() for f in range(3)]
    return prim, bf


def run_device(kind, seed, n, steps, precision):
    P = go.Problem(kind, seed, *n)
    try:
        cell, faces = active_slices(P)
        vol = P.geo(VOL_CENTER)[cell].copy()
        init = fetch_active(P)
        P.device_open(precision)
        P.device_step(steps)
        P.device_fetch()
        sc = P.device_scalars()
        return init, fetch_active(P), vol, sc
    finally:
        P.close()


def fp32_vs_fp64(s32, s64):
    (p32, b32), (p64, b64) = s32, s64
    c32, c64 = conserved(p32), conserved(p64)
    rel = {k: float(np.abs(c32[k] - c64[k]).sum() / np.abs(c64[k]).sum())
           for k in ("rho", "eng")}
    rel["p"] = float(np.abs(p32[4] - p64[4]).sum() / np.abs(p64[4]).sum())
    rel["B"] = float(sum(np.abs(x - y).sum() for x, y in zip(b32, b64))
                     / sum(np.abs(y).sum() for y in b64))
    md = sum(np.abs(c32[k] - c64[k]).sum() for k in ("rhov1", "rhov2", "rhov3"))
    mn = sum(np.abs(c64[k]).sum() for k in ("rhov1", "rhov2", "rhov3"))
    rel["rhov"] = float(md / mn)
    return rel


FIRST_STEPS = [
    # kind, seed, ni, nj, nk, steps; the 256^3 case spends about a minute on
    # the host (B0 quadratures of the reference arrays, one oracle step)
    pytest.param(go.SINUSOID_BLAST, 1, 256, 256, 256, 1, marks=pytest.mark.timeout(400)),
    (go.DIPOLE_SHELL, 2, 192, 96, 192, 2),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed,ni,nj,nk,steps", FIRST_STEPS)
def test_full_size_first_steps_match_oracle(kind, seed, ni, nj, nk, steps):
    print(f"full-size oracle case kind={kind} seed={seed} {ni}x{nj}x{nk} steps={steps}",
          flush=True)
    P = go.Problem(kind, seed, ni, nj, nk, full_b0=True)
    try:
        O = run_oracle(P, steps)
        print("oracle done", flush=True)
        try:
            P.device_open(0)
            P.device_step(steps)
            P.device_fetch()
            sc = P.device_scalars()
            nbad = exact_mismatches(P, O)
            rel = compare_states(P, O)
            print({k: f"{v:.2e}" for k, v in rel.items()}, "divB", sc["divB"],
                  "bitwise mismatches", nbad, flush=True)
            assert nbad == 0
            assert sc["nonfinite"] == 0.0
            assert abs(sc["time"] - O.scalar("time_sim")) <= 1e-12 * O.scalar("time_sim")
            assert abs(sc["dt"] - O.scalar("dt")) <= 1e-10 * O.scalar("dt")
        finally:
            O.close()
    finally:
        P.close()


@pytest.mark.gpu
def test_box_256_conservation_and_fp32_bulk():
    n, steps = (256, 256, 256), 100
    print(f"box 256^3, {steps} steps, FP64", flush=True)
    (p0, b0), s64, vol, sc64 = run_device(go.SINUSOID_BLAST, 1, n, steps, 0)
    print("FP64 scalars", sc64, flush=True)
    assert sc64["nonfinite"] == 0.0
    assert sc64["divB"] < 1e-10
    m0 = float((p0[0] * vol).sum())
    m1 = float((s64[0][0] * vol).sum())
    print("mass", m0, m1, "relative change", abs(m1 - m0) / m0, flush=True)
    assert abs(m1 - m0) <= 1e-12 * m0
    # net flux of B_i through each i plane (sum over j, k), and cyclic
    b1 = s64[1]
    for a in range(3):
        other = tuple(x for x in range(3) if x != a)
        f0 = b0[a].sum(axis=other)
        f1 = b1[a].sum(axis=other)
        scale = np.abs(b0[a]).sum(axis=other) + np.abs(b1[a]).sum(axis=other) + 1e-300
        drift = float(np.max(np.abs(f1 - f0) / scale))
        print(f"face-B flux through {'ijk'[a]} planes: max relative drift {drift:.2e}", flush=True)
        assert drift <= 1e-12
    print(f"box 256^3, {steps} steps, FP32-bulk", flush=True)
    _, s32, _, sc32 = run_device(go.SINUSOID_BLAST, 1, n, steps, 1)
    rel = fp32_vs_fp64(s32, s64)
    print("FP32-bulk vs FP64", {k: f"{v:.2e}" for k, v in rel.items()}, "divB", sc32["divB"],
          flush=True)
    assert sc32["nonfinite"] == 0.0
    assert sc32["divB"] < 1e-10
    for k in ("rho", "p", "eng", "B"):
        assert rel[k] <= 1e-5, (k, rel[k])
    assert rel["rhov"] <= 1e-4, rel


@pytest.mark.gpu
def test_shell_192_fp64_and_fp32_bulk():
    n, steps = (192, 96, 192), 200
    print(f"shell 192x96x192, {steps} steps, FP64", flush=True)
    _, s64, _, sc64 = run_device(go.DIPOLE_SHELL, 2, n, steps, 0)
    print("FP64 scalars", sc64, flush=True)
    assert sc64["nonfinite"] == 0.0
    assert sc64["divB"] < 1e-10
    print(f"shell 192x96x192, {steps} steps, FP32-bulk", flush=True)
    _, s32, _, sc32 = run_device(go.DIPOLE_SHELL, 2, n, steps, 1)
    rel = fp32_vs_fp64(s32, s64)
    print("FP32-bulk vs FP64", {k: f"{v:.2e}" for k, v in rel.items()}, "divB", sc32["divB"],
          flush=True)
    assert sc32["nonfinite"] == 0.0
    assert sc32["divB"] < 1e-10
    for k in ("rho", "p", "eng", "B"):
        assert rel[k] <= 1e-5, (k, rel[k])
    assert rel["rhov"] <= 1e-4, rel
