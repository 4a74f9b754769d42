"""Projection-step benchmark on one MI355X.

A step is one pass of the hot path over a device-resident scene:
RK2 semi-Lagrangian advection -> RHS -> pressure solve to the stated
relative residual -> gradient correction -> max|div u| gate -> commit.
Inputs are synthetic (seeded Fourier-mode spectrum), already resident in
HBM when the timed region starts.

Prints ONE JSON line (rank 0, N = 1).  `--gpus N` with N > 1 exits: this
build measures one GPU (a scene shards by whole scenes across GPUs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import rheon_amd as R  # noqa: E402
from rheon_amd import _native as N  # noqa: E402
from rheon_amd.scenes import GridSpec, courant_dt, fourier_modes  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def mg_cycle_bytes(spec: GridSpec, s: int, pre: int = 2, post: int = 2) -> float:
    """Compulsory bytes of one defect-correction cycle (outer f64 residual +
    V(pre, post) in storage of s bytes): per finest cell
      outer residual  read p, b (f64) + write b0, x0           16 + 2s
      extra pre-sweeps  (pre-1) x (read x, b + write x)         3s each
      restrict  read x, b + write coarse b, x (1/8 cells)       2s + 2s/8
      prolong   read x + coarse x (1/8) + write x               2s + s/8
      post-sweeps (post-1) x 3s, last one adds to p (f64 r/w)   2s + 16
    and coarse levels l >= 1 the same minus the f64 terms, on N/8^l cells."""
    n = spec.cell_len
    finest = (16 + 2 * s) + (pre - 1) * 3 * s + (2 * s + 2 * s / 8) + (2 * s + s / 8) + (post - 1) * 3 * s + (2 * s + 16)
    coarse_per_cell = (pre - 1) * 3 * s + (2 * s + 2 * s / 8) + (2 * s + s / 8) + post * 3 * s
    coarse = sum(coarse_per_cell * n / 8 ** l for l in range(1, 8))
    return finest * n + coarse


def algorithmic_bytes(spec: GridSpec, face_bytes: int, method: str, solver_bytes: int, iterations: int) -> dict:
    """Compulsory HBM bytes of one step (per-kernel minimal-fusion model)."""
    n = spec.cell_len
    F = spec.face_total
    sv = solver_bytes
    if method in ("mixed-mg-dc", "mg-f64-dc"):
        per_iter = mg_cycle_bytes(spec, sv)
    else:
        per_iter = {"jacobi-pcg-v1": 11 * sv * n, "rbsgs-pcg": 17 * sv * n}.get(method, 0)
    return {
        "advect": 2 * face_bytes * F,
        "rhs": face_bytes * F + 8 * n,
        "solve": per_iter * iterations,
        "correct": 8 * n + 2 * face_bytes * F,
        "divmax": face_bytes * F,
    }


# Algorithmic bytes per launch of each profiled kernel, per cell (s = vector bytes).
KERNEL_BYTES_PER_CELL = {
    "pcg_apply_dir": lambda s: 2 * s,      # read d, write q
    "pcg_update": lambda s: 6 * s,         # read p, d, r, q; write p, r
    "pcg_direction": lambda s: 3 * s,      # read r (or z), d; write d
    "pcg_sgs_black": lambda s: 3 * s,      # read r, z; write z (interleaved colours)
    "pcg_sgs_red": lambda s: 3 * s,        # read r, z; write z
    "pcg_true_residual": lambda s: 8 + 8 + s,  # read p (s), b (8), write product (s)
    "mg_outer_residual": lambda s: 16 + 2 * s,  # read p, b (f64); write b0, x0
    "mg_smooth": lambda s: 3 * s,              # read x, b; write x
    "mg_smooth_update": lambda s: 2 * s + 16,  # read x, b; read+write p (f64)
    "mg_restrict": lambda s: 2 * s + 2 * s / 8,  # read x, b; write coarse b, x
    "mg_prolong": lambda s: 2 * s + s / 8,       # read x, coarse x; write x
}


def cpu_baseline(spec: GridSpec, vel, dt: float, iterations_needed: int, budget_s: float = 20.0) -> dict:
    """The oracle (C restatement of the reference, its parallel.rs
    decomposition) on this host's cores, on a bounded sample: one advection,
    the RHS, a few JPCG iterations and the correction, extrapolated to the
    JPCG iteration count the reference method needs for this step."""
    import oracle as O

    threads = min(os.cpu_count() or 1, 16)
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    t0 = time.perf_counter()
    adv = O.advect_velocity(g, vel, dt, threads=threads)
    t_adv = time.perf_counter() - t0
    t0 = time.perf_counter()
    rhs = O.build_rhs(g, adv, dt)
    t_rhs = time.perf_counter() - t0
    k = 4
    t0 = time.perf_counter()
    r = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(0.0, 0.0, 0.0, k), threads=threads)
    t_iter = (time.perf_counter() - t0) / k
    t0 = time.perf_counter()
    out = O.correct(g, 1.0, r.pressure, dt, adv)
    O.actual_divergence_max(g, out)
    t_corr = time.perf_counter() - t0
    step = t_adv + t_rhs + t_corr + t_iter * iterations_needed
    return {
        "value": 1.0 / step,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{spec.counts[0]}x{spec.counts[1]}x{spec.counts[2]}: 1 advection ({t_adv:.2f}s), RHS, "
                   f"{k} jacobi-pcg-v1 iterations ({t_iter * 1e3:.0f} ms each), correction+div; "
                   f"extrapolated to {iterations_needed} iterations (the reference method's count for this "
                   f"step, counted by the bit-identical device jacobi-pcg-v1)"),
        "seconds_per_step_estimate": step,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--method", default="mixed-mg-dc",
                    choices=["jacobi-pcg-v1", "rbsgs-pcg", "mixed-mg-dc", "mg-f64-dc"])
    ap.add_argument("--rel", type=float, default=1e-10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: a projection scene lives on one device and many scenes shard "
              "across GPUs by whole scenes (no collectives); run with --gpus 1.", file=sys.stderr)
        return 2

    spec = GridSpec.unit_box(args.n)
    vel = fourier_modes(spec, seed=2026, modes=32, kmax=16, amplitude=1.0, dtype=np.float32)
    dt = courant_dt(spec, vel, 0.5)
    mg = args.method in ("mixed-mg-dc", "mg-f64-dc")
    s_vec = 4 if args.method == "mixed-mg-dc" else 8  # bytes per Krylov / V-cycle value
    settings = R.PressureSettings(args.rel, 1e-14, 1e-4, 200 if mg else 20000, poll_interval=2 if mg else 32)
    scene = R.Scene(spec, density=1.0, precision=R.F32, method=args.method,
                    solver_precision=R.F32 if args.method == "mixed-mg-dc" else R.F64,
                    settings=settings, actual_divergence_limit=1e-3, max_courant=0.0)
    scene.set_velocity(vel)
    N.lib().rh_device_synchronize()

    for _ in range(args.warmup):
        rep = scene.step(dt)
        if rep.status != R.OK:
            print(f"warmup step failed: {R.STATUS_NAMES.get(rep.status)}", file=sys.stderr)
            return 1

    lib = N.lib()
    lib.rh_profile_enable(1)
    lib.rh_device_synchronize()
    lat, iters, divs = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        rep = scene.step(dt)
        lat.append(time.perf_counter() - s0)
        if rep.status != R.OK:
            print(f"step failed: {R.STATUS_NAMES.get(rep.status)}", file=sys.stderr)
            return 1
        iters.append(rep.pressure.iterations)
        divs.append(rep.actual_divergence_max)
    lib.rh_device_synchronize()
    elapsed = time.perf_counter() - t0
    lib.rh_profile_enable(0) if False else None

    # per-kernel live timings (HIP events on the solver stream)
    import ctypes as C
    kernels = {}
    for name in list(KERNEL_BYTES_PER_CELL) + ["advect_velocity", "build_rhs", "correct", "divergence_max",
                                               "mg_coarse_levels"]:
        n_l, sec = C.c_int64(), C.c_double()
        if lib.rh_profile_query(name.encode(), C.byref(n_l), C.byref(sec)) == 0 and n_l.value:
            kernels[name] = (n_l.value, sec.value)
    lib.rh_profile_enable(0)

    dom = max((k for k in kernels if k in KERNEL_BYTES_PER_CELL), key=lambda k: kernels[k][1])
    launches, secs = kernels[dom]
    avg = secs / launches
    per_launch = KERNEL_BYTES_PER_CELL[dom](s_vec) * spec.cell_len
    achieved = per_launch / avg / 1e9

    mean_iters = float(np.mean(iters))
    ab = algorithmic_bytes(spec, 4, args.method, s_vec, int(round(mean_iters)))
    step_bytes = sum(ab.values())
    ms = elapsed / args.steps * 1e3

    result = {
        "metric": "projection steps/s at 256^3 to fixed rel residual + max|div u| gate (HBM GB/s, p50/p99)",
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 faces + f32 V-cycles, f64 outer residual/pressure" if args.method == "mixed-mg-dc"
                  else "f32 faces, f64 pressure"),
        "data": "synthetic: 32 seeded Fourier modes |k| in [1,16], amplitude ~|k|^-5/3, max|u| = 1",
        "config": {
            "workload": f"projection step {args.n}^3, {args.method} to rel {args.rel:g}",
            "grid": list(spec.counts),
            "method": args.method,
            "relative_residual": args.rel,
            "obstacle": "none",
            "iterations_mean": mean_iters,
            "max_div_u": max(divs),
            "latency_ms_p50": float(np.percentile(np.array(lat) * 1e3, 50)),
            "latency_ms_p99": float(np.percentile(np.array(lat) * 1e3, 99)),
            "algorithmic_GBps": step_bytes / (elapsed / args.steps) / 1e9,
            "cell_updates_per_s": spec.cell_len * args.steps / elapsed,
            "kernels": {k: {"launches": v[0], "avg_us": v[1] / v[0] * 1e6} for k, v in kernels.items()},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": per_launch,
            "avg_launch_us": avg * 1e6,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the reference method's iteration count for this step (device jacobi-pcg-v1 is bit-identical)
        if args.method == "jacobi-pcg-v1":
            need = int(round(mean_iters))
        else:
            ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
            rhs = R.build_rhs(spec, R.advect_velocity(spec, scene.get_velocity(), dt), dt)
            need = ws.solve_rhs(1.0, rhs, dt, settings).iterations
            ws.close()
        result["cpu_baseline"] = cpu_baseline(spec, scene.get_velocity(), dt, need)
    line = json.dumps(result)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
