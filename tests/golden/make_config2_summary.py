"""Writes tests/golden/config2_128.json: a summary of the oracle's converged
pressure for the 128^3 FP32 provider workload recipe (Taylor-Green + LCG
noise 0.1, seed 6, f32 faces, dt = 0.5 h / max|u|, RHS of those faces),
solved by the oracle's f64 jacobi-pcg-v1 to rel 1e-10 (abs 1e-14,
predicted-divergence gate off).  Holds the RHS and pressure digests, the
iteration count and residuals, and the pressure at fixed cells with its
max |p|, against which the device's f32 solvers are held to the stated
tolerances.

Run from the repository root (about a minute on 8 threads):
    python tests/golden/make_config2_summary.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from rheon_amd.scenes import GridSpec, courant_dt, taylor_green_noise  # noqa: E402

N = 128
SETTINGS = (1e-10, 1e-14, 1.0, 20000)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    spec = GridSpec.unit_box(N)
    vel = taylor_green_noise(spec, 6, 0.1, np.float32)
    dt = courant_dt(spec, vel, 0.5)
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    rhs = O.build_rhs(g, vel, dt)
    res = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(*SETTINGS), threads=threads)
    cells = np.array(sorted(set(int(v) for v in (np.arange(96) * 21851 + 777) % spec.cell_len)
                            | {0, spec.cell_len - 1, N - 1, N * N - 1}), np.int64)
    data = {
        "recipe": ("128^3 unit box, Taylor-Green + LCG noise 0.1 seed 6 (f32 faces), dt = 0.5 h / max|u|, RHS of "
                   "those faces; oracle jacobi-pcg-v1 f64 to rel 1e-10, abs 1e-14, predicted-divergence gate off"),
        "settings": list(SETTINGS),
        "dt": repr(dt),
        "inputs": digest(*vel),
        "rhs": digest(rhs),
        "status": res.status,
        "iterations": res.iterations,
        "true_residual_l2": repr(res.true_residual_l2),
        "pressure": digest(res.pressure),
        "pressure_max_abs": repr(float(np.max(np.abs(res.pressure)))),
        "sample_cells": [int(c) for c in cells],
        "sample_pressure": [repr(float(res.pressure[c])) for c in cells],
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config2_128.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print(path, res.iterations, res.status, f"{time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
