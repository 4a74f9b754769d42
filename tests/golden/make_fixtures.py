"""Writes tests/golden/fixtures.json: digests of the oracle's outputs on the
small workload recipes that no reference-held vector covers (Taylor-Green +
noise at 32^3 with the 200-iteration Jacobi-PCG step, and a Fourier field
around a solid sphere at 32^3).  The oracle itself is pinned by the
reference's own known-answer tests and goldens (tests/test_oracle.py); these
fixtures freeze its outputs so that later changes to the oracle or to the
device path are caught against committed data.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from rheon_amd.scenes import (GridSpec, centred_sphere, close_obstacle_faces, courant_dt,  # noqa: E402
                              fourier_modes, taylor_green_noise)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def config1():
    spec = GridSpec.unit_box(32)
    vel = taylor_green_noise(spec, 1, 0.1, np.float64)
    dt = courant_dt(spec, vel, 0.5)
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    adv = O.advect_velocity(g, vel, dt)
    rhs = O.build_rhs(g, adv, dt)
    res = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(0.0, 0.0, 0.0, 200))
    out = O.correct(g, 1.0, res.pressure, dt, adv)
    return {
        "recipe": "32^3 unit box, Taylor-Green + LCG noise 0.1 seed 1, f64, dt = 0.5 h / max|u|, "
                  "advect -> RHS -> 200 jacobi-pcg-v1 iterations -> correct",
        "dt": repr(dt),
        "status": res.status,
        "iterations": res.iterations,
        "advected": digest(*adv),
        "rhs": digest(rhs),
        "pressure": digest(res.pressure),
        "corrected": digest(*out),
        "max_div": repr(O.actual_divergence_max(g, out)),
    }


def sphere():
    spec = GridSpec.unit_box(32)
    solid = centred_sphere(spec, 0.2)
    vel = close_obstacle_faces(spec, solid, fourier_modes(spec, 11, dtype=np.float32))
    dt = courant_dt(spec, vel, 0.5)
    g = O.Grid(spec.counts, spec.spacing, spec.origin)
    ob = O.obstacle_mask(g, solid)
    adv = close_obstacle_faces(spec, solid, O.advect_velocity(g, vel, dt))
    rhs = O.build_rhs(g, adv, dt)
    res = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, O.Settings(1e-9, 1e-14, 1e-4, 2000))
    out = O.correct_masked(g, 1.0, ob, res.pressure, dt, adv)
    return {
        "recipe": "32^3 unit box, 32 Fourier modes seed 11 (f32 faces), solid sphere r=0.2 at the centre, "
                  "dt = 0.5 h / max|u|, advect (obstacle faces closed) -> RHS -> masked jacobi-pcg-v1 to "
                  "rel 1e-9 -> masked correct",
        "dt": repr(dt),
        "solid_cells": int(solid.sum()),
        "gauge": ob.gauge,
        "status": res.status,
        "iterations": res.iterations,
        "true_residual_l2": repr(res.true_residual_l2),
        "advected": digest(*adv),
        "pressure": digest(res.pressure),
        "corrected": digest(*out),
        "max_div": repr(O.actual_divergence_max_masked(g, ob, out)),
    }


def main():
    data = {"config1_32": config1(), "sphere_32": sphere()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")
    print(path)


if __name__ == "__main__":
    main()
