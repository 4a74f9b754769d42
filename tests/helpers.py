"""Shared inputs for the parity tests."""
import numpy as np

import oracle as O
from rheon_amd.scenes import GridSpec, courant_dt, taylor_green_noise


def ogrid(spec: GridSpec) -> O.Grid:
    return O.Grid(spec.counts, spec.spacing, spec.origin)


def wavy_faces(spec: GridSpec, dtype=np.float32, seed: int = 0):
    """0.8 sin(0.017 i + d) per face (tests/advection_contract.rs:9-13 pattern),
    wall-normal faces zeroed."""
    out = []
    for d in range(3):
        x = np.arange(spec.face_len(d))
        f = 0.8 * np.sin(x * 0.017 + d + seed)
        f[~spec.interior_mask(d).ravel()] = 0.0
        out.append(np.ascontiguousarray(f.astype(dtype)))
    return out


def config1(n: int = 32, seed: int = 1, dtype=np.float64):
    """Taylor-Green + LCG noise in [-0.1, 0.1] on the unit box, dt = 0.5 h / max|u|."""
    spec = GridSpec.unit_box(n)
    vel = taylor_green_noise(spec, seed, 0.1, dtype)
    return spec, vel, courant_dt(spec, vel, 0.5)


def rel_err(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / den
