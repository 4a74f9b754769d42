"""Committed fixtures (tests/golden/fixtures.json, written by
tests/golden/make_fixtures.py from the oracle): the oracle must keep
reproducing them, and the device path must match them bit for bit."""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest

from rheon_amd.scenes import GridSpec, centred_sphere, close_obstacle_faces, courant_dt, fourier_modes, \
    taylor_green_noise

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = json.load(open(os.path.join(HERE, "golden", "fixtures.json")))


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_oracle_reproduces_committed_fixtures():
    spec = importlib.util.spec_from_file_location("make_fixtures", os.path.join(HERE, "golden", "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert mod.config1() == FIXTURES["config1_32"]
    assert mod.sphere() == FIXTURES["sphere_32"]


@pytest.mark.gpu
def test_device_config1_matches_fixture():
    import rheon_amd as R
    fx = FIXTURES["config1_32"]
    spec = GridSpec.unit_box(32)
    vel = taylor_green_noise(spec, 1, 0.1, np.float64)
    dt = courant_dt(spec, vel, 0.5)
    assert repr(dt) == fx["dt"]
    adv = R.advect_velocity(spec, vel, dt)
    assert digest(*adv) == fx["advected"]
    rhs = R.build_rhs(spec, adv, dt)
    assert digest(rhs) == fx["rhs"]
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, 200))
    assert rep.status == fx["status"] and rep.iterations == fx["iterations"]
    p = ws.get_pressure()
    assert digest(p) == fx["pressure"]
    out = R.correct_velocity(spec, 1.0, p, dt, adv)
    assert digest(*out) == fx["corrected"]
    assert repr(R.divergence_max(spec, out)) == fx["max_div"]


@pytest.mark.gpu
def test_device_sphere_step_matches_fixture():
    import rheon_amd as R
    fx = FIXTURES["sphere_32"]
    spec = GridSpec.unit_box(32)
    solid = centred_sphere(spec, 0.2)
    vel = close_obstacle_faces(spec, solid, fourier_modes(spec, 11, dtype=np.float32))
    dt = courant_dt(spec, vel, 0.5)
    sc = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64,
                 settings=R.PressureSettings(1e-9, 1e-14, 1e-4, 2000), actual_divergence_limit=1e-3)
    sc.set_solid(solid)
    sc.set_velocity(vel)
    rep = sc.step(dt)
    assert rep.status == R.OK
    assert rep.pressure.iterations == fx["iterations"]
    assert repr(rep.pressure.true_residual_l2) == fx["true_residual_l2"]
    assert digest(sc.get_pressure()) == fx["pressure"]
    assert digest(*sc.get_velocity()) == fx["corrected"]
    assert repr(rep.actual_divergence_max) == fx["max_div"]
