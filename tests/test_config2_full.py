"""The 128^3 FP32 drop-in provider workload at its stated size: Taylor-Green +
noise (seed 6), the RHS of those faces solved to a relative residual of 1e-5
by f32 Jacobi-PCG, f32 red-black SGS-PCG and the 4-level weighted-Jacobi
V(2,2) multigrid (and the default Chebyshev V(3,1)).

tests/golden/config2_128.json holds the oracle's f64 jacobi-pcg-v1 solution
to rel 1e-10 (written by tests/golden/make_config2_summary.py).  The device
f64 solve must match it bit for bit; the f32 Jacobi iterates must stay
within 1e-5 relative of the oracle's f64 iterates; every f32 solve must
reach a true residual of 1e-5 ||b|| (checked with the oracle's operator) and
a pressure within the stated tolerance (times max|p|) of the converged one."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec, courant_dt, taylor_green_noise

from helpers import ogrid

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "config2_128.json")))
# solves stopped at a true residual of 1e-5 ||b||, against the converged f64
# pressure, relative to its max|p|: Krylov methods leave the smooth error
# modes (measured 2.7e-3 Jacobi, 2.5e-3 RB-SGS), multigrid removes them
# (measured 1.3e-5 both smoothers)
KRYLOV_PRESSURE_TOL = 1e-2
MG_PRESSURE_TOL = 1e-4
JACOBI_ITERATE_TOL = 1e-5


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def case():
    spec = GridSpec.unit_box(128)
    vel = taylor_green_noise(spec, 6, 0.1, np.float32)
    dt = courant_dt(spec, vel, 0.5)
    assert digest(*vel) == GOLD["inputs"] and repr(dt) == GOLD["dt"]
    rhs = R.build_rhs(spec, vel, dt)
    assert digest(rhs) == GOLD["rhs"]
    return spec, vel, dt, rhs


def true_rel(spec, p, rhs):
    r = rhs - O.apply_full(ogrid(spec), 1.0, p, threads=8)
    return float(np.sqrt(r @ r) / np.sqrt(rhs @ rhs))


def pressure_err(p):
    cells = np.array(GOLD["sample_cells"])
    want = np.array([float(v) for v in GOLD["sample_pressure"]])
    return float(np.max(np.abs(p[cells] - want))) / float(GOLD["pressure_max_abs"])


@pytest.mark.gpu
def test_f64_reference_solve_bits_128(case):
    spec, vel, dt, rhs = case
    rel, ab, div, mx = GOLD["settings"]
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(rel, ab, div, int(mx)))
    assert rep.status == GOLD["status"] and rep.iterations == GOLD["iterations"]
    assert repr(rep.true_residual_l2) == GOLD["true_residual_l2"]
    assert digest(ws.get_pressure()) == GOLD["pressure"]
    ws.close()


@pytest.mark.gpu
def test_f32_jacobi_iterates_against_oracle_128(case):
    """f32 Krylov vectors (f64 scalars and solution): the iterate after k
    iterations within 1e-5 relative of the oracle's f64 iterate."""
    spec, vel, dt, rhs = case
    g = ogrid(spec)
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F32)
    for k in (1, 10, 50, 200):
        ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k))
        ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(0.0, 0.0, 0.0, k), threads=8)
        p = ws.get_pressure()
        err = float(np.max(np.abs(p - ref.pressure)) / np.max(np.abs(ref.pressure)))
        assert err <= JACOBI_ITERATE_TOL, (k, err)
    ws.close()


def first_iteration_to(ws, rhs, dt, target, cap=8192):
    """First k whose true residual is <= target ||b|| (stopping tests off):
    the reference's termination stalls on this field at rel 1e-5 (its true
    residual includes the gauge row), so the count is found directly."""
    bnorm = float(np.sqrt(rhs @ rhs))

    def rel(k):
        return ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k, poll_interval=32)).true_residual_l2 / bnorm

    lo, hi = 1, 64
    while rel(hi) > target:
        assert hi < cap
        lo, hi = hi, hi * 2
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if rel(mid) > target:
            lo = mid
        else:
            hi = mid
    rel(hi)
    return hi


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_f32_krylov_to_1e5_128(case, method):
    spec, vel, dt, rhs = case
    ws = R.PressureWorkspace(spec, method, R.F32)
    k = first_iteration_to(ws, rhs, dt, 1e-5)
    p = ws.get_pressure()
    assert true_rel(spec, p, rhs) <= 1e-5 * (1 + 1e-6)
    err = pressure_err(p)
    print(method, "iterations", k, "pressure error", err)
    assert err <= KRYLOV_PRESSURE_TOL, err
    ws.close()


@pytest.mark.gpu
@pytest.mark.parametrize("smoother,sweeps", [(R.SMOOTHER_JACOBI, (2, 2)), (R.SMOOTHER_CHEBYSHEV, (3, 1))])
def test_four_level_multigrid_to_1e5_128(case, smoother, sweeps):
    """The stated multigrid: 4 levels (128 -> 16), weighted-Jacobi V(2,2),
    f32 V-cycles in f64 defect correction, to rel 1e-5."""
    spec, vel, dt, rhs = case
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32, mg_levels=4, mg_smoother=smoother, mg_sweeps=sweeps)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-5, 1e-14, 1.0, 100))
    assert rep.status == R.OK, R.STATUS_NAMES.get(rep.status)
    p = ws.get_pressure()
    rel = true_rel(spec, p, rhs)
    assert rel <= 1e-5 * (1 + 1e-6)
    assert rep.true_residual_l2 / float(np.sqrt(rhs @ rhs)) == pytest.approx(rel, rel=1e-6)
    err = pressure_err(p)
    print("multigrid", smoother, sweeps, "cycles", rep.iterations, "pressure error", err)
    assert err <= MG_PRESSURE_TOL, err
    ws.close()
