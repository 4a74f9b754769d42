"""HIP path against the oracle, through the C ABI (include/rheon_gpu.h).

Integer/bit-level work and the f64 arithmetic paths are compared bit for
bit: the kernels follow the reference's operation order with no FMA
contraction, and dot products are double-double sums that reproduce the
reference's Neumaier results.  Tolerances are written where a path is not
bit-exact by design (f32 Krylov storage: 1e-5 relative per iterate).
"""
import math

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec, courant_dt, fourier_modes, lcg_faces, taylor_green_noise

from helpers import config1, ogrid, rel_err, wavy_faces

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 9, 1), (7, 1, 2), (33, 17, 5), (65, 3, 9), (32, 32, 32)]


def bits(a):
    return np.ascontiguousarray(a).tobytes()


def test_device_is_gfx950():
    n, arch = R.device_info()
    assert n >= 1
    assert arch.startswith("gfx950"), arch


@pytest.mark.parametrize("counts", SHAPES + [(64, 64, 64)])
def test_apply_full_bits(counts):
    spec = GridSpec(counts, (0.13, 0.27, 0.41), (-1.0, 2.0, 0.3))
    rng = np.random.default_rng(sum(counts))
    x = rng.standard_normal(spec.cell_len)
    assert bits(R.apply_full(spec, 1.37, x)) == bits(O.apply_full(ogrid(spec), 1.37, x))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("counts", SHAPES)
def test_build_rhs_correct_divergence_bits(counts, dtype):
    spec = GridSpec(counts, (0.07, 0.11, 0.19), (-2.0, 4.0, 0.3))
    g = ogrid(spec)
    vel = wavy_faces(spec, dtype)
    dt = 0.013
    assert bits(R.build_rhs(spec, vel, dt)) == bits(O.build_rhs(g, vel, dt))
    rng = np.random.default_rng(5)
    p = rng.standard_normal(spec.cell_len)
    got = R.correct_velocity(spec, 1.3, p, dt, vel)
    want = O.correct(g, 1.3, p, dt, vel)
    for a in range(3):
        assert bits(got[a]) == bits(want[a])
    assert R.divergence_max(spec, got) == O.actual_divergence_max(g, want)
    assert R.velocity_rate(spec, vel) == O.velocity_rate(g, vel)


def test_operator_rejections():
    spec = GridSpec((3, 1, 1), (1.0,) * 3)
    bad = [np.array([0.0, 1.0, 0.0, 1.0], np.float32), np.zeros(6, np.float32), np.zeros(6, np.float32)]
    with pytest.raises(R.RheonError) as e:
        R.build_rhs(spec, bad, 0.5)
    assert e.value.status == R.NONZERO_WALL_VELOCITY
    bad[0] = np.array([0.0, np.nan, 0.0, 0.0], np.float32)
    with pytest.raises(R.RheonError) as e:
        R.build_rhs(spec, bad, 0.5)
    assert e.value.status == R.NONFINITE_INPUT
    with pytest.raises(R.RheonError) as e:
        R.build_rhs(spec, [np.zeros(4, np.float32), np.zeros(6, np.float32), np.zeros(6, np.float32)], 0.0)
    assert e.value.status == R.INVALID_TIME_STEP


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("counts", [(1, 1, 1), (9, 7, 5), (33, 17, 5), (32, 32, 32)])
def test_advect_velocity_bits(counts, dtype):
    spec = GridSpec(counts, (0.07, 0.11, 0.19), (-2.0, 4.0, 0.3))
    vel = wavy_faces(spec, dtype)
    for dt in (0.01, 0.5, 3.0):
        got = R.advect_velocity(spec, vel, dt)
        want = O.advect_velocity(ogrid(spec), vel, dt)
        for a in range(3):
            assert bits(got[a]) == bits(want[a]), (a, dt)


# Power-of-two spacings, where every face position is an exact lattice
# point of the sampler's arithmetic (first RK2 stage weights exactly 0, 1/2
# or 1), and one with an origin off the half-cell lattice.
LATTICE_SPECS = [
    GridSpec((32, 32, 32), (1 / 32,) * 3),
    GridSpec((24, 16, 40), (1 / 16, 1 / 32, 1 / 8), (-0.5, 0.25, 1.0)),
    GridSpec((17, 9, 33), (0.25, 0.5, 0.125), (-2.0, 4.0, 0.5)),
    GridSpec((20, 12, 8), (1 / 16, 1 / 16, 1 / 16), (0.1, 0.3, -0.7)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("which", range(len(LATTICE_SPECS)))
def test_advect_velocity_bits_lattice_grids(which, dtype):
    spec = LATTICE_SPECS[which]
    vel = wavy_faces(spec, dtype)
    for dt in (0.001, 0.01, 0.5, 3.0):
        got = R.advect_velocity(spec, vel, dt)
        want = O.advect_velocity(ogrid(spec), vel, dt)
        for a in range(3):
            assert bits(got[a]) == bits(want[a]), (a, dt)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 96])
def test_advect_velocity_bits_fourier_lattice(n):
    """The config-3 field recipe (Fourier modes) on unit boxes: 64^3 is a
    lattice grid, 96^3 (h = 1/96) is not; both bit for bit."""
    spec = GridSpec.unit_box(n)
    vel = fourier_modes(spec, seed=5, modes=32, kmax=16, amplitude=1.0, dtype=np.float32)
    dt = courant_dt(spec, vel, 0.5)
    got = R.advect_velocity(spec, vel, dt)
    want = O.advect_velocity(ogrid(spec), vel, dt, threads=8)
    for a in range(3):
        assert bits(got[a]) == bits(want[a]), a


@pytest.mark.gpu
def test_advect_velocity_lattice_zero_and_nonfinite():
    """Signed zeros and a non-finite value next to a sampled face."""
    spec = LATTICE_SPECS[0]
    vel = wavy_faces(spec, np.float32)
    vel = [v.copy() for v in vel]
    vel[1][::7] = -0.0
    vel[2][::5] = 0.0
    got = R.advect_velocity(spec, vel, 0.01)
    want = O.advect_velocity(ogrid(spec), vel, 0.01)
    for a in range(3):
        assert bits(got[a]) == bits(want[a]), a
    # a component that is -0.0 almost everywhere: every first-stage sum of it
    # is the +0.0 the reference's running sum gives
    negz = [v.copy() for v in vel]
    negz[0][:] = -0.0
    negz[0][::11] = 0.25
    negz[2][:] = -0.0
    got = R.advect_velocity(spec, negz, 0.3)
    want = O.advect_velocity(ogrid(spec), negz, 0.3)
    for a in range(3):
        assert bits(got[a]) == bits(want[a]), a
    bad = [v.copy() for v in vel]
    bad[1][spec.face_counts(1)[0] * 3 + 40] = np.inf
    with pytest.raises(R.RheonError):
        R.advect_velocity(spec, bad, 0.01)


def test_advect_tracer_bits():
    # tests/advection_contract.rs:6-41 scene: sparse tracer, clamping, tiny donor
    spec = GridSpec((41, 29, 17), (0.07, 0.11, 0.19), (-2.0, 4.0, 0.3))
    vel = []
    for d in range(3):
        x = np.arange(spec.face_len(d))
        vel.append((0.8 * np.sin(x * 0.017 + d)).astype(np.float32))
    nx, ny, _ = spec.counts
    old = np.zeros(spec.cell_len, np.float32)
    old[nx * (8 + ny * 5)] = 1.0
    old[nx - 1 + nx * (9 + ny * 6)] = -0.5
    old[18 + nx * (9 + ny * 6)] = np.frombuffer(np.uint32(1).tobytes(), np.float32)[0]
    for dt in (0.01, 0.5, 3.0):
        assert bits(R.advect_tracer(spec, old, vel, dt)) == bits(O.advect_tracer(ogrid(spec), old, vel, dt))


@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
@pytest.mark.parametrize("counts", [(2, 2, 1), (1, 9, 1), (33, 17, 5), (32, 32, 32)])
def test_precondition_bits(counts, method):
    spec = GridSpec(counts, (0.3, 0.2, 0.5))
    rng = np.random.default_rng(9)
    r = rng.standard_normal(spec.cell_len)
    gr, gz, grz = R.precondition(spec, 1.1, r, method)
    orr, oz, orz = O.precondition(ogrid(spec), 1.1, r, method)
    assert bits(gz) == bits(oz)
    assert gr[0] == 0.0
    assert grz == orz


def jpcg_settings(**kw):
    s = R.PressureSettings(1e-11, 1e-12, 1e-9, 1000)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def osettings(s: R.PressureSettings) -> O.Settings:
    return O.Settings(s.relative_residual, s.absolute_residual, s.divergence_limit, s.max_iterations)


@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_pcg_iterates_bit_for_bit(method):
    """JPCG (jacobi-pcg-v1) and RB-SGS PCG: the iterate after k iterations equals
    the oracle's bit for bit (f64 storage)."""
    spec, vel, dt = config1(32, seed=3)
    g = ogrid(spec)
    rhs = O.build_rhs(g, O.advect_velocity(g, vel, dt), dt)
    ws = R.PressureWorkspace(spec, method, R.F64)
    for k in (1, 2, 3, 7, 16, 17, 60, 200):
        s = jpcg_settings(relative_residual=0.0, absolute_residual=0.0, max_iterations=k)
        rep = ws.solve_rhs(1.0, rhs, dt, s)
        ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(s), method=method)
        assert rep.status == R.ITERATION_LIMIT == ref.status
        assert rep.error_iteration == k
        assert bits(ws.get_pressure()) == bits(ref.pressure), k
        assert rep.true_residual_l2 == ref.true_residual_l2


@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_pcg_replacement_path_bit_for_bit(method):
    """A threshold below the attainable residual (rel 1e-17): from about
    iteration 370 on, candidate iterations take the true residual, are not
    accepted and restart from it (the oracle's iterates then differ from
    the rel = 0 run); the iterate, status and residual equal the oracle's
    bit for bit."""
    spec, vel, dt = config1(32, seed=3)
    g = ogrid(spec)
    rhs = O.build_rhs(g, O.advect_velocity(g, vel, dt), dt)
    ws = R.PressureWorkspace(spec, method, R.F64)
    for k in (380, 420, 500):
        s = jpcg_settings(relative_residual=1e-17, absolute_residual=0.0, max_iterations=k)
        rep = ws.solve_rhs(1.0, rhs, dt, s)
        ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(s), method=method)
        assert rep.status == R.ITERATION_LIMIT == ref.status
        assert rep.error_iteration == k
        assert bits(ws.get_pressure()) == bits(ref.pressure), k
        assert rep.true_residual_l2 == ref.true_residual_l2
    plain = O.solve_rhs(g, 1.0, rhs, dt, osettings(jpcg_settings(relative_residual=0.0, absolute_residual=0.0,
                                                                  max_iterations=500)), method=method)
    assert bits(plain.pressure) != bits(ref.pressure)  # the replacement path was taken


@pytest.mark.parametrize("n", [32, 64])
def test_steady_scene_golden_on_device(n):
    """Reference-published run (docs/plans/accepted-pressure-warm-start): the
    device JPCG reproduces the iteration count and residual bits."""
    from test_oracle import GOLDEN, steady_scene
    spec, g, vel, dt, _ = steady_scene(n)
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
    rep = ws.solve_velocity(1.25, vel, dt, R.PressureSettings(1e-8, 1e-12, 1e-7, 3000))
    gold = GOLDEN[n]
    assert rep.ok
    assert rep.iterations == gold["iterations"]
    assert rep.true_residual_l2 == gold["l2"]
    assert rep.true_residual_max == gold["rmax"]
    out = R.correct_velocity(spec, 1.25, ws.get_pressure(), dt, vel)
    assert R.divergence_max(spec, out) == gold["div"]


def test_config1_projection_step():
    """32^3 f64 Taylor-Green + noise: advect -> RHS -> exactly 200 JPCG
    iterations (no early exit) -> correct -> max|div u|, every stage
    compared with the oracle bit for bit."""
    spec, vel, dt = config1(32, seed=1, dtype=np.float64)
    g = ogrid(spec)
    adv = R.advect_velocity(spec, vel, dt)
    oadv = O.advect_velocity(g, vel, dt)
    for a in range(3):
        assert bits(adv[a]) == bits(oadv[a])
    rhs = R.build_rhs(spec, adv, dt)
    assert bits(rhs) == bits(O.build_rhs(g, oadv, dt))
    s = jpcg_settings(relative_residual=0.0, absolute_residual=0.0, divergence_limit=0.0, max_iterations=200)
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F64)
    rep = ws.solve_rhs(1.0, rhs, dt, s)
    ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(s))
    assert rep.status == R.ITERATION_LIMIT and ref.status == O.ITERATION_LIMIT
    p = ws.get_pressure()
    assert bits(p) == bits(ref.pressure)
    out = R.correct_velocity(spec, 1.0, p, dt, adv)
    oout = O.correct(g, 1.0, ref.pressure, dt, oadv)
    for a in range(3):
        assert bits(out[a]) == bits(oout[a])
    assert R.divergence_max(spec, out) == O.actual_divergence_max(g, oout)


def test_scene_step_matches_oracle_and_is_transactional():
    spec, vel, dt = config1(32, seed=2, dtype=np.float32)
    g = ogrid(spec)
    s = R.PressureSettings(1e-9, 1e-12, 1e-7, 2000)
    scene = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64, settings=s,
                    actual_divergence_limit=1e-3)
    scene.set_velocity(vel)
    rep = scene.step(dt)
    assert rep.status == R.OK, R.STATUS_NAMES.get(rep.status)
    oadv = O.advect_velocity(g, vel, dt)
    ref = O.solve_velocity(g, 1.0, oadv, dt, osettings(s))
    assert ref.ok and rep.pressure.iterations == ref.iterations
    assert bits(scene.get_pressure()) == bits(ref.pressure)
    oout = O.correct(g, 1.0, ref.pressure, dt, oadv)
    got = scene.get_velocity()
    for a in range(3):
        assert bits(got[a]) == bits(oout[a])
    assert rep.actual_divergence_max == O.actual_divergence_max(g, oout)
    # a failing gate leaves the accepted faces untouched
    strict = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64, settings=s,
                     actual_divergence_limit=0.0)
    strict.set_velocity(vel)
    assert strict.step(dt).status == R.DIVERGENCE_LIMIT
    assert all(bits(a) == bits(b) for a, b in zip(strict.get_velocity(), vel))


def test_scene_step_validates_uploaded_faces():
    """Accepted faces committed by a step are not re-validated, faces set
    from the host always are: a non-finite upload after good steps is
    rejected by the next step, and the accepted faces stay as uploaded."""
    spec, vel, dt = config1(32, seed=2, dtype=np.float32)
    s = R.PressureSettings(1e-8, 1e-12, 1e-6, 400)
    scene = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=s, actual_divergence_limit=1e-3)
    scene.set_velocity(vel)
    for _ in range(2):
        assert scene.step(dt).status == R.OK
    bad = [np.array(a, copy=True) for a in scene.get_velocity()]
    bad[1][len(bad[1]) // 2] = np.nan
    scene.set_velocity(bad)
    assert scene.step(dt).status == R.NONFINITE_INPUT
    got = scene.get_velocity()
    assert np.isnan(got[1][len(bad[1]) // 2])


def test_pcg_status_paths():
    spec = GridSpec((2, 1, 1), (1.0,) * 3)
    ws = R.PressureWorkspace(spec)
    assert ws.solve_rhs(1.0, np.array([1.0, 0.0]), 1.0, jpcg_settings()).status == R.INCOMPATIBLE_RHS
    spec = GridSpec((4, 1, 1), (1.0,) * 3)
    ws = R.PressureWorkspace(spec)
    rep = ws.solve_rhs(1.0, np.array([-0.003, 0.001, 0.001, 0.001]), 1.0, R.PressureSettings(1.0, 1.0, 0.0015, 0))
    assert rep.status == R.ITERATION_LIMIT and rep.error_iteration == 0
    spec = GridSpec((1, 1, 1), (1.0,) * 3)
    rep = R.PressureWorkspace(spec).solve_rhs(1.0, np.array([0.0]), 1.0, jpcg_settings())
    assert rep.ok and rep.iterations == 0 and rep.true_residual_max == 0.0
    spec = GridSpec((3, 1, 1), (1.0,) * 3)
    ws = R.PressureWorkspace(spec)
    rep = ws.solve_rhs(1.0, np.array([-4.0, 6.0, -2.0]), 0.5, jpcg_settings())
    assert rep.ok and abs(ws.get_pressure()[1] - 4.0) < 1e-10
    assert ws.solve_rhs(1.0, np.array([-4.0, 6.0, -2.0]), 0.5, jpcg_settings(), cancelled=lambda: True).status == R.CANCELLED
    assert ws.solve_rhs(1.0, np.array([-4.0, 6.0, -2.0]), 0.5,
                        jpcg_settings(relative_residual=math.nan)).status == R.INVALID_SETTINGS
    assert ws.solve_rhs(1.0, np.array([-4.0, np.inf, -2.0]), 0.5, jpcg_settings()).status == R.NONFINITE_INPUT


def test_cancellation_between_polls():
    spec, vel, dt = config1(32, seed=4)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    ws = R.PressureWorkspace(spec)
    calls = []

    def cancel():
        calls.append(1)
        return len(calls) >= 3

    rep = ws.solve_rhs(1.0, rhs, dt, jpcg_settings(relative_residual=0.0, absolute_residual=0.0, max_iterations=1000,
                                                   poll_interval=4), cancelled=cancel)
    assert rep.status == R.CANCELLED
    assert 0 < rep.error_iteration < 1000


def test_f32_jacobi_iterates_within_tolerance():
    """f32 Krylov storage, f64 scalars: iterate k within 1e-5 relative of the
    f64 reference iterate (the stated FP32 tolerance)."""
    spec, vel, dt = config1(64, seed=5, dtype=np.float32)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ws = R.PressureWorkspace(spec, "jacobi-pcg-v1", R.F32)
    worst = 0.0
    for k in (1, 5, 20, 50, 100):
        s = jpcg_settings(relative_residual=0.0, absolute_residual=0.0, max_iterations=k)
        ws.solve_rhs(1.0, rhs, dt, s)
        ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(s), threads=8)
        e = rel_err(ws.get_pressure(), ref.pressure)
        worst = max(worst, e)
        assert e < 1e-5, (k, e)


def test_rbsgs_and_jacobi_reach_the_same_pressure():
    # tests/implementation_contract.rs:143-165 analogue: both methods converge
    # to the same gauge-fixed pressure at a tight tolerance.
    spec, vel, dt = config1(32, seed=6)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    s = jpcg_settings(relative_residual=1e-12, divergence_limit=1e-6)
    ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(s))
    out = {}
    for m in ("jacobi-pcg-v1", "rbsgs-pcg"):
        ws = R.PressureWorkspace(spec, m, R.F64)
        rep = ws.solve_rhs(1.0, rhs, dt, s)
        assert rep.ok, m
        out[m] = (rep.iterations, ws.get_pressure())
    assert out["rbsgs-pcg"][0] < out["jacobi-pcg-v1"][0]
    for m, (_, p) in out.items():
        assert rel_err(p, ref.pressure) < 1e-9, m


def test_provider_contract_and_workspace_qualification():
    spec, vel, dt = config1(32, seed=7)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    req = R.PressureRequest(spec, 1.0, dt, rhs)
    prov = R.GpuPressureProvider("rbsgs-pcg", R.F64, R.PressureSettings(1e-10, 1e-12, 1e-7, 2000),
                                 shape=(32, 32, 32))
    meta = prov.metadata()
    assert meta["stage"] == "pressure" and meta["shape_xyz"] == [32, 32, 32]
    # a provider for any shape says so; one for another shape declines the request
    assert R.GpuPressureProvider("jacobi-pcg-v1").metadata()["shape_xyz"] == [0, 0, 0]
    with pytest.raises(R.ProviderRejection) as e:
        R.GpuPressureProvider("jacobi-pcg-v1", shape=(16, 32, 32)).supports(req)
    assert e.value.status == R.UNSUPPORTED
    assert meta["layout"] == "x-fastest-cell-centered" and meta["scalar"] == "f64"
    prov.supports(req)
    out = np.full(spec.cell_len, np.nan)
    prov.predict(req, out)
    assert np.all(np.isfinite(out)) and out[0] == 0.0
    settings = R.PressureSettings(1e-9, 1e-12, 1e-7, 2000)
    ws = R.ProviderWorkspace(spec)
    rep = ws.solve(req, settings, prov)
    assert rep.selected == "rbsgs-pcg" and rep.rejection == R.OK and rep.candidate is not None
    assert rep.attempted == meta
    assert rep.accepted.true_residual_l2 <= rep.candidate.threshold
    # a provider with a looser stopping rule fails the host quality gate and
    # the native device jacobi-pcg-v1 fallback answers, as the reference's.
    loose = R.GpuPressureProvider("jacobi-pcg-v1", R.F64, R.PressureSettings(1e-3, 1e-12, 1.0, 2000))
    rep = ws.solve(req, settings, loose)
    assert rep.rejection == R.QUALITY_GATE and rep.selected == "jacobi-pcg-v1"
    assert rep.attempted == loose.metadata()
    ref = O.solve_rhs(g, 1.0, rhs, dt, osettings(settings))
    assert rep.accepted.iterations == ref.iterations
    assert bits(ws.pressure()) == bits(ref.pressure)
    # unsupported request: wrong length
    with pytest.raises(R.ProviderRejection) as e:
        prov.supports(R.PressureRequest(spec, 1.0, dt, rhs[:-1]))
    assert e.value.status == R.UNSUPPORTED
    # no provider: straight fallback
    rep = ws.solve(req, settings, None)
    assert rep.attempted is None and rep.selected == "jacobi-pcg-v1"


def test_fourier_field_projection_is_divergence_free():
    spec = GridSpec.unit_box(48)
    vel = fourier_modes(spec, 11, dtype=np.float32)
    dt = 0.5 / 48
    s = R.PressureSettings(1e-10, 1e-12, 1e-7, 4000)
    ws = R.PressureWorkspace(spec, "rbsgs-pcg", R.F64)
    rep = ws.solve_velocity(1.0, vel, dt, s)
    assert rep.ok
    out = R.correct_velocity(spec, 1.0, ws.get_pressure(), dt, vel)
    assert R.divergence_max(spec, out) < 1e-3
