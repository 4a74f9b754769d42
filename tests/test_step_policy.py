"""The carrier step's time-step policy, Courant gate and passive tracer on the
device (src/simulation.rs:926-979, :1268-1331, src/advection.rs:72-162),
checked against the oracle through the C ABI."""
import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec, centred_sphere, courant_dt, taylor_green_noise

from helpers import ogrid


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64).tobytes()


def blob(spec: GridSpec, centre, radius, peak=0.9):
    """A tracer blob: smooth bump inside `radius` of `centre`, exactly 0 outside."""
    n = spec.counts
    i = np.arange(spec.cell_len)
    x = [((i // m) % c + 0.5) * h for m, c, h in zip((1, n[0], n[0] * n[1]), n, spec.spacing)]
    r2 = sum((xd - cd) ** 2 for xd, cd in zip(x, centre))
    t = np.where(r2 < radius ** 2, peak * (1.0 - r2 / radius ** 2), 0.0)
    return t.astype(np.float32)


SPECS = [GridSpec.unit_box(32), GridSpec((40, 24, 16), (1 / 40,) * 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("which", range(len(SPECS)))
def test_tracer_step_matches_oracle(which):
    """Tracer advected with the corrected candidate faces, bit for bit (the
    first spec is a lattice grid, the second is not); cells outside the
    displaced support are exactly 0; integral and kinetic energy agree with
    the reference's serial sums to rounding."""
    spec = SPECS[which]
    vel = taylor_green_noise(spec, seed=3 + which)
    dt = courant_dt(spec, vel, 0.3)  # dt * rate about 0.6, under max_courant = 1
    sc = R.Scene(spec, 1.25, R.F32, "jacobi-pcg-v1", solver_precision=R.F64, settings=R.PressureSettings(1e-9, 1e-14, 1e-4, 3000),
                 actual_divergence_limit=1e-4, max_courant=1.0)
    sc.set_velocity(vel)
    t0 = blob(spec, (0.3, 0.4, 0.25), 0.15)
    sc.set_tracer(t0)
    g = ogrid(spec)
    for step in range(3):
        old = sc.get_tracer()
        rep = sc.step(dt)
        assert rep.status == R.OK, R.STATUS_NAMES.get(rep.status)
        cand = sc.get_velocity()  # committed = corrected candidate faces
        want = O.advect_tracer(g, old, cand, dt)
        got = sc.get_tracer()
        assert bits(got) == bits(want), step
        vol = spec.spacing[0] * spec.spacing[1] * spec.spacing[2]
        serial = float(np.cumsum(want.astype(np.float64))[-1]) * vol
        assert rep.tracer_integral == pytest.approx(serial, rel=1e-12, abs=1e-300)
        energy = float(np.cumsum(np.concatenate([c.astype(np.float64) ** 2 for c in cand]))[-1])
        assert rep.kinetic_energy == pytest.approx(0.5 * 1.25 * vol * energy, rel=1e-12)
        assert rep.generation == step + 1
        assert rep.time == pytest.approx(dt * (step + 1), rel=1e-15)
    # the blob stays local: far cells are exact +0.0
    assert np.count_nonzero(got) < spec.cell_len
    assert not np.any(np.signbit(got[got == 0]))
    sc.close()


@pytest.mark.gpu
def test_tracer_out_of_range_rejects_step():
    """A candidate tracer value above 1 fails the step with ArithmeticFailure
    (src/simulation.rs:1322-1331) and nothing is committed."""
    spec = SPECS[0]
    vel = taylor_green_noise(spec, seed=8)
    dt = courant_dt(spec, vel, 0.5)
    sc = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64, settings=R.PressureSettings(1e-9, 1e-14, 1e-4, 3000),
                 actual_divergence_limit=1e-4)
    sc.set_velocity(vel)
    t0 = np.full(spec.cell_len, 1.5, np.float32)
    sc.set_tracer(t0)
    rep = sc.step(dt)
    assert rep.status == R.ARITHMETIC_FAILURE
    assert bits(sc.get_tracer()) == bits(t0)
    got = sc.get_velocity()
    assert all(bits(a) == bits(b) for a, b in zip(got, vel))
    sc.close()


@pytest.mark.gpu
def test_courant_gate_one_scene_and_batched():
    """dt * rate(candidate) above max_courant: CourantLimit, accepted faces
    unchanged (src/simulation.rs:1268-1277), one scene and batched."""
    spec = GridSpec.unit_box(32)
    vel = taylor_green_noise(spec, seed=4)
    # courant_dt(0.3) bounds each axis's displacement by 0.3 cells; the rate
    # sums three axes, so dt * rate is about 0.6
    dt = courant_dt(spec, vel, 0.3)
    st = R.PressureSettings(1e-8, 1e-14, 1e-3, 40)
    sc = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=st, actual_divergence_limit=1e-3, max_courant=0.2)
    sc.set_velocity(vel)
    rep = sc.step(dt)
    assert rep.status == R.COURANT_LIMIT
    rate = O.velocity_rate(ogrid(spec), vel)
    assert rep.courant > 0.2 and rep.courant == pytest.approx(dt * rate, rel=0.05)
    assert all(bits(a) == bits(b) for a, b in zip(sc.get_velocity(), vel))
    # the same scene passes with a gate above its Courant number
    ok = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=st, actual_divergence_limit=1e-3, max_courant=1.0)
    ok.set_velocity(vel)
    r2 = ok.step(dt)
    assert r2.status == R.OK and r2.courant == rep.courant
    batch = R.SceneBatch([spec, spec], settings=st, use_graph=True, actual_divergence_limit=1e-3, max_courant=0.2)
    batch.scene(0).set_velocity(vel)
    slow = [0.1 * v for v in vel]
    batch.scene(1).set_velocity(slow)
    batch.set_dt([dt, dt])
    reps = batch.step(1)
    assert reps[0].status == R.COURANT_LIMIT and reps[0].courant == rep.courant
    assert reps[1].status == R.OK
    assert all(bits(a) == bits(b) for a, b in zip(batch.scene(0).get_velocity(), vel))
    assert not all(bits(a) == bits(b) for a, b in zip(batch.scene(1).get_velocity(), slow))
    batch.close()
    sc.close()
    ok.close()


@pytest.mark.gpu
def test_requested_dt_policy():
    """dt = min(requested, 0.5 max_courant / rate) with the rate of the
    accepted faces (oracle velocity_rate); the next step reuses the rate the
    Courant gate measured on the faces it committed."""
    spec = GridSpec((24, 16, 40), (1 / 16, 1 / 32, 1 / 8), (-0.5, 0.25, 1.0))
    vel = taylor_green_noise(spec, seed=12)
    g = ogrid(spec)
    sc = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64, settings=R.PressureSettings(1e-9, 1e-14, 1e-3, 5000),
                 actual_divergence_limit=1e-3, max_courant=0.8)
    sc.set_velocity(vel)
    for step in range(3):
        acc = sc.get_velocity()
        rate = O.velocity_rate(g, acc)
        want = min(10.0, 0.5 * 0.8 / rate)
        rep = sc.step_requested(10.0)
        assert rep.status == R.OK, R.STATUS_NAMES.get(rep.status)
        assert rep.dt == want, step
    # a requested dt below the Courant bound is taken as is
    rep = sc.step_requested(1e-4)
    assert rep.status == R.OK and rep.dt == 1e-4
    # max_courant outside (0, 1] is an invalid configuration for the policy
    bad = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", max_courant=0.0)
    bad.set_velocity(vel)
    assert bad.step_requested(0.1).status == R.INVALID_SETTINGS
    bad.close()
    sc.close()


@pytest.mark.gpu
def test_tracer_with_sphere_and_dense_support():
    """Tracer around a solid sphere, with a blob that covers the box (the
    support is then the whole box): bit for bit against the oracle."""
    spec = GridSpec.unit_box(32)
    solid = centred_sphere(spec, 0.2)
    vel = taylor_green_noise(spec, seed=9)
    dt = courant_dt(spec, vel, 0.5)
    sc = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=R.PressureSettings(1e-10, 1e-14, 1e-3, 40),
                 actual_divergence_limit=1e-3)
    sc.set_solid(solid)
    sc.set_velocity(vel)
    t0 = blob(spec, (0.5, 0.5, 0.5), 0.9, 0.5)
    sc.set_tracer(t0)
    old = sc.get_tracer()
    rep = sc.step(dt)
    assert rep.status == R.OK
    want = O.advect_tracer(ogrid(spec), old, sc.get_velocity(), dt)
    assert bits(sc.get_tracer()) == bits(want)
    sc.close()


@pytest.mark.gpu
def test_rhs_failure_after_a_committed_step():
    """A candidate whose RHS overflows (src/operator.rs:460-509: a non-finite
    b is ArithmeticFailure) fails the step with the same status whether the
    candidate's check is read before the solve (faces handed out through
    device_velocity) or with the step's last read-back (faces the scene
    committed itself), and nothing is committed either way."""
    spec = GridSpec.unit_box(32)
    vel = [np.ascontiguousarray(v * np.float32(1e30)) for v in taylor_green_noise(spec, seed=12)]
    st = R.PressureSettings(1e-8, 1e-14, 1e30, 60)
    statuses = []
    for external in (False, True):
        sc = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=st, actual_divergence_limit=1e30, max_courant=0.0)
        sc.set_velocity(vel)
        first = sc.step(1e-3)
        assert first.status == R.OK
        before = sc.get_velocity()
        if external:
            R._native.lib().rh_scene_device_velocity(sc.handle, *(R._native.C.byref(R._native.C.c_void_p())
                                                                  for _ in range(3)))
        rep = sc.step(1e-300)
        statuses.append(rep.status)
        assert rep.status != R.OK
        after = sc.get_velocity()
        assert all(bits(a) == bits(b) for a, b in zip(after, before))
        sc.close()
    assert statuses[0] == statuses[1] == R.ARITHMETIC_FAILURE
