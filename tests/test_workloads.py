"""The batched and serving workloads at their stated sizes.

Batched: 64 scenes (axes log-uniform on [32, 128], multiples of 8),
Taylor-Green + noise with seed = scene index, 10 steps of the FP32 multigrid
solve in one graph replay per step with autotuned launch shapes; every scene
equals the looped one-scene path bit for bit, and one scene's last pressure
is held to the oracle's f64 solution of that step.

Serving: 40 requests (axes log-uniform on [32, 192], 10-200 steps, Fourier
fields with 0-3 solid spheres) arriving in a Poisson stream; every request
finishes with max|div u| <= 1e-6 at every step (the step gate), and one
request's final pressure equals the one-scene run of the same request bit for
bit and the oracle's f64 solution of its last step to the stated tolerance."""
import time

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import batch_workload, close_obstacle_faces, serving_requests

from helpers import ogrid

# multigrid pressure (rel 1e-8 defect correction) against the f64 oracle
# solve of the same RHS to rel 1e-12, relative to max|p|
MG_PRESSURE_TOL = 1e-6


def oracle_pressure(spec, faces, dt, solid=None, density=1.0):
    """The oracle's step pressure from accepted faces: advection (obstacle
    faces closed), RHS, f64 jacobi-pcg-v1 (masked with obstacles) to rel 1e-12."""
    g = ogrid(spec)
    adv = O.advect_velocity(g, faces, dt, threads=8)
    st = O.Settings(1e-12, 1e-16, 1.0, 50000)
    if solid is None:
        return O.solve_rhs(g, density, O.build_rhs(g, adv, dt), dt, st, threads=8)
    adv = close_obstacle_faces(spec, solid, adv)
    ob = O.obstacle_mask(g, solid)
    return O.solve_rhs_masked(g, density, ob, O.build_rhs(g, adv, dt), dt, st, threads=8)


@pytest.mark.gpu
def test_batch_64_scenes_10_steps():
    specs, vels, dts = batch_workload(64)
    st = R.PressureSettings(1e-8, 1e-14, 1e-3, 40)
    batch = R.SceneBatch(specs, settings=st, use_graph=True, actual_divergence_limit=1e-3, max_cycles=20)
    for i in range(len(specs)):
        batch.scene(i).set_velocity(vels[i])
    batch.set_dt(dts)
    best, table = batch.autotune(steps=1)
    assert len(table) >= 12 and best["seconds_per_step"] == min(t["seconds_per_step"] for t in table)
    for i in range(len(specs)):
        batch.scene(i).set_velocity(vels[i])
    reps = batch.step(10)
    assert all(r.status == R.OK for r in reps), [R.STATUS_NAMES.get(r.status) for r in reps]
    # the smallest scene: its state before the last step, for the oracle
    pick = int(np.argmin([sp.cell_len for sp in specs]))
    before_last = None
    for i, sp in enumerate(specs):
        sc = R.Scene(sp, 1.0, R.F32, "mixed-mg-dc", settings=st, actual_divergence_limit=1e-3)
        sc.set_velocity(vels[i])
        for k in range(10):
            if i == pick and k == 9:
                before_last = sc.get_velocity()
            r = sc.step(dts[i])
            assert r.status == R.OK, (i, k)
        assert np.array_equal(batch.scene(i).get_pressure().view(np.uint64), sc.get_pressure().view(np.uint64)), i
        a, b = batch.scene(i).get_velocity(), sc.get_velocity()
        for c in range(3):
            assert np.array_equal(a[c].view(np.uint32), b[c].view(np.uint32)), (i, c)
        sc.close()
    ref = oracle_pressure(specs[pick], before_last, dts[pick])
    assert ref.ok
    p = batch.scene(pick).get_pressure()
    err = float(np.max(np.abs(p - ref.pressure)) / np.max(np.abs(ref.pressure)))
    assert err <= MG_PRESSURE_TOL, err
    batch.close()


@pytest.mark.gpu
def test_serving_40_requests_poisson():
    reqs = serving_requests(40, 5)
    rng = np.random.default_rng(5)
    gaps = rng.exponential(1.0 / 8.0, size=len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    settings = R.PressureSettings(1e-8, 1e-14, 5e-7, 40)
    server = R.ProjectionServer(settings=settings, actual_divergence_limit=1e-6, hbm_budget=16 << 30,
                                block_divisor=8)
    ids, submitted = [], 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while submitted < len(reqs) and arrivals[submitted] <= now:
            spec, vel, dt, steps, solid, _ = reqs[submitted]
            ids.append(server.submit(spec, vel, dt, steps, solid))
            submitted += 1
        states = [server.state(i)["state"] for i in ids]
        if submitted == len(reqs) and all(s in ("done", "failed") for s in states):
            break
        if not any(s in ("queued", "running") for s in states):
            time.sleep(max(0.0, arrivals[submitted] - now))
            continue
        server.tick()
        assert time.perf_counter() - t0 < 100.0
    st = [server.state(i) for i in ids]
    assert all(s["state"] == "done" for s in st), [(i, R.STATUS_NAMES.get(s["status"])) for i, s in zip(ids, st)]
    assert all(s["steps_done"] == r[3] for s, r in zip(st, reqs))
    assert max(s["actual_divergence_max"] for s in st) <= 1e-6
    # one request (with an obstacle) alone through the one-scene path
    pick = min((i for i, r in enumerate(reqs) if r[4] is not None), key=lambda i: reqs[i][0].cell_len * reqs[i][3])
    spec, vel, dt, steps, solid, _ = reqs[pick]
    faces, p = server.result(ids[pick])
    sc = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc", settings=settings, actual_divergence_limit=1e-6)
    sc.set_solid(solid)
    sc.set_velocity(vel)
    before_last = None
    for k in range(steps):
        if k == steps - 1:
            before_last = sc.get_velocity()
        assert sc.step(dt).status == R.OK, k
    assert np.array_equal(sc.get_pressure().view(np.uint64), p.view(np.uint64))
    got = sc.get_velocity()
    for c in range(3):
        assert np.array_equal(got[c].view(np.uint32), faces[c].view(np.uint32)), c
    ref = oracle_pressure(spec, before_last, dt, solid)
    assert ref.ok
    err = float(np.max(np.abs(p - ref.pressure)) / np.max(np.abs(ref.pressure)))
    assert err <= MG_PRESSURE_TOL, err
    sc.close()
    server.close()
