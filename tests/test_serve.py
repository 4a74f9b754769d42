"""Serving loop: admission by planned device bytes, continuous batching, and
the same result for a request alone and in a mix."""
import numpy as np
import pytest

from rheon_amd.scenes import close_obstacle_faces, courant_dt, fourier_modes, isotropic_spec, spheres_solid


def request(counts, seed, spheres):
    spec = isotropic_spec(counts)
    solid = spheres_solid(spec, spheres) if spheres else None
    vel = fourier_modes(spec, seed, modes=8, kmax=6, amplitude=0.02, dtype=np.float32)
    if solid is not None:
        vel = close_obstacle_faces(spec, solid, vel)
    return spec, vel, courant_dt(spec, vel, 0.5), solid


A = ((48, 40, 56), 3, [(0.4, 0.35, 0.5, 0.15)])
B = ((64, 32, 40), 5, [])
C_ = ((40, 56, 48), 8, [(0.3, 0.4, 0.5, 0.1), (0.7, 0.5, 0.6, 0.12)])
D = ((32, 32, 72), 9, [(0.2, 0.2, 0.5, 0.1)])


def run_until_done(server, ids, max_ticks=200):
    ticks = []
    for _ in range(max_ticks):
        if all(server.state(i)["state"] in ("done", "failed") for i in ids):
            return ticks
        ticks.append(server.tick())
    raise AssertionError("requests did not finish")


@pytest.mark.gpu
def test_result_same_alone_and_in_mix():
    import rheon_amd as R
    spec, vel, dt, solid = request(*A)
    alone = R.ProjectionServer()
    a = alone.submit(spec, vel, dt, 6, solid)
    run_until_done(alone, [a])
    assert alone.state(a)["state"] == "done"
    ref_vel, ref_p = alone.result(a)
    alone.close()

    mix = R.ProjectionServer()
    others = []
    for req, steps in ((B, 9), (C_, 4)):
        s, v, d, so = request(*req)
        others.append(mix.submit(s, v, d, steps, so))
    mix.tick()
    mix.tick()
    a2 = mix.submit(spec, vel, dt, 6, solid)  # arrives while others run
    s, v, d, so = request(*D)
    others.append(mix.submit(s, v, d, 7, so))
    run_until_done(mix, [a2] + others)
    for i in [a2] + others:
        assert mix.state(i)["state"] == "done", mix.state(i)
    got_vel, got_p = mix.result(a2)
    for k in range(3):
        assert np.array_equal(got_vel[k].view(np.uint32), ref_vel[k].view(np.uint32))
    assert np.array_equal(got_p.view(np.uint64), ref_p.view(np.uint64))
    mix.close()


@pytest.mark.gpu
def test_admission_respects_the_memory_budget():
    import rheon_amd as R
    reqs = [request(*A), request(*B), request(*C_)]
    probe = R.ProjectionServer()
    plans = [probe.plan_bytes(r[0]) for r in reqs]
    probe.close()
    budget = plans[0] + plans[1] + plans[2] // 2  # the third waits for a slot
    server = R.ProjectionServer(hbm_budget=budget)
    ids = [server.submit(s, v, d, 3 + i, so) for i, (s, v, d, so) in enumerate(reqs)]
    t = server.tick()
    assert t["admitted"] == 2 and t["queued"] == 1
    assert server.state(ids[2])["state"] == "queued"
    ticks = run_until_done(server, ids)
    assert all(tk["bytes_in_use"] <= budget for tk in ticks)
    assert all(server.state(i)["state"] == "done" for i in ids)
    assert server.state(ids[2])["admitted"] >= server.state(ids[0])["finished"]
    server.close()


@pytest.mark.gpu
def test_failed_request_does_not_disturb_others():
    import rheon_amd as R
    server = R.ProjectionServer()
    s, v, d, so = request(*B)
    bad = [x.copy() for x in v]
    bad[1][s.face_counts(1)[0] * 3 + 2] = np.inf
    good = server.submit(s, v, d, 3, so)
    broken = server.submit(s, bad, d, 3, so)
    run_until_done(server, [good, broken])
    assert server.state(good)["state"] == "done"
    st = server.state(broken)
    assert st["state"] == "failed" and st["status"] == R.NONFINITE_INPUT
    server.close()
