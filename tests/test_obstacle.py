"""Solid obstacles: the voxelised-sphere discretisation of StaticObstaclePressure
(src/obstacle_pressure.rs:224-309 with open areas in {0, A} and fluid volumes
in {0, V}; solid faces stay 0, :566-578).

The reference holds no vector for an obstacle scene, so the oracle's masked
operator is pinned by (a) bit-identity with the box operator when nothing is
solid, (b) an independent face-by-face assembly of the matrix, (c) a direct
dense solve of the gauge-eliminated system, and (d) the invariants of the
operator (constant null space, symmetric preconditioner).  The GPU path is
then held to the oracle: bit-exact for the f64 PCG family, to the converged
tolerance for multigrid.
"""
import numpy as np
import pytest

import oracle as O
from helpers import ogrid, rel_err
from rheon_amd.scenes import (GridSpec, centred_sphere, close_obstacle_faces, courant_dt, fourier_modes,
                              spheres_solid)


def face_matrix(spec: GridSpec, density: float, solid: np.ndarray) -> np.ndarray:
    """Dense K assembled face by face: every interior face between two fluid
    cells adds w = A/(rho h) to both diagonals and -w off the diagonal."""
    nx, ny, nz = spec.counts
    n = nx * ny * nz
    h = spec.spacing
    area = (h[1] * h[2], h[0] * h[2], h[0] * h[1])
    w = [area[a] * (1.0 / (density * h[a])) for a in range(3)]
    stride = (1, nx, nx * ny)
    K = np.zeros((n, n))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                pos = (i, j, k)
                for a in range(3):
                    if pos[a] + 1 >= spec.counts[a]:
                        continue
                    d = c + stride[a]
                    if solid[c] or solid[d]:
                        continue
                    K[c, c] += w[a]
                    K[d, d] += w[a]
                    K[c, d] -= w[a]
                    K[d, c] -= w[a]
    return K


def sphere_case(n=12, radius=0.3, seed=3):
    spec = GridSpec.unit_box(n)
    solid = centred_sphere(spec, radius)
    vel = close_obstacle_faces(spec, solid, fourier_modes(spec, seed, modes=8, kmax=4, dtype=np.float32))
    return spec, solid, vel, courant_dt(spec, vel, 0.5)


def test_sphere_mask_and_gauge():
    spec = GridSpec.unit_box(16)
    solid = centred_sphere(spec, 0.2)
    assert 0 < solid.sum() < solid.size
    ob = O.obstacle_mask(ogrid(spec), solid)
    assert ob.gauge == 0  # a centred sphere leaves cell 0 fluid
    fluid = solid == 0
    assert np.array_equal((ob.mask & 64) != 0, fluid)
    f3 = fluid.reshape(16, 16, 16)
    m3 = ob.mask.reshape(16, 16, 16)
    # lower x face open iff both cells fluid, x = 0 closed (wall)
    assert np.array_equal((m3[:, :, 1:] & 1) != 0, f3[:, :, 1:] & f3[:, :, :-1])
    assert not np.any(m3[:, :, 0] & 1)
    assert np.array_equal((m3[:, 1:, :] & 4) != 0, f3[:, 1:, :] & f3[:, :-1, :])
    assert np.array_equal((m3[:-1, :, :] & 32) != 0, f3[:-1, :, :] & f3[1:, :, :])


def test_apply_masked_without_solids_is_the_box_operator():
    spec = GridSpec.unit_box(9, 7, 5)
    g = ogrid(spec)
    ob = O.obstacle_mask(g, np.zeros(spec.cell_len, np.uint8))
    x = np.random.default_rng(0).standard_normal(spec.cell_len)
    a = O.apply_masked(g, 1.3, ob, x)
    b = O.apply_full(g, 1.3, x)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_apply_masked_matches_face_assembly():
    spec = GridSpec((10, 9, 8), (0.1, 1.0 / 9, 0.125))
    solid = spheres_solid(spec, [(0.5, 0.5, 0.5, 0.3), (0.1, 0.9, 0.2, 0.2)])
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    x = np.random.default_rng(1).standard_normal(spec.cell_len)
    y = O.apply_masked(g, 0.7, ob, x)
    K = face_matrix(spec, 0.7, solid)
    assert rel_err(y, K @ x) < 1e-13
    assert np.all(y[solid == 1] == 0.0)
    # constant null space on the fluid
    assert np.max(np.abs(O.apply_masked(g, 0.7, ob, np.ones(spec.cell_len)))) < 1e-12


@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_masked_solve_without_solids_is_bitwise_the_box_solve(method):
    spec = GridSpec.unit_box(10)
    g = ogrid(spec)
    vel = fourier_modes(spec, 5, modes=6, kmax=4, dtype=np.float64)
    dt = courant_dt(spec, vel)
    rhs = O.build_rhs(g, vel, dt)
    st = O.Settings(1e-10, 1e-14, 1e-6, 500)
    ob = O.obstacle_mask(g, np.zeros(spec.cell_len, np.uint8))
    a = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, st, method)
    b = O.solve_rhs(g, 1.0, rhs, dt, st, method)
    assert a.status == b.status == O.OK
    assert a.iterations == b.iterations
    assert np.array_equal(a.pressure.view(np.uint64), b.pressure.view(np.uint64))


@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_masked_solve_matches_direct_solution(method):
    spec, solid, vel, dt = sphere_case()
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    rhs = O.build_rhs(g, vel, dt)
    assert np.all(rhs[solid == 1] == 0.0)
    res = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, O.Settings(1e-12, 1e-16, 1e-6, 2000), method)
    assert res.status == O.OK, res
    # direct: fluid rows/columns with the gauge eliminated
    K = face_matrix(spec, 1.0, solid)
    keep = np.flatnonzero(solid == 0)
    keep = keep[keep != ob.gauge]
    p = np.zeros(spec.cell_len)
    p[keep] = np.linalg.solve(K[np.ix_(keep, keep)], rhs[keep])
    assert res.pressure[ob.gauge] == 0.0
    assert np.all(res.pressure[solid == 1] == 0.0)
    assert rel_err(res.pressure, p) < 1e-8
    # the correction closes the divergence on the fluid and keeps solid faces 0
    out = O.correct_masked(g, 1.0, ob, res.pressure, dt, vel)
    assert O.actual_divergence_max_masked(g, ob, out) < 1e-4
    closed = close_obstacle_faces(spec, solid, out)
    for a in range(3):
        assert np.array_equal(closed[a], out[a])


def test_masked_rbsgs_preconditioner_is_symmetric():
    spec, solid, _, _ = sphere_case(n=10, radius=0.25)
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    rng = np.random.default_rng(7)
    r1, r2 = (rng.standard_normal(spec.cell_len) * (solid == 0) for _ in range(2))
    r1, z1, _ = O.precondition_masked(g, 1.0, ob, r1, "rbsgs-pcg")
    r2, z2, _ = O.precondition_masked(g, 1.0, ob, r2, "rbsgs-pcg")
    assert abs(r1 @ z2 - r2 @ z1) <= 1e-12 * abs(r1 @ z2)
    assert r1 @ z1 > 0.0
    assert np.all(z1[solid == 1] == 0.0) and z1[ob.gauge] == 0.0


# ---------------------------------------------------------------- GPU path


def oracle_masked_step(spec, solid, vel, dt, method, settings):
    """Oracle carrier with obstacles: advect, close obstacle faces, RHS, masked
    solve, masked correction, fluid divergence."""
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    adv = close_obstacle_faces(spec, solid, O.advect_velocity(g, vel, dt))
    rhs = O.build_rhs(g, adv, dt)
    res = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, settings, method)
    assert res.status == O.OK
    out = O.correct_masked(g, 1.0, ob, res.pressure, dt, adv)
    return res, out, O.actual_divergence_max_masked(g, ob, out)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["jacobi-pcg-v1", "rbsgs-pcg"])
def test_gpu_masked_pcg_bitwise(method):
    import rheon_amd as R
    spec, solid, vel, dt = sphere_case(n=16, radius=0.3)
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    rhs = O.build_rhs(g, vel, dt)
    st = O.Settings(1e-10, 1e-14, 1e-6, 1000)
    ref = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, st, method)
    ws = R.PressureWorkspace(spec, method, R.F64)
    ws.set_solid(solid)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1e-6, 1000))
    assert rep.ok and ref.status == O.OK
    assert rep.iterations == ref.iterations
    p = ws.get_pressure()
    assert np.array_equal(p.view(np.uint64), ref.pressure.view(np.uint64))
    assert rep.true_residual_l2 == ref.true_residual_l2


@pytest.mark.gpu
def test_gpu_scene_step_with_sphere_bitwise():
    import rheon_amd as R
    spec = GridSpec.unit_box(32)
    solid = centred_sphere(spec, 0.2)
    vel = close_obstacle_faces(spec, solid, fourier_modes(spec, 11, dtype=np.float32))
    dt = courant_dt(spec, vel, 0.5)
    st = O.Settings(1e-9, 1e-14, 1e-4, 2000)
    res, out, div = oracle_masked_step(spec, solid, vel, dt, "jacobi-pcg-v1", st)
    sc = R.Scene(spec, 1.0, R.F32, "jacobi-pcg-v1", solver_precision=R.F64,
                 settings=R.PressureSettings(1e-9, 1e-14, 1e-4, 2000), actual_divergence_limit=1e-3)
    sc.set_solid(solid)
    sc.set_velocity(vel)
    rep = sc.step(dt)
    assert rep.status == R.OK, rep
    assert rep.pressure.iterations == res.iterations
    got = sc.get_velocity()
    for a in range(3):
        assert np.array_equal(got[a].view(np.uint32), out[a].view(np.uint32))
    p = sc.get_pressure()
    assert np.array_equal(p.view(np.uint64), res.pressure.view(np.uint64))
    assert rep.actual_divergence_max == div


@pytest.mark.gpu
@pytest.mark.parametrize("n,prec", [(32, "f32"), (64, "f32"), (32, "f64")])
def test_gpu_mg_with_sphere_converges_to_oracle(n, prec):
    import rheon_amd as R
    spec = GridSpec.unit_box(n)
    solid = centred_sphere(spec, 0.2)
    vel = close_obstacle_faces(spec, solid, fourier_modes(spec, 2, dtype=np.float32))
    dt = courant_dt(spec, vel, 0.5)
    g = ogrid(spec)
    ob = O.obstacle_mask(g, solid)
    rhs = O.build_rhs(g, vel, dt)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32 if prec == "f32" else R.F64)
    ws.set_solid(solid)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1e-3, 60))
    assert rep.ok, rep
    assert rep.iterations <= 30, rep.iterations
    p = ws.get_pressure()
    assert p[ob.gauge] == 0.0 and np.all(p[solid == 1] == 0.0)
    ref = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, O.Settings(1e-12, 1e-16, 1e-3, 4000), "rbsgs-pcg", threads=8)
    assert ref.status == O.OK
    assert rel_err(p, ref.pressure) < 1e-7
    # residual of the returned pressure, checked on the host
    r = rhs - O.apply_masked(g, 1.0, ob, p)
    assert np.sqrt(r @ r) <= 1e-10 * np.sqrt(rhs @ rhs) * 1.0001


@pytest.mark.gpu
def test_gpu_more_than_254_fluid_regions_are_rejected():
    """Split fluid is solved region by region (tests/test_components.py); the
    region labels are one byte, so at most 254 regions are accepted."""
    import rheon_amd as R
    spec = GridSpec.unit_box(16)
    solid = np.ones((16, 16, 16), np.uint8)
    solid[::2, ::2, ::2] = 0  # 512 isolated fluid cells
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32)
    with pytest.raises(R.RheonError) as e:
        ws.set_solid(solid.ravel())
    assert e.value.status == R.UNSUPPORTED
    solid = np.ones((16, 16, 16), np.uint8)
    solid[::4, ::4, ::4] = 0  # 64 regions: accepted
    ws.set_solid(solid.ravel())


@pytest.mark.gpu
def test_gpu_scene_closes_obstacle_faces_on_upload():
    import rheon_amd as R
    spec = GridSpec.unit_box(16)
    solid = centred_sphere(spec, 0.3)
    vel = fourier_modes(spec, 4, dtype=np.float32)
    sc = R.Scene(spec, 1.0, R.F32, "mixed-mg-dc")
    sc.set_solid(solid)
    sc.set_velocity(vel)
    got = sc.get_velocity()
    want = close_obstacle_faces(spec, solid, vel)
    for a in range(3):
        assert np.array_equal(got[a], want[a])
