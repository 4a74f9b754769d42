"""The reference's sgs-pcg-v1 on the device: exact lexicographic symmetric
Gauss-Seidel sweeps run as 8^3 tile waves (src/pressure.rs:467-532,
src/sgs.rs:56-225), held to the oracle's serial sweeps (sgs_c) bit for bit:
the preconditioner alone, every PCG iterate, and the reference's published
iteration count for the 32^3 steady scene."""
import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec

from helpers import config1, ogrid
from test_oracle import steady_scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64).tobytes()


def test_sgs_known_answer_on_device():
    # src/pressure.rs:824-842: 2x2x1, r = [-12, 2, 4, 6] -> z = [0, 3.25, 4.25, 4.5], rz = 50.5
    spec = GridSpec((2, 2, 1), (1.0, 1.0, 1.0))
    r, z, rz = R.precondition(spec, 1.0, np.array([-12.0, 2.0, 4.0, 6.0]), "sgs-pcg-v1")
    assert list(z) == [0.0, 3.25, 4.25, 4.5] and rz == 50.5 and r[0] == 0.0


@pytest.mark.parametrize("counts,spacing", [
    ((1, 1, 1), (1.0, 1.0, 1.0)), ((9, 7, 5), (0.2, 0.25, 0.3)), ((33, 17, 5), (0.07, 0.11, 0.19)),
    ((20, 12, 8), (1 / 16,) * 3), ((32, 32, 32), (1 / 32,) * 3), ((3, 40, 9), (0.5, 0.1, 0.2))])
def test_sgs_preconditioner_bits(counts, spacing):
    spec = GridSpec(counts, spacing)
    rng = np.random.default_rng(sum(counts))
    r = rng.standard_normal(spec.cell_len)
    rd, zd, rzd = R.precondition(spec, 1.3, r, "sgs-pcg-v1")
    ro, zo, rzo = O.precondition(ogrid(spec), 1.3, r, "sgs-pcg-v1")
    assert bits(rd) == bits(ro)
    assert bits(zd) == bits(zo)
    assert rzd == rzo


@pytest.mark.parametrize("n", [16, 32])
def test_sgs_pcg_iterates_bits(n):
    spec, vel, dt = config1(n)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ws = R.PressureWorkspace(spec, "sgs-pcg-v1", R.F64)
    for k in (1, 2, 5, 17, 40):
        rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k))
        ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(0.0, 0.0, 0.0, k), method="sgs-pcg-v1")
        assert rep.status == ref.status and rep.iterations == ref.iterations, k
        assert bits(ws.get_pressure()) == bits(ref.pressure), k
    st = R.PressureSettings(1e-9, 1e-12, 1e-7, 2000)
    rep = ws.solve_rhs(1.0, rhs, dt, st)
    ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(1e-9, 1e-12, 1e-7, 2000), method="sgs-pcg-v1")
    assert rep.ok and ref.ok and rep.iterations == ref.iterations
    assert rep.true_residual_l2 == ref.true_residual_l2
    assert bits(ws.get_pressure()) == bits(ref.pressure)
    ws.close()


def test_sgs_steady_scene_reference_count():
    # measurement-tables.md:79: 97 sgs-pcg-v1 iterations per solve of the 32^3 steady scene
    spec, g, vel, dt, acc = steady_scene(32)
    ws = R.PressureWorkspace(spec, "sgs-pcg-v1", R.F64)
    rep = ws.solve_velocity(1.25, vel, dt, R.PressureSettings(1e-8, 1e-12, 1e-7, 3000))
    ref = O.solve_velocity(g, 1.25, vel, dt, O.Settings(1e-8, 1e-12, 1e-7, 3000), method="sgs-pcg-v1")
    assert rep.ok and rep.iterations == 97 == ref.iterations
    assert rep.true_residual_l2 == ref.true_residual_l2
    assert bits(ws.get_pressure()) == bits(ref.pressure)
    ws.close()


def test_sgs_rejects_obstacles():
    spec = GridSpec.unit_box(16)
    ws = R.PressureWorkspace(spec, "sgs-pcg-v1", R.F64)
    solid = np.zeros(spec.cell_len, np.uint8)
    solid[100] = 1
    with pytest.raises(R.RheonError) as e:
        ws.set_solid(solid)
    assert e.value.status == R.UNSUPPORTED
    ws.close()
