"""CPU checks of the C ABI library and the host-side scene recipes (no GPU calls)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd import _native as N
from rheon_amd.scenes import GridSpec, fourier_modes, lcg_faces, lcg_states, taylor_green_noise

from helpers import ogrid


def test_library_exports_every_header_symbol():
    lib = N.lib()
    declared = N.header_symbols()
    assert len(declared) >= 30
    assert [s for s in declared if not hasattr(lib, s)] == []
    assert set(declared) == set(N.SIGNATURES)
    assert lib.rh_abi_version() == 2


@pytest.mark.parametrize("counts,spacing,origin,ok", [
    ((64, 64, 64), (1 / 64,) * 3, (0.0,) * 3, True),
    ((3, 2, 4), (0.2, 0.3, 0.4), (-1.0, 2.0, 0.0), True),
    ((0, 1, 1), (1.0,) * 3, (0.0,) * 3, False),
    ((1, 1, 1), (math.nan, 1.0, 1.0), (0.0,) * 3, False),
    ((2, 2, 2), (1.0,) * 3, (1e30, 0.0, 0.0), False),
])
def test_grid_validation_matches_oracle(counts, spacing, origin, ok):
    g = R.make_grid(GridSpec(counts, spacing, origin))
    n = C.c_uint64()
    st = N.lib().rh_grid_validate(C.byref(g), C.byref(n))
    assert (st == R.OK) == ok
    try:
        O.Grid(counts, spacing, origin)
        assert ok
    except O.OracleError:
        assert not ok


def test_buffer_plan_matches_reference_numbers():
    # tests/geometry_contract.rs:60-73
    g = R.make_grid(GridSpec.unit_box(64))
    v = [C.c_uint64() for _ in range(4)]
    assert N.lib().rh_buffer_plan(C.byref(g), *[C.byref(x) for x in v]) == R.OK
    assert [x.value for x in v] == [6_389_760, 2_097_152, 12_582_912, 21_069_824]


def test_lcg_matches_reference_recurrence():
    # tests/pressure_contract.rs:201-210 written as a plain loop
    seed = 0x1234_5678
    s = seed
    expect = []
    for _ in range(1000):
        s = (s * 6364136223846793005 + 1) % (1 << 64)
        expect.append(s)
    assert lcg_states(seed, 1000).tolist() == expect


def test_lcg_faces_order_and_walls():
    spec = GridSpec((8, 7, 6), (0.2, 0.3, 0.4))
    f = lcg_faces(spec, 0x1234_5678)
    s = 0x1234_5678
    for d in range(3):
        fx, fy, fz = spec.face_counts(d)
        for k in range(fz):
            for j in range(fy):
                for i in range(fx):
                    p = (i, j, k)
                    idx = i + fx * (j + fy * k)
                    if 0 < p[d] < spec.counts[d]:
                        s = (s * 6364136223846793005 + 1) % (1 << 64)
                        assert f[d][idx] == np.float32((s >> 40) / float(1 << 24) - 0.5)
                    else:
                        assert f[d][idx] == 0.0


@pytest.mark.parametrize("maker", ["tg", "fourier"])
def test_scene_fields_have_zero_walls_and_are_finite(maker):
    spec = GridSpec.unit_box(24, 16, 8)
    vel = taylor_green_noise(spec, 3) if maker == "tg" else fourier_modes(spec, 3)
    g = ogrid(spec)
    # build_rhs validates finiteness and zero wall-normal faces
    O.build_rhs(g, vel, 0.01)
    assert max(float(np.abs(v).max()) for v in vel) > 0.1


def test_taylor_green_without_noise_is_nearly_divergence_free():
    spec = GridSpec.unit_box(32)
    vel = taylor_green_noise(spec, 0, noise=0.0, dtype=np.float64)
    div = O.actual_divergence_max(ogrid(spec), vel)
    # MAC sampling of sin/cos: div = O(h^2) truncation only, not O(1/h)
    assert div < 0.05 * 32
