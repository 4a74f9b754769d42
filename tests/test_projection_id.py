"""ProjectionRequest::new on the host (src/solver_provider/composition/
projection.rs:58-131): input validation and the SHA-256 input id, through the
C ABI (host code only, no device).  The id is checked against hashlib over
the reference's byte sequence, and the binding test of the reference
(projection_tests.rs:354-418) is restated: every input fact and every face
bit changes the id."""
import hashlib
import math
import struct

import numpy as np
import pytest

import rheon_amd as R
from rheon_amd.scenes import GridSpec

CONTRACT = "sealed-mac-projection-si-f32-faces-f64-cell-zero-gauge-v1"
OPERATOR = "sealed-constant-density-mac-cell-zero-gauge-v1"


def nonzero_fixture(counts=(4, 3, 2), spacing=(0.25, 0.3, 0.4), origin=(0.0, 0.0, 0.0)):
    """projection_tests.rs:19-38: 0.001 sin(0.3 index) on interior faces."""
    spec = GridSpec(counts, spacing, origin)
    vel = []
    for d in range(3):
        f = np.zeros(spec.face_len(d), np.float32)
        m = spec.interior_mask(d).ravel()
        idx = np.arange(f.size)
        f[m] = (0.001 * np.sin(0.3 * idx[m])).astype(np.float32)
        vel.append(f)
    return spec, vel


def make_input(spec, vel, **kw):
    base = dict(density=1.25, interval_start=0.0, dt=0.01, run_id=101, source_revision="c1b1-frozen-reference",
                generation=0, settings=R.PressureSettings(1e-10, 1e-12, 1e-8, 1000), actual_divergence_limit=1e-5)
    base.update(kw)
    return R.ProjectionInput(spec, vel, **base)


def expected_id(inp: R.ProjectionInput) -> bytes:
    """hashlib over the byte sequence of projection.rs:87-126."""
    h = hashlib.sha256()

    def text(s):
        b = s.encode()
        h.update(struct.pack("<Q", len(b)))
        h.update(b)

    spec = inp.spec
    text(CONTRACT)
    h.update(struct.pack("<I", 1))
    text(OPERATOR)
    metas = [(tuple(spec.counts), "x-fastest-cell-centered", "f64", "Pa")]
    for d, layout in enumerate(("x-fastest-x-face", "x-fastest-y-face", "x-fastest-z-face")):
        metas.append((spec.face_counts(d), layout, "f32", "m/s"))
    for shape, layout, scalar, units in metas:
        h.update(struct.pack("<3Q", *shape))
        text(layout)
        text(scalar)
        text(units)
    h.update(struct.pack("<3Q", *spec.counts))
    h.update(struct.pack("<3d", *spec.spacing))
    h.update(struct.pack("<3d", *spec.origin))
    h.update(struct.pack("<Q", inp.run_id))
    text(inp.source_revision)
    h.update(struct.pack("<Q", inp.generation))
    st = inp.settings
    h.update(struct.pack("<7d", inp.density, inp.interval_start, inp.dt, st.relative_residual, st.absolute_residual,
                         st.divergence_limit, inp.actual_divergence_limit))
    h.update(struct.pack("<Q", st.max_iterations))
    for v in inp.velocity:
        h.update(struct.pack("<Q", v.size))
        h.update(np.ascontiguousarray(v, "<f4").tobytes())
    return h.digest()


def test_input_id_is_sha256_of_the_reference_byte_sequence():
    spec, vel = nonzero_fixture()
    inp = make_input(spec, vel)
    assert R.projection_input_id(inp) == expected_id(inp)
    spec2, vel2 = nonzero_fixture((5, 3, 2))
    inp2 = make_input(spec2, vel2, run_id=7, source_revision="révision-中", generation=3)
    assert R.projection_input_id(inp2) == expected_id(inp2)


def test_input_id_binds_every_input_fact_and_face_bit():
    spec, vel = nonzero_fixture()
    base = make_input(spec, vel)
    ids = {R.projection_input_id(base)}
    st = base.settings
    changes = [dict(dt=0.02), dict(interval_start=0.125), dict(generation=1), dict(density=2.5),
               dict(settings=R.PressureSettings(2e-10, st.absolute_residual, st.divergence_limit, 1000)),
               dict(settings=R.PressureSettings(1e-10, 2e-12, st.divergence_limit, 1000)),
               dict(settings=R.PressureSettings(1e-10, 1e-12, 2e-8, 1000)),
               dict(settings=R.PressureSettings(1e-10, 1e-12, 1e-8, 1001)),
               dict(actual_divergence_limit=2e-5), dict(run_id=102),
               dict(source_revision="another-declared-revision")]
    for kw in changes:
        ids.add(R.projection_input_id(make_input(spec, vel, **kw)))
    for other in (GridSpec((4, 3, 2), (0.25, 0.3, 0.4), (0.1, 0.0, 0.0)), GridSpec((4, 3, 2), (0.26, 0.3, 0.4))):
        ids.add(R.projection_input_id(make_input(other, vel)))
    one = [v.copy() for v in vel]
    i = 1  # face (1, 0, 0) of the x lattice
    one[0][i] = np.frombuffer(struct.pack("<I", struct.unpack("<I", one[0][i].tobytes())[0] + 1), np.float32)[0]
    ids.add(R.projection_input_id(make_input(spec, one)))
    negz = [v.copy() for v in vel]
    negz[0][0] = -0.0  # a signed zero is a distinct actual input
    ids.add(R.projection_input_id(make_input(spec, negz)))
    spec5, vel5 = nonzero_fixture((5, 3, 2))
    ids.add(R.projection_input_id(make_input(spec5, [np.zeros_like(v) for v in vel5])))
    assert len(ids) == 2 + len(changes) + 4
    assert R.projection_input_id(base) == R.projection_input_id(make_input(spec, vel))


@pytest.mark.parametrize("kw,status", [
    (dict(domain=R.DOMAIN_FREE_SURFACE), R.REFUSED),
    (dict(source_revision=""), R.REFUSED),
    (dict(source_revision="bell\x07"), R.REFUSED),
    (dict(source_revision="c1\u0085"), R.REFUSED),
    (dict(source_revision="x" * 4097), R.REFUSED),
    (dict(dt=0.0), R.INVALID_SETTINGS),
    (dict(dt=math.inf), R.INVALID_SETTINGS),
    (dict(interval_start=-1.0), R.INVALID_SETTINGS),
    (dict(interval_start=1e20, dt=1e-10), R.INVALID_SETTINGS),  # end == start
    (dict(interval_start=1.7e308, dt=1e308), R.INVALID_SETTINGS),  # end not finite
    (dict(settings=R.PressureSettings(1e-10, 1e-12, 1e-8, 0)), R.INVALID_SETTINGS),
    (dict(actual_divergence_limit=-1.0), R.INVALID_SETTINGS),
    (dict(density=0.0), R.INVALID_DENSITY),
])
def test_invalid_inputs_are_typed_errors(kw, status):
    # projection_tests.rs:626-706 and the constructor's checks
    spec, vel = nonzero_fixture()
    with pytest.raises(R.RheonError) as e:
        R.projection_input_id(make_input(spec, vel, **kw))
    assert e.value.status == status
    assert R.projection_input_id(make_input(spec, vel, source_revision="x" * 4096))


def test_invalid_faces_are_operator_errors():
    spec, vel = nonzero_fixture()
    bad = [v.copy() for v in vel]
    bad[0][0] = 1.0  # a wall face
    with pytest.raises(R.RheonError) as e:
        R.projection_input_id(make_input(spec, bad))
    assert e.value.status == R.NONZERO_WALL_VELOCITY
    bad = [v.copy() for v in vel]
    bad[1][5] = np.nan
    with pytest.raises(R.RheonError) as e:
        R.projection_input_id(make_input(spec, bad))
    assert e.value.status == R.NONFINITE_INPUT
    with pytest.raises(R.RheonError) as e:
        R.projection_input_id(make_input(spec, [vel[0][:-1], vel[1], vel[2]]))
    assert e.value.status == R.INVALID_ARGUMENT
