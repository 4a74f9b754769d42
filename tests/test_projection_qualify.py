"""ProjectionWorkspace::qualify on the device, checked against the oracle
(src/solver_provider/composition/projection.rs:350-542; the cases restate
projection_tests.rs).  "direct" is the oracle's own projection of the input:
RHS, the native pressure solve, the f32 correction and the actual divergence.
Device native results must equal it bit for bit; a supplied candidate is
accepted only when its velocity is the bit-exact correction of the input by
its pressure (the momentum witness) and both pass requalification."""
import struct

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec

from helpers import ogrid
from test_projection_id import make_input, nonzero_fixture

pytestmark = pytest.mark.gpu

METHODS = ["jacobi-pcg-v1", "rbsgs-pcg"]


def zero_fixture(value=0.0):
    """projection_tests.rs:39-43: 4^3 cells of 0.25, every face `value`."""
    spec = GridSpec((4, 4, 4), (0.25,) * 3)
    return spec, [np.full(spec.face_len(d), value, np.float32) for d in range(3)]


def u32(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def u64(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def osettings(inp):
    s = inp.settings
    return O.Settings(s.relative_residual, s.absolute_residual, s.divergence_limit, s.max_iterations)


def corrected(inp, pressure):
    return O.correct(ogrid(inp.spec), inp.density, np.asarray(pressure, np.float64), inp.dt, inp.velocity)


def direct(inp, method):
    g = ogrid(inp.spec)
    rhs = O.build_rhs(g, inp.velocity, inp.dt)
    res = O.solve_rhs(g, inp.density, rhs, inp.dt, osettings(inp), method=method)
    assert res.ok
    vel = corrected(inp, res.pressure)
    return dict(pressure=res.pressure, velocity=vel, report=res, div=O.actual_divergence_max(g, vel))


def candidate_of(inp, pressure, velocity, **kw):
    return R.ProjectionCandidate(R.projection_input_id(inp), np.asarray(pressure, np.float64), velocity,
                                 algorithm="classical-supplied-projection", version="1",
                                 source_revision="c1b1-direct-existing-dispatcher", **kw)


def assert_native(ws, rep, d):
    p, vel = ws.result()
    assert np.array_equal(u64(p), u64(d["pressure"]))
    for a in range(3):
        assert np.array_equal(u32(vel[a]), u32(d["velocity"][a])), a
    acc = rep["pressure"].accepted
    r = d["report"]
    assert (acc.true_residual_l2, acc.true_residual_max, acc.predicted_divergence_max) == \
        (r.true_residual_l2, r.true_residual_max, r.predicted_divergence_max)
    assert acc.iterations == r.iterations
    assert rep["actual_divergence_max"] == d["div"]


@pytest.mark.parametrize("method", METHODS)
def test_native_and_supplied_match_the_oracle_for_independent_requests(method):
    spec, vel = nonzero_fixture()
    base = make_input(spec, vel)
    d = direct(base, method)
    assert d["report"].iterations > 0
    ws = R.ProjectionWorkspace(spec, 1 << 20, method)
    for generation in range(20):
        inp = make_input(spec, vel, generation=generation, interval_start=generation * 0.01)
        st, rep = ws.qualify(inp, None)
        assert st == R.OK
        assert_native(ws, rep, d)
        assert rep["selection"] == "native_baseline" and rep["rejection"] is None
        assert rep["input_id"] == R.projection_input_id(inp)
        assert rep["pressure"].selected == method and rep["pressure"].attempted is None
        cand = candidate_of(inp, d["pressure"], d["velocity"])
        st, rep = ws.qualify(inp, cand)
        assert st == R.OK
        p, out = ws.result()
        assert np.array_equal(u64(p), u64(d["pressure"]))
        assert all(np.array_equal(u32(out[a]), u32(d["velocity"][a])) for a in range(3))
        assert rep["selection"] == "supplied" and rep["rejection"] is None and rep["supplied_identity"]
        assert rep["pressure"].accepted.iterations == 0
        acc, r = rep["pressure"].accepted, d["report"]
        assert (acc.true_residual_l2, acc.true_residual_max, acc.predicted_divergence_max) == \
            (r.true_residual_l2, r.true_residual_max, r.predicted_divergence_max)
        assert rep["actual_divergence_max"] == d["div"]
        bridge = rep["pressure"].attempted
        assert bridge["id"] == "supplied-projection-pressure-witness-v1" == rep["pressure"].selected
    ws.close()


@pytest.mark.parametrize("method", METHODS)
def test_divergence_free_wrong_dynamics_fail_the_momentum_witness(method):
    spec, vel = zero_fixture(0.0)
    inp = make_input(spec, vel)
    d = direct(inp, method)
    wrong = [v.copy() for v in d["velocity"]]
    fx = [spec.face_counts(a) for a in range(3)]

    def face(a, i, j, k):
        return i + fx[a][0] * (j + fx[a][1] * k)

    for a, (i, j, k), value in [(0, (2, 1, 1), 0.125), (1, (2, 2, 1), 0.125), (0, (2, 2, 1), -0.125),
                                (1, (1, 2, 1), -0.125)]:
        wrong[a][face(a, i, j, k)] = value
    assert O.actual_divergence_max(ogrid(spec), wrong) == 0.0
    ws = R.ProjectionWorkspace(spec, 1 << 20, method)
    st, rep = ws.qualify(inp, candidate_of(inp, d["pressure"], wrong))
    assert st == R.OK
    assert_native(ws, rep, d)
    assert rep["selection"] == "native_fallback"
    assert rep["rejection"] == "momentum_witness" and rep["witness"] == (0, face(0, 2, 1, 1))
    assert rep["pressure"].attempted is None
    ws.close()


def test_signed_zero_correction_and_wall_bits():
    spec, vel = zero_fixture(-0.0)
    inp = make_input(spec, vel)
    d = direct(inp, "jacobi-pcg-v1")
    p = d["pressure"].copy()
    p[0] = -0.0
    good = corrected(inp, p)
    assert u32(good[0])[0] == u32(np.float32(-0.0))
    ws = R.ProjectionWorkspace(spec, 1 << 20)
    st, rep = ws.qualify(inp, candidate_of(inp, p, good))
    assert st == R.OK and rep["selection"] == "supplied"
    pr, out = ws.result()
    assert np.array_equal(u64(pr), u64(p))
    assert all(np.array_equal(u32(out[a]), u32(good[a])) for a in range(3))
    wrong = [v.copy() for v in good]
    wrong[0][0] = 0.0
    st, rep = ws.qualify(inp, candidate_of(inp, p, wrong))
    assert st == R.OK and rep["selection"] == "native_fallback"
    assert rep["rejection"] == "momentum_witness" and rep["witness"] == (0, 0)
    assert u32(ws.result()[1][0])[0] == u32(np.float32(-0.0))
    ws.close()


@pytest.mark.parametrize("method", METHODS)
def test_candidate_numerical_failures_keep_typed_reasons(method):
    spec, vel = nonzero_fixture()
    inp = make_input(spec, vel)
    d = direct(inp, method)
    interior = 1  # face (1, 0, 0) of the x lattice
    ws = R.ProjectionWorkspace(spec, 1 << 20, method)
    for case in range(8):
        p = d["pressure"].copy()
        v = [x.copy() for x in d["velocity"]]
        if case == 0:
            p = p[:-1]
            want = ("field", R.INVALID_ARGUMENT)
        elif case == 1:
            p[1] = np.nan
            want = ("field", R.NONFINITE_INPUT)
        elif case == 2:
            v[0] = v[0][:-1]
            want = ("field", R.INVALID_ARGUMENT)
        elif case == 3:
            v[0][interior] = np.nan
            want = ("field", R.NONFINITE_INPUT)
        elif case == 4:
            v[0][0] = 1.0
            want = ("field", R.NONZERO_WALL_VELOCITY)
        elif case == 5:
            p[:] = 0.0
            v = corrected(inp, p)
            want = ("pressure", R.QUALITY_GATE)
        elif case == 6:
            p[:] = 1.0
            v = corrected(inp, p)
            want = ("pressure", R.NONZERO_GAUGE)
        else:
            v[0][interior] = np.frombuffer(struct.pack("<I", int(u32(v[0])[interior]) ^ 1), np.float32)[0]
            want = ("momentum_witness", None)
        st, rep = ws.qualify(inp, candidate_of(inp, p, v))
        assert st == R.OK, case
        assert_native(ws, rep, d)
        assert rep["selection"] == "native_fallback", case
        assert rep["supplied_identity"]
        assert rep["rejection"] == want[0], case
        if want[0] == "momentum_witness":
            assert rep["witness"] == (0, interior)
        else:
            assert rep["rejection_status"] == want[1], case
        if case == 4:
            assert rep["wall"] == (0, 0)
        assert rep["pressure"].selected == method
        if want[0] == "pressure":
            assert rep["pressure"].rejection == want[1] and rep["pressure"].attempted is not None
            if case == 5:
                assert rep["pressure"].candidate.true_residual_l2 > 0.0
        else:
            assert rep["pressure"].attempted is None and rep["pressure"].rejection == R.OK
        st, rep = ws.qualify(inp, candidate_of(inp, d["pressure"], d["velocity"]))
        assert st == R.OK and rep["selection"] == "supplied" and rep["rejection"] is None
    ws.close()


def test_refusals_come_before_any_work():
    spec, vel = nonzero_fixture()
    inp = make_input(spec, vel)
    d = direct(inp, "jacobi-pcg-v1")
    ws = R.ProjectionWorkspace(spec, 1 << 20)

    def never():
        raise AssertionError("a refusal polled cancellation")

    # binding: a candidate for another input
    other = make_input(spec, vel, generation=1)
    st, rep = ws.qualify(inp, candidate_of(other, d["pressure"], d["velocity"]), never)
    assert st == R.REFUSED and rep["refusal"] == "input_binding"
    # representation: every field of the header
    for field in range(18):
        rp = R.projection_representation(spec)
        if field == 0:
            rp.version += 1
        elif field == 1:
            rp.operator_id = b"unsupported-operator"
        else:
            f = field - 2
            meta = rp.pressure if f // 4 == 0 else rp.velocity[f // 4 - 1]
            if f % 4 == 0:
                meta.shape_xyz[0] += 1
            elif f % 4 == 1:
                meta.layout = b"unsupported-layout"
            elif f % 4 == 2:
                meta.scalar = b"f16"
            else:
                meta.units = b"unsupported"
        st, rep = ws.qualify(inp, candidate_of(inp, d["pressure"], d["velocity"], representation=rp), never)
        assert st == R.REFUSED and rep["refusal"] == "representation", field
    for kw in (dict(algorithm=""), dict(version="bad\nversion"), dict(source_revision=""), dict(algorithm="a" * 257)):
        c = candidate_of(inp, d["pressure"], d["velocity"])
        for k, v in kw.items():
            setattr(c, k, v)
        st, rep = ws.qualify(inp, c, never)
        assert st == R.REFUSED and rep["refusal"] == "identity", kw
    for domain in (R.DOMAIN_FREE_SURFACE, R.DOMAIN_EMBEDDED_OBSTACLE, R.DOMAIN_MOVING_WALL):
        st, rep = ws.qualify(make_input(spec, vel, domain=domain), None, never)
        assert st == R.REFUSED and rep["refusal"] == "domain"
    zspec, zvel = zero_fixture()
    st, rep = ws.qualify(make_input(zspec, zvel), None, never)
    assert st == R.REFUSED and rep["refusal"] == "workspace_grid"
    st, rep = ws.qualify(inp, None)
    assert st == R.OK
    assert_native(ws, rep, d)
    ws.close()


@pytest.mark.parametrize("method", METHODS)
def test_actual_divergence_rejection_keeps_both_pressure_calls(method):
    spec, vel = zero_fixture(0.0)
    inp = make_input(spec, vel, actual_divergence_limit=0.0)
    p = np.zeros(spec.cell_len)
    p[1] = 1e-13
    v = corrected(inp, p)
    cand_div = O.actual_divergence_max(ogrid(spec), v)
    assert cand_div > 0.0
    native = direct(make_input(spec, vel), method)
    ws = R.ProjectionWorkspace(spec, 1 << 20, method)
    st, rep = ws.qualify(inp, candidate_of(inp, p, v))
    assert st == R.OK
    assert_native(ws, rep, native)
    assert rep["selection"] == "native_fallback" and rep["supplied_identity"]
    assert rep["rejection"] == "actual_divergence" and rep["rejected_divergence"] == cand_div
    assert rep["actual_divergence_max"] == 0.0
    assert rep["pressure"].selected == method and rep["pressure"].attempted is None
    disc = rep["discarded"]
    assert disc.selected == "supplied-projection-pressure-witness-v1" and disc.attempted is not None
    assert disc.rejection == R.OK and disc.accepted.iterations == 0
    assert 0.0 < disc.accepted.true_residual_l2 < 1e-12
    assert disc.accepted.predicted_divergence_max <= 1e-8
    ws.close()


def test_cancellation_returns_no_result_and_allows_retry():
    spec, vel = nonzero_fixture()
    inp = make_input(spec, vel)
    d = direct(inp, "jacobi-pcg-v1")
    ws = R.ProjectionWorkspace(spec, 1 << 20)
    for stop_at in range(1, 5):
        calls = [0]

        def cancelled():
            calls[0] += 1
            return calls[0] >= stop_at

        st, _ = ws.qualify(inp, candidate_of(inp, d["pressure"], d["velocity"]), cancelled)
        assert st == R.CANCELLED, stop_at
        with pytest.raises(R.RheonError):
            ws.result()
    st, rep = ws.qualify(inp, candidate_of(inp, d["pressure"], d["velocity"]))
    assert st == R.OK and rep["selection"] == "supplied"
    ws.close()
