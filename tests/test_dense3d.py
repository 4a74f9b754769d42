"""Dense3d frames: write_dense3d_frame (src/dense3d_sequence.rs:415-564).

The number format is Rust's exact `{v:.9e}` (f32 fields) and `{v:.17e}` (f64
fields, diagnostics): the exact binary value rounded to 10 or 18 significant
digits, ties to even, exponent without '+' or padding, '-' for -0.0.  Three
implementations are held to each other and to the reference's output:

* the oracle (oracle/dense3d_oracle.c): glibc's exact "%.*e" with the
  exponent rewritten;
* Python's own exact conversion ("%.*e", correctly rounded, ties to even),
  rewritten the same way here;
* the product's formatter (rheon_amd/csrc/dense3d.hpp), on the host through
  rh_format_sci and on the device inside rh_dense3d_frame;

and the oracle and the host formatter reproduce every number the reference
printed with these specifiers in its published tables
(tests/golden/rust_sci.json, tests/golden/make_rust_sci.py).  The frame
layout follows the reference's format strings; the round trip of every
field bit and the decimal u64 stamps mirror
tests/dense3d_sequence_contract.rs:118-214.
"""
import hashlib
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as O
import rheon_amd as R
from rheon_amd.scenes import GridSpec

HERE = os.path.dirname(os.path.abspath(__file__))
U64_MAX = 2**64 - 1


def py_rust(x: float, digits: int) -> str:
    """Python's exact %e conversion in Rust's exponent form."""
    s = "%.*e" % (digits - 1, x)
    mant, ex = s.split("e")
    return f"{mant}e{int(ex)}"


def f32(x: float) -> float:
    return struct.unpack("<f", struct.pack("<f", x))[0]


# ------------------------------------------------------------ number format
def test_reference_printed_numbers():
    d = json.load(open(os.path.join(HERE, "golden", "rust_sci.json")))
    assert len(d["sci18"]) > 1000 and d["sci10"]
    for digits, key in ((18, "sci18"), (10, "sci10")):
        for s in d[key]:
            v = float(s)
            assert O.format_sci(v, digits) == s, s
            assert R.format_sci(v, digits) == s, s


def special_values():
    vals = [0.0, -0.0, 1.0, -1.0, 0.1, 0.5, 1.5, 2.5, 1e-10, 1e-11, 1e-18, 1e-19, 1e-300, 5e-324,
            2.2250738585072014e-308, 1.7976931348623157e308, 1e22, 1e23, 9.999999999999999e22,
            123456789012345678.0, 1000000000000000.125, 99999999999999999.0, 9.9999999995, 0.99999999995]
    for x in (3.4028234663852886e38, 1.401298464324817e-45, 1.1754943508222875e-38, 1.0009765625,
              1.0029296875, 16777215.0, 1e9, 9999999999.0, 0.00048828125):
        vals += [x, -x]
    return vals


def test_format_matches_exact_conversion():
    rng = np.random.default_rng(11)
    vals = special_values()
    bits = rng.integers(0, 2**63, 20000, dtype=np.int64).view(np.float64)
    vals += [float(v) for v in bits if math.isfinite(v)]
    vals += [float(v) for v in rng.uniform(-1, 1, 20000) * 10.0 ** rng.integers(-40, 40, 20000)]
    fb = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    f32s = [float(v) for v in fb if math.isfinite(v)] + [k / 1024 for k in range(1025, 9000, 2)]
    for x in vals:
        want = py_rust(x, 18)
        assert O.format_sci(x, 18) == want, x
        assert R.format_sci(x, 18) == want, x
    for x in f32s + [f32(v) for v in vals if abs(v) < 3.4e38]:
        want = py_rust(x, 10)
        assert O.format_sci(x, 10) == want, x
        assert R.format_sci(x, 10) == want, x


def test_ties_round_to_even():
    # 1025/1024 = 1.0009765625 (11 digits, tie at 10): the even neighbour
    assert R.format_sci(1.0009765625, 10) == "1.000976562e0"
    assert R.format_sci(1.0029296875, 10) == "1.002929688e0"
    assert R.format_sci(-1.0009765625, 10) == "-1.000976562e0"
    # 10^15 + 1/8: 19 significant digits, tie at 18
    assert R.format_sci(1000000000000000.125, 18) == "1.00000000000000012e15"
    assert R.format_sci(1000000000000000.375, 18) == "1.00000000000000038e15"
    # carry into a new decade
    assert R.format_sci(9.9999999995, 10) == py_rust(9.9999999995, 10)
    assert R.format_sci(-0.0, 18) == "-0.00000000000000000e0"


def test_round_trip_bits():
    rng = np.random.default_rng(3)
    d = rng.integers(0, 2**63, 5000, dtype=np.int64).view(np.float64)
    for x in d[np.isfinite(d)]:
        assert float(R.format_sci(float(x), 18)) == float(x)
    f = rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    for x in f[np.isfinite(f)]:
        assert np.float32(float(R.format_sci(float(x), 10))) == x


# ------------------------------------------------------------- frame layout
def face_len(n, a):
    return (n[0] + (a == 0)) * (n[1] + (a == 1)) * (n[2] + (a == 2))


def make_state(n, seed, specials=False):
    rng = np.random.default_rng(seed)
    vel = [rng.standard_normal(face_len(n, a)).astype(np.float32) for a in range(3)]
    cells = n[0] * n[1] * n[2]
    tracer = rng.random(cells).astype(np.float32)
    fraction = rng.random(cells)
    pressure = rng.standard_normal(cells) * 10.0 ** rng.integers(-12, 12, cells)
    if specials:
        sv = np.array([v for v in special_values() if abs(v) < 3.4e38], np.float64)
        for a in range(3):
            k = min(sv.size, vel[a].size)
            vel[a][:k] = sv[:k].astype(np.float32)
        k = min(sv.size, cells)
        pressure[:k] = sv[:k]
        fraction[: min(4, cells)] = [0.0, -0.0, 1.0, 5e-324][: min(4, cells)]
        tracer[: min(3, cells)] = [-0.0, 1e-45, 3.4e38][: min(3, cells)]
    return vel, tracer, fraction, pressure


def make_report(time, generation, liquid_stamp, carrier_stamp):
    return R.Dense3dReport(
        carrier_time=time, carrier_dt=0.0625, carrier_generation=generation, pressure_iterations=17,
        pressure_true_residual_max=1.25e-13, carrier_actual_divergence_max=3.0e-9, liquid_time=time,
        liquid_dt=0.0625, liquid_stamp=liquid_stamp, liquid_flow=carrier_stamp,
        liquid_actual_divergence_max=2.5e-10, liquid_volume_before=0.125, liquid_volume_after=0.125000000001,
        liquid_mass_after=100.0, inward_boundary_volume=0.0, outward_boundary_volume=1e-300, source_volume=-0.0,
        volume_balance_error=-2.0e-17, volume_rounding_budget=64 * 2.220446049250313e-16 * 0.25)


def oracle_frame(n, state, frame, report=None, *, carrier_stamp=(U64_MAX - 1, 0), liquid_stamp=(U64_MAX, 0),
                 time=0.0, generation=0, threads=1):
    vel, tracer, fraction, pressure = state
    return O.dense3d_frame(n, vel, tracer, fraction, pressure, frame=frame, carrier_time=time, liquid_time=time,
                           carrier_generation=generation, carrier_stamp=carrier_stamp, liquid_stamp=liquid_stamp,
                           report=report, threads=threads)


def test_oracle_frame_layout_and_round_trip():
    n = (16, 8, 4)
    state = make_state(n, 1, specials=True)
    line = oracle_frame(n, state, 0)
    assert line.endswith(b"}\n") and line.count(b"\n") == 1 and b" " not in line
    doc = json.loads(line)
    assert list(doc) == ["frame", "time_s", "dt_s", "carrier_stamp", "liquid_stamp", "fields", "diagnostics"]
    assert doc["carrier_stamp"] == {"id": str(U64_MAX - 1), "version": "0"}
    assert doc["liquid_stamp"] == {"id": str(U64_MAX), "version": "0"}
    assert doc["diagnostics"] is None and b'"diagnostics":null' in line
    assert list(doc["fields"]) == ["velocity_x", "velocity_y", "velocity_z", "tracer", "fraction", "pressure"]
    vel, tracer, fraction, pressure = state
    text = line.decode()

    def field(name):
        return text.split(f'"{name}":[', 1)[1].split("]", 1)[0].split(",")

    for a, name in enumerate(["velocity_x", "velocity_y", "velocity_z"]):
        got = np.array([np.float32(float(s)) for s in field(name)], np.float32)
        assert got.view(np.uint32).tolist() == vel[a].view(np.uint32).tolist()
    got = np.array([np.float32(float(s)) for s in field("tracer")], np.float32)
    assert got.view(np.uint32).tolist() == tracer.view(np.uint32).tolist()
    for name, arr in (("fraction", fraction), ("pressure", pressure)):
        got = np.array([float(s) for s in field(name)])
        assert got.view(np.uint64).tolist() == arr.view(np.uint64).tolist()
    # a later frame carries the report's diagnostics in the reference's order
    rep = make_report(0.0625, 1, (U64_MAX, 1), (U64_MAX - 1, 1))
    line = oracle_frame(n, state, 1, rep, carrier_stamp=(U64_MAX - 1, 1), liquid_stamp=(U64_MAX, 1), time=0.0625,
                        generation=1)
    doc = json.loads(line)
    assert list(doc["diagnostics"]) == [
        "pressure_iterations", "pressure_residual_m3_s2", "carrier_divergence_s_inv", "liquid_divergence_s_inv",
        "volume_before_m3", "volume_after_m3", "mass_after_kg", "inward_m3", "outward_m3", "source_m3",
        "balance_m3", "rounding_budget_m3"]
    assert b'"dt_s":6.25000000000000000e-2' in line and b'"time_s":6.25000000000000000e-2' in line
    assert b'"source_m3":-0.00000000000000000e0' in line


def test_oracle_threads_give_the_same_bytes():
    n = (40, 40, 48)
    state = make_state(n, 2)
    assert oracle_frame(n, state, 0, threads=1) == oracle_frame(n, state, 0, threads=4)


def preflight_cases(n):
    """(description, kwargs changes, expected reference message)."""
    return [
        ("nan velocity", {"vel_nan": 1}, "nonfinite velocity"),
        ("inf tracer", {"tracer": math.inf}, "nonfinite or invalid cell field"),
        ("fraction above 1", {"fraction": 1.5}, "nonfinite or invalid cell field"),
        ("nan pressure", {"pressure": math.nan}, "nonfinite or invalid cell field"),
        ("clock", {"liquid_time": 1.0}, "state clock/stamp"),
        ("stamp", {"stamp_version": 1}, "state clock/stamp"),
        ("initial frame", {"frame": 1}, "initial frame"),
        ("report continuity", {"report_dt": 0.03125}, "report continuity"),
        ("nonfinite report", {"report_nan": True}, "nonfinite report"),
        ("transport only", {"transport_only": False}, "transport-only view required"),
    ]


def build_case(n, change):
    vel, tracer, fraction, pressure = make_state(n, 5)
    vel = [v.copy() for v in vel]
    if "vel_nan" in change:
        vel[change["vel_nan"]][3] = np.nan
    if "tracer" in change:
        tracer = tracer.copy()
        tracer[2] = change["tracer"]
    if "fraction" in change:
        fraction = fraction.copy()
        fraction[1] = change["fraction"]
    if "pressure" in change:
        pressure = pressure.copy()
        pressure[0] = change["pressure"]
    kw = dict(frame=change.get("frame", 0), carrier_time=0.0, liquid_time=change.get("liquid_time", 0.0),
              carrier_generation=0, carrier_stamp=(7, change.get("stamp_version", 0)),
              liquid_stamp=(9, 0), transport_only=change.get("transport_only", True), report=None)
    if "report_dt" in change or "report_nan" in change:
        kw.update(frame=3, carrier_time=0.5, liquid_time=0.5, carrier_generation=3, carrier_stamp=(7, 3))
        rep = make_report(0.5, 3, (9, 0), (7, 3))
        if "report_dt" in change:
            rep.liquid_dt = change["report_dt"]
        if "report_nan" in change:
            rep.liquid_mass_after = math.nan
        kw["report"] = rep
    return (vel, tracer, fraction, pressure), kw


@pytest.mark.parametrize("case", range(10))
def test_oracle_preflight(case):
    n = (5, 4, 3)
    desc, change, msg = preflight_cases(n)[case]
    (vel, tracer, fraction, pressure), kw = build_case(n, change)
    with pytest.raises(O.Dense3dInvalidState, match=msg):
        O.dense3d_frame(n, vel, tracer, fraction, pressure, **kw)


# ----------------------------------------------------------------- device
def device_frame(n, state, frame, report=None, *, carrier_stamp=(U64_MAX - 1, 0), liquid_stamp=(U64_MAX, 0),
                 time=0.0, generation=0):
    vel, tracer, fraction, pressure = state
    spec = GridSpec(tuple(n), tuple(1.0 / c for c in n))
    st = R.Dense3dState(velocity=tuple(vel), tracer=tracer, fraction=fraction, pressure=pressure, carrier_time=time,
                        liquid_time=time, carrier_generation=generation, carrier_stamp=carrier_stamp,
                        liquid_stamp=liquid_stamp)
    return R.write_dense3d_frame(spec, st, frame, report)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(1, 1, 1), (3, 2, 5), (16, 8, 4), (40, 30, 20), (33, 65, 17)])
def test_device_frame_equals_oracle(n):
    state = make_state(n, sum(n), specials=True)
    assert device_frame(n, state, 0) == oracle_frame(n, state, 0)
    rep = make_report(0.75, 12, (U64_MAX, 5), (U64_MAX - 1, 12))
    kw = dict(carrier_stamp=(U64_MAX - 1, 12), liquid_stamp=(U64_MAX, 5), time=0.75, generation=12)
    assert device_frame(n, state, 12, rep, **kw) == oracle_frame(n, state, 12, rep, **kw)


@pytest.mark.gpu
def test_device_frame_slow_path_magnitudes():
    """Tiny and huge magnitudes (the multi-word branch of the formatter) and
    rounding ties in every field."""
    n = (24, 20, 18)
    vel, tracer, fraction, pressure = make_state(n, 9)
    rng = np.random.default_rng(9)
    cells = n[0] * n[1] * n[2]
    pressure = rng.standard_normal(cells) * 10.0 ** rng.integers(-320, 308, cells)
    pressure[::7] = 1000000000000000.125
    for a in range(3):
        vel[a] = (rng.standard_normal(vel[a].size) * 10.0 ** rng.integers(-45, 38, vel[a].size)).astype(np.float32)
        vel[a][::5] = np.float32(1.0009765625)
    fraction = rng.random(cells) * 10.0 ** -rng.integers(0, 320, cells)
    state = (vel, tracer, fraction, pressure)
    assert device_frame(n, state, 0) == oracle_frame(n, state, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(10))
def test_device_preflight(case):
    n = (5, 4, 3)
    desc, change, msg = preflight_cases(n)[case]
    (vel, tracer, fraction, pressure), kw = build_case(n, change)
    spec = GridSpec(n, tuple(1.0 / c for c in n))
    st = R.Dense3dState(velocity=tuple(vel), tracer=tracer, fraction=fraction, pressure=pressure,
                        carrier_time=kw["carrier_time"], liquid_time=kw["liquid_time"],
                        carrier_generation=kw["carrier_generation"], carrier_stamp=kw["carrier_stamp"],
                        liquid_stamp=kw["liquid_stamp"], transport_only=kw["transport_only"])
    with pytest.raises(R.RheonError) as e:
        R.write_dense3d_frame(spec, st, kw["frame"], kw["report"])
    assert e.value.status == R.INVALID_STATE and msg in str(e.value)


@pytest.mark.gpu
def test_device_buffer_limit_reports_the_length():
    n = (6, 5, 4)
    state = make_state(n, 4)
    want = oracle_frame(n, state, 0)
    vel, tracer, fraction, pressure = state
    spec = GridSpec(n, tuple(1.0 / c for c in n))
    st = R.Dense3dState(velocity=tuple(vel), tracer=tracer, fraction=fraction, pressure=pressure,
                        carrier_stamp=(U64_MAX - 1, 0), liquid_stamp=(U64_MAX, 0))
    small = R.DeviceArray(len(want) - 1, np.uint8)
    with pytest.raises(R.RheonError) as e:
        R.dense3d_frame_device(spec, st, 0, None, small)
    assert e.value.status == R.BUFFER_LIMIT and e.value.needed == len(want)
    out = R.DeviceArray(len(want), np.uint8)
    assert R.dense3d_frame_device(spec, st, 0, None, out) == len(want)
    assert out.download().tobytes() == want


@pytest.mark.gpu
def test_device_frame_256_equals_oracle():
    """A 256^3 frame (the bench's size; random fields): byte-identical to
    the oracle's line (compared by SHA-256)."""
    n = (256, 256, 256)
    rng = np.random.default_rng(256)
    cells = 256 ** 3
    vel = [rng.standard_normal(face_len(n, a), dtype=np.float32) for a in range(3)]
    state = (vel, rng.random(cells, dtype=np.float32), rng.random(cells), rng.standard_normal(cells))
    dev = device_frame(n, state, 0)
    ref = oracle_frame(n, state, 0, threads=min(16, os.cpu_count() or 1))
    assert len(dev) == len(ref)
    assert hashlib.sha256(dev).digest() == hashlib.sha256(ref).digest()


@pytest.mark.parametrize("bad", ["face", "cell"])
def test_shape_refused_before_the_device(bad):
    """Arrays shorter than the grid's are refused on the host ("face shape" /
    "cell shape", src/dense3d_sequence.rs:431-442): the device kernels read
    every array to the grid's length."""
    n = (6, 5, 4)
    vel, tracer, fraction, pressure = make_state(n, 8)
    if bad == "face":
        vel[1] = vel[1][:-1]
    else:
        pressure = pressure[:-1]
    spec = GridSpec(n, tuple(1.0 / c for c in n))
    st = R.Dense3dState(velocity=tuple(vel), tracer=tracer, fraction=fraction, pressure=pressure)
    with pytest.raises(R.Dense3dShapeError, match=f"{bad} shape"):
        R.write_dense3d_frame(spec, st, 0)
