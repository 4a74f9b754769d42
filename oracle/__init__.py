"""CPU oracle for Rheon's fixed-box MAC projection step.

TEST INFRASTRUCTURE ONLY.  This package is the checker for the HIP path in
``rheon_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product never calls
it and has no CPU fallback through it.

It wraps ``oracle/liboracle.so`` (built from ``rheon_oracle.c`` by
``oracle/Makefile``), a C restatement of the reference routines
(src/geometry.rs, src/sampling.rs, src/advection.rs, src/operator.rs,
src/pressure.rs; line citations beside each C function).  Arrays are numpy,
x-fastest, exactly the reference's flat layouts.

Parity status of each function (how it is pinned) is listed in the test
module tests/test_oracle.py; every function here has at least one check
against a reference known-answer test, a reference-published golden run, or
a closed form.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK = 0
CANCELLED = 6
INVALID_ARGUMENT = 7
INVALID_SETTINGS = 8
INCOMPATIBLE_RHS = 9
INVALID_INITIAL_PRESSURE = 24
BREAKDOWN = 10
ITERATION_LIMIT = 11
NONFINITE_INPUT = 12
NONZERO_WALL_VELOCITY = 13
ARITHMETIC_FAILURE = 14
INVALID_TIME_STEP = 15
INVALID_DENSITY = 16
NONFINITE_POSITION = 17
GEOMETRY = 20

JACOBI, SGS, RBSGS = 0, 1, 2
METHODS = {"jacobi-pcg-v1": JACOBI, "sgs-pcg-v1": SGS, "rbsgs-pcg": RBSGS}


class _Grid(C.Structure):
    _fields_ = [
        ("n", C.c_int64 * 3),
        ("fn", (C.c_int64 * 3) * 3),
        ("cell_len", C.c_int64),
        ("face_len", C.c_int64 * 3),
        ("h", C.c_double * 3),
        ("origin", C.c_double * 3),
        ("upper", C.c_double * 3),
        ("volume", C.c_double),
    ]


class _Settings(C.Structure):
    _fields_ = [
        ("relative_residual", C.c_double),
        ("absolute_residual", C.c_double),
        ("divergence_limit", C.c_double),
        ("max_iterations", C.c_int64),
    ]


class _Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("true_residual_l2", C.c_double),
        ("true_residual_max", C.c_double),
        ("predicted_divergence_max", C.c_double),
        ("error_iteration", C.c_int64),
        ("error_value", C.c_double),
    ]


def build() -> None:
    """Compile liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_G = C.POINTER(_Grid)


class _D3Meta(C.Structure):
    _fields_ = [("frame", C.c_uint64), ("time", C.c_double), ("has_report", C.c_int32), ("pad", C.c_int32),
                ("dt", C.c_double), ("carrier_id", C.c_uint64), ("carrier_version", C.c_uint64),
                ("liquid_id", C.c_uint64), ("liquid_version", C.c_uint64), ("pressure_iterations", C.c_uint64),
                ("diag", C.c_double * 11)]


def _declare(L):
    L.or_format_sci.argtypes = [C.c_double, C.c_int, C.c_char_p]
    L.or_format_sci.restype = C.c_int
    L.or_dense3d_frame.argtypes = [C.POINTER(C.c_int64), _P, _P, _P, _P, _P, _P, C.POINTER(_D3Meta), _P, C.c_int64,
                                   C.c_int]
    L.or_dense3d_frame.restype = C.c_int64
    L.or_grid_init.argtypes = [_G, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_buffer_plan.argtypes = [_G, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.or_buffer_plan.restype = C.c_int64
    for suf in ("f32", "f64"):
        getattr(L, f"or_sample_{suf}").argtypes = [_G, C.c_int, _P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        getattr(L, f"or_backtrace_{suf}").argtypes = [_G, _P, _P, _P, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)]
        getattr(L, f"or_advect_velocity_{suf}").argtypes = [_G, _P, _P, _P, C.c_double, _P, _P, _P, C.c_int]
        getattr(L, f"or_build_rhs_{suf}").argtypes = [_G, _P, _P, _P, C.c_double, _P]
        getattr(L, f"or_divergence_{suf}").argtypes = [_G, _P, _P, _P, _P]
        getattr(L, f"or_correct_{suf}").argtypes = [_G, C.c_double, _P, C.c_double, _P, _P, _P]
        getattr(L, f"or_actual_divergence_max_{suf}").argtypes = [_G, _P, _P, _P, C.POINTER(C.c_double)]
        getattr(L, f"or_velocity_rate_{suf}").argtypes = [_G, _P, _P, _P]
        getattr(L, f"or_velocity_rate_{suf}").restype = C.c_double
    L.or_advect_tracer_f32.argtypes = [_G, _P, _P, _P, _P, C.c_double, _P, C.c_int]
    L.or_apply_full.argtypes = [_G, C.c_double, _P, _P, C.c_int]
    L.or_diagonal.argtypes = [_G, C.c_double, C.POINTER(C.c_int64)]
    L.or_diagonal.restype = C.c_double
    L.or_compensated_sum.argtypes = [_P, C.c_int64]
    L.or_compensated_sum.restype = C.c_double
    L.or_dot.argtypes = [_P, _P, C.c_int64]
    L.or_dot.restype = C.c_double
    for name in ("or_precondition_sgs", "or_precondition_rbsgs"):
        getattr(L, name).argtypes = [_G, C.c_double, _P, _P, C.POINTER(C.c_double)]
    L.or_precondition_jacobi.argtypes = [_G, C.c_double, _P, _P, C.POINTER(C.c_double), C.c_int]
    L.or_solve_rhs.argtypes = [_G, C.c_double, _P, C.c_double, C.POINTER(_Settings), C.c_int, C.c_int, _P, C.POINTER(_Report)]
    L.or_solve_rhs_initial.argtypes = [_G, C.c_double, _P, _P, C.c_double, C.POINTER(_Settings), C.c_int, C.c_int, _P,
                                       C.POINTER(_Report)]
    L.or_apply_acceleration.argtypes = [_G, C.c_double, C.POINTER(C.c_double * 3), C.c_void_p, C.c_void_p, _P, _P, _P]
    L.or_obstacle_mask.argtypes = [_G, _P, _P, C.POINTER(C.c_int64)]
    L.or_obstacle_components.argtypes = [_G, _P, _P]
    L.or_obstacle_components.restype = C.c_int32
    L.or_apply_masked.argtypes = [_G, C.c_double, _P, _P, _P]
    L.or_precondition_masked.argtypes = [_G, C.c_double, _P, C.c_int64, C.c_int, _P, _P, C.POINTER(C.c_double)]
    L.or_solve_rhs_masked.argtypes = [_G, C.c_double, _P, C.c_int64, _P, C.c_double, C.POINTER(_Settings), C.c_int,
                                      C.c_int, _P, C.POINTER(_Report)]
    L.or_correct_f32_masked.argtypes = [_G, C.c_double, _P, _P, C.c_double, _P, _P, _P]
    L.or_actual_divergence_max_f32_masked.argtypes = [_G, _P, _P, _P, _P, C.POINTER(C.c_double)]


class OracleError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        super().__init__(f"oracle status {status} {what}")
        self.status = status


def _ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class Grid:
    """GridGeometry (src/geometry.rs:84-267)."""

    def __init__(self, counts, spacing, origin=(0.0, 0.0, 0.0)):
        self.s = _Grid()
        st = lib().or_grid_init(C.byref(self.s), (C.c_uint64 * 3)(*[int(c) for c in counts]), _d3(spacing), _d3(origin))
        if st != OK:
            raise OracleError(st, "GridGeometry::new")
        self.counts = tuple(int(c) for c in counts)
        self.spacing = tuple(float(h) for h in spacing)
        self.origin = tuple(float(o) for o in origin)

    @property
    def cell_len(self) -> int:
        return int(self.s.cell_len)

    def face_counts(self, axis: int):
        return tuple(int(self.s.fn[axis][d]) for d in range(3))

    def face_len(self, axis: int) -> int:
        return int(self.s.face_len[axis])

    @property
    def upper(self):
        return tuple(self.s.upper)

    @property
    def volume(self) -> float:
        return float(self.s.volume)

    def buffer_plan(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        total = lib().or_buffer_plan(C.byref(self.s), C.byref(a), C.byref(b), C.byref(c))
        return {"velocity_pair_bytes": a.value, "tracer_pair_bytes": b.value,
                "pressure_workspace_bytes": c.value, "total_bytes": total}


def _suffix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"unsupported storage {dt}")


def sample(grid: Grid, component: int, values: np.ndarray, pos) -> float:
    out = C.c_double()
    st = getattr(lib(), f"or_sample_{_suffix(values.dtype)}")(C.byref(grid.s), component, _ptr(values), _d3(pos), C.byref(out))
    if st != OK:
        raise OracleError(st, "sample")
    return out.value


def backtrace(grid: Grid, vel, pos, dt: float):
    out = (C.c_double * 3)()
    u, v, w = vel
    st = getattr(lib(), f"or_backtrace_{_suffix(u.dtype)}")(C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w), _d3(pos), dt, out)
    if st != OK:
        raise OracleError(st, "backtrace")
    return [out[0], out[1], out[2]]


def advect_velocity(grid: Grid, vel, dt: float, threads: int = 1):
    u, v, w = vel
    outs = [np.empty_like(x) for x in vel]
    st = getattr(lib(), f"or_advect_velocity_{_suffix(u.dtype)}")(
        C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w), dt, _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]), threads)
    if st != OK:
        raise OracleError(st, "advect_velocity")
    return outs


def advect_tracer(grid: Grid, old: np.ndarray, vel, dt: float, threads: int = 1):
    u, v, w = vel
    out = np.empty_like(old)
    st = lib().or_advect_tracer_f32(C.byref(grid.s), _ptr(old), _ptr(u), _ptr(v), _ptr(w), dt, _ptr(out), threads)
    if st != OK:
        raise OracleError(st, "advect_tracer")
    return out


def build_rhs(grid: Grid, vel, dt: float) -> np.ndarray:
    u, v, w = vel
    out = np.empty(grid.cell_len, np.float64)
    st = getattr(lib(), f"or_build_rhs_{_suffix(u.dtype)}")(C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w), dt, _ptr(out))
    if st != OK:
        raise OracleError(st, "build_rhs")
    return out


def apply_full(grid: Grid, density: float, x: np.ndarray, threads: int = 1) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    st = lib().or_apply_full(C.byref(grid.s), density, _ptr(x), _ptr(y), threads)
    if st != OK:
        raise OracleError(st, "apply_full")
    return y


def diagonal(grid: Grid, density: float, p) -> float:
    return lib().or_diagonal(C.byref(grid.s), density, (C.c_int64 * 3)(*p))


def divergence(grid: Grid, vel) -> np.ndarray:
    u, v, w = vel
    out = np.empty(grid.cell_len, np.float64)
    st = getattr(lib(), f"or_divergence_{_suffix(u.dtype)}")(C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w), _ptr(out))
    if st != OK:
        raise OracleError(st, "divergence")
    return out


def actual_divergence_max(grid: Grid, vel) -> float:
    u, v, w = vel
    out = C.c_double()
    st = getattr(lib(), f"or_actual_divergence_max_{_suffix(u.dtype)}")(C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w), C.byref(out))
    if st != OK:
        raise OracleError(st, "actual_divergence_max")
    return out.value


def correct(grid: Grid, density: float, p: np.ndarray, dt: float, vel):
    """Return corrected copies of vel (correct_candidate_in_place / _f64)."""
    outs = [np.array(x, copy=True) for x in vel]
    st = getattr(lib(), f"or_correct_{_suffix(outs[0].dtype)}")(
        C.byref(grid.s), density, _ptr(p), dt, _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]))
    if st != OK:
        raise OracleError(st, "correct")
    return outs


def velocity_rate(grid: Grid, vel) -> float:
    u, v, w = vel
    return getattr(lib(), f"or_velocity_rate_{_suffix(u.dtype)}")(C.byref(grid.s), _ptr(u), _ptr(v), _ptr(w))


def compensated_sum(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_compensated_sum(_ptr(x), x.size)


def dot(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().or_dot(_ptr(a), _ptr(b), a.size)


def precondition(grid: Grid, density: float, r: np.ndarray, method: str = "jacobi-pcg-v1", threads: int = 1):
    """Return (r with gauge row zeroed, z, rz)."""
    r = np.array(r, np.float64, copy=True)
    z = np.zeros_like(r)
    rz = C.c_double()
    m = METHODS[method]
    if m == JACOBI:
        st = lib().or_precondition_jacobi(C.byref(grid.s), density, _ptr(r), _ptr(z), C.byref(rz), threads)
    elif m == SGS:
        st = lib().or_precondition_sgs(C.byref(grid.s), density, _ptr(r), _ptr(z), C.byref(rz))
    else:
        st = lib().or_precondition_rbsgs(C.byref(grid.s), density, _ptr(r), _ptr(z), C.byref(rz))
    if st != OK:
        raise OracleError(st, "precondition")
    return r, z, rz.value


@dataclass
class Settings:
    """PressureSettings (src/pressure.rs:33-39)."""
    relative_residual: float = 1e-9
    absolute_residual: float = 1e-12
    divergence_limit: float = 1e-7
    max_iterations: int = 2000


@dataclass
class SolveResult:
    status: int
    pressure: np.ndarray
    iterations: int
    true_residual_l2: float
    true_residual_max: float
    predicted_divergence_max: float
    error_iteration: int
    error_value: float

    @property
    def ok(self) -> bool:
        return self.status == OK


def solve_rhs(grid: Grid, density: float, rhs: np.ndarray, dt: float, settings: Settings,
              method: str = "jacobi-pcg-v1", threads: int = 1, initial: np.ndarray | None = None) -> SolveResult:
    """PressureWorkspace::solve_rhs (src/pressure.rs:365-393 -> :577-740), or
    solve_rhs_with_initial_pressure (:413-429) when ``initial`` is given.

    Returns the status instead of raising; ``pressure`` holds the last iterate
    (also on IterationLimit, as the reference's workspace does)."""
    rhs = np.ascontiguousarray(rhs, np.float64)
    p = np.zeros(grid.cell_len, np.float64)
    s = _Settings(settings.relative_residual, settings.absolute_residual, settings.divergence_limit,
                  int(settings.max_iterations))
    rep = _Report()
    if initial is None:
        st = lib().or_solve_rhs(C.byref(grid.s), density, _ptr(rhs), dt, C.byref(s), METHODS[method], threads,
                                _ptr(p), C.byref(rep))
    else:
        initial = np.ascontiguousarray(initial, np.float64)
        assert initial.size == grid.cell_len
        st = lib().or_solve_rhs_initial(C.byref(grid.s), density, _ptr(rhs), _ptr(initial), dt, C.byref(s),
                                        METHODS[method], threads, _ptr(p), C.byref(rep))
    return SolveResult(st, p, rep.iterations, rep.true_residual_l2, rep.true_residual_max,
                       rep.predicted_divergence_max, rep.error_iteration, rep.error_value)


def apply_acceleration(grid: Grid, vel, dt: float, value, lower=None, upper=None):
    """Copies of f32 faces with one body force applied (src/forces.rs:97-205)."""
    outs = [np.array(x, np.float32, copy=True) for x in vel]
    val = (C.c_double * 3)(*value)
    lo = (C.c_double * 3)(*lower) if lower is not None else None
    hi = (C.c_double * 3)(*upper) if upper is not None else None
    st = lib().or_apply_acceleration(C.byref(grid.s), dt, val, lo, hi, _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]))
    if st != OK:
        raise OracleError(st, "apply_acceleration")
    return outs


def solve_velocity(grid: Grid, density: float, vel, dt: float, settings: Settings,
                   method: str = "jacobi-pcg-v1", threads: int = 1) -> SolveResult:
    """PressureWorkspace::solve_velocity (src/pressure.rs:344-361)."""
    return solve_rhs(grid, density, build_rhs(grid, vel, dt), dt, settings, method, threads)


# --------------------------------------------------------------------------
# Solid obstacles (src/obstacle_pressure.rs:224-309 with all-or-nothing face
# apertures): a per-cell open-face mask and the lowest-index fluid cell as the
# gauge.
# --------------------------------------------------------------------------

@dataclass
class Obstacle:
    mask: np.ndarray  # uint8 per cell: bits 0..5 open faces x-,x+,y-,y+,z-,z+; bit 6 fluid; bit 7 gauge
    gauge: int        # the first gauge (lowest-index fluid cell)

    @property
    def gauges(self) -> np.ndarray:
        """One gauge per connected fluid region, in label order."""
        return np.flatnonzero(self.mask & 128)


def obstacle_mask(grid: Grid, solid: np.ndarray) -> Obstacle:
    solid = np.ascontiguousarray(solid, np.uint8).reshape(-1)
    assert solid.size == grid.cell_len
    mask = np.zeros(grid.cell_len, np.uint8)
    gauge = C.c_int64()
    st = lib().or_obstacle_mask(C.byref(grid.s), _ptr(solid), _ptr(mask), C.byref(gauge))
    if st != OK:
        raise OracleError(st, "obstacle_mask")
    return Obstacle(mask, gauge.value)


def obstacle_components(grid: Grid, mask: np.ndarray):
    """(labels, count): connected fluid regions of a mask (-1 = solid)."""
    mask = np.ascontiguousarray(mask, np.uint8)
    labels = np.empty(grid.cell_len, np.int32)
    count = lib().or_obstacle_components(C.byref(grid.s), _ptr(mask), _ptr(labels))
    if count < 0:
        raise OracleError(INVALID_ARGUMENT, "obstacle_components")
    return labels, int(count)


def apply_masked(grid: Grid, density: float, ob: Obstacle, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    st = lib().or_apply_masked(C.byref(grid.s), density, _ptr(ob.mask), _ptr(x), _ptr(y))
    if st != OK:
        raise OracleError(st, "apply_masked")
    return y


def precondition_masked(grid: Grid, density: float, ob: Obstacle, r: np.ndarray, method: str = "jacobi-pcg-v1"):
    r = np.array(r, np.float64, copy=True)
    z = np.zeros_like(r)
    rz = C.c_double()
    st = lib().or_precondition_masked(C.byref(grid.s), density, _ptr(ob.mask), ob.gauge, METHODS[method],
                                      _ptr(r), _ptr(z), C.byref(rz))
    if st != OK:
        raise OracleError(st, "precondition_masked")
    return r, z, rz.value


def solve_rhs_masked(grid: Grid, density: float, ob: Obstacle, rhs: np.ndarray, dt: float, settings: Settings,
                     method: str = "jacobi-pcg-v1", threads: int = 1) -> SolveResult:
    rhs = np.ascontiguousarray(rhs, np.float64)
    p = np.zeros(grid.cell_len, np.float64)
    s = _Settings(settings.relative_residual, settings.absolute_residual, settings.divergence_limit,
                  int(settings.max_iterations))
    rep = _Report()
    st = lib().or_solve_rhs_masked(C.byref(grid.s), density, _ptr(ob.mask), ob.gauge, _ptr(rhs), dt, C.byref(s),
                                   METHODS[method], threads, _ptr(p), C.byref(rep))
    return SolveResult(st, p, rep.iterations, rep.true_residual_l2, rep.true_residual_max,
                       rep.predicted_divergence_max, rep.error_iteration, rep.error_value)


def correct_masked(grid: Grid, density: float, ob: Obstacle, p: np.ndarray, dt: float, vel):
    outs = [np.array(x, np.float32, copy=True) for x in vel]
    st = lib().or_correct_f32_masked(C.byref(grid.s), density, _ptr(ob.mask), _ptr(p), dt,
                                     _ptr(outs[0]), _ptr(outs[1]), _ptr(outs[2]))
    if st != OK:
        raise OracleError(st, "correct_masked")
    return outs


def actual_divergence_max_masked(grid: Grid, ob: Obstacle, vel) -> float:
    u, v, w = vel
    out = C.c_double()
    st = lib().or_actual_divergence_max_f32_masked(C.byref(grid.s), _ptr(ob.mask), _ptr(u), _ptr(v), _ptr(w),
                                                   C.byref(out))
    if st != OK:
        raise OracleError(st, "actual_divergence_max_masked")
    return out.value


def zero_blocked_faces(grid: Grid, ob: Obstacle, vel):
    """Faces with a solid cell on either side are closed (velocity 0); box
    walls are not touched (they carry the wall condition)."""
    nx, ny, nz = (int(v) for v in grid.s.n)
    fluid = (ob.mask & 64).astype(bool).reshape(nz, ny, nx)
    outs = []
    for axis, comp in enumerate(vel):
        a = np.array(comp, copy=True).reshape(nz + (axis == 2), ny + (axis == 1), nx + (axis == 0))
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        ax = 2 - axis
        sl_lo[ax] = slice(0, -1)
        sl_hi[ax] = slice(1, None)
        inner = [slice(None)] * 3
        inner[ax] = slice(1, -1)
        left = fluid[tuple(sl_lo)]
        right = fluid[tuple(sl_hi)]
        open_ = left & right
        a[tuple(inner)] = np.where(open_, a[tuple(inner)], 0)
        outs.append(a.reshape(-1))
    return outs



# ------------------------------------------------------------ dense3d frames
class Dense3dInvalidState(Exception):
    """Dense3dError::InvalidState (src/dense3d_sequence.rs:21-27)."""


def format_sci(v: float, digits: int) -> str:
    """Rust's `{v:.(digits-1)e}` (or_format_sci: glibc's exact conversion)."""
    buf = C.create_string_buffer(64)
    n = lib().or_format_sci(float(v), int(digits), buf)
    return buf.raw[:n].decode()


def dense3d_frame(counts, velocity, tracer, fraction, pressure, *, frame, carrier_time=0.0, liquid_time=0.0,
                  carrier_generation=0, carrier_stamp=(0, 0), liquid_stamp=(0, 0), transport_only=True,
                  report=None, threads: int = 1) -> bytes:
    """write_dense3d_frame (src/dense3d_sequence.rs:415-564): the preflight
    checks in the reference's order (:422-495), then the line (C
    serialization).  ``report``: an object with the rh_dense3d_report field
    names (carrier_time, carrier_dt, ...), or None for the constructor frame."""
    n = [int(c) for c in counts]
    cells = n[0] * n[1] * n[2]

    def invalid(ok, msg):
        if not ok:
            raise Dense3dInvalidState(msg)

    invalid(transport_only, "transport-only view required")
    vel = []
    for axis in range(3):
        a = np.ascontiguousarray(velocity[axis], np.float32)
        flen = (n[0] + (axis == 0)) * (n[1] + (axis == 1)) * (n[2] + (axis == 2))
        invalid(a.size == flen, "face shape")
        invalid(bool(np.all(np.isfinite(a))), "nonfinite velocity")
        vel.append(a)
    tr = np.ascontiguousarray(tracer, np.float32)
    fr = np.ascontiguousarray(fraction, np.float64)
    pr = np.ascontiguousarray(pressure, np.float64)
    invalid(tr.size == cells and fr.size == cells and pr.size == cells, "cell shape")
    invalid(bool(np.all(np.isfinite(tr)) and np.all(np.isfinite(pr)) and np.all(np.isfinite(fr))
                 and np.all((fr >= 0.0) & (fr <= 1.0))), "nonfinite or invalid cell field")
    invalid(np.isfinite(carrier_time) and carrier_time == liquid_time and carrier_stamp[1] == carrier_generation,
            "state clock/stamp")
    m = _D3Meta()
    m.frame = int(frame)
    m.time = float(carrier_time)
    if report is not None:
        r = report
        invalid(r.carrier_time == carrier_time and r.liquid_time == r.carrier_time and r.carrier_dt == r.liquid_dt
                and r.carrier_dt > 0.0 and r.carrier_generation == carrier_generation
                and tuple(r.liquid_stamp) == tuple(liquid_stamp) and tuple(r.liquid_flow) == tuple(carrier_stamp),
                "report continuity")
        diag = [r.pressure_true_residual_max, r.carrier_actual_divergence_max, r.liquid_actual_divergence_max,
                r.liquid_volume_before, r.liquid_volume_after, r.liquid_mass_after, r.inward_boundary_volume,
                r.outward_boundary_volume, r.source_volume, r.volume_balance_error, r.volume_rounding_budget]
        invalid(all(np.isfinite(x) for x in [r.carrier_dt] + diag), "nonfinite report")
        m.has_report = 1
        m.dt = float(r.carrier_dt)
        m.pressure_iterations = int(r.pressure_iterations)
        for k, x in enumerate(diag):
            m.diag[k] = float(x)
    else:
        invalid(frame == 0 and carrier_time == 0.0 and carrier_generation == 0, "initial frame")
    m.carrier_id, m.carrier_version = (int(x) for x in carrier_stamp)
    m.liquid_id, m.liquid_version = (int(x) for x in liquid_stamp)
    nn = (C.c_int64 * 3)(*n)
    args = [nn] + [_ptr(a) for a in vel] + [_ptr(tr), _ptr(fr), _ptr(pr), C.byref(m)]
    need = -lib().or_dense3d_frame(*args, None, 0, int(threads))
    out = C.create_string_buffer(need)
    got = lib().or_dense3d_frame(*args, out, need, int(threads))
    assert got == need
    return out.raw[:got]
