"""rheon_amd: Rheon's fixed-box MAC projection step on AMD Instinct MI355X.

The compute path is hand-written HIP for gfx950 in ``rheon_amd/csrc``,
exposed through the C ABI of ``include/rheon_gpu.h`` and bound here with
ctypes.  This module mirrors the reference's host-side interface for the
path (src/solver_provider.rs PressureProvider / ProviderWorkspace,
src/pressure.rs PressureWorkspace / PressureSettings / PressureReport,
src/simulation.rs step sequence) so callers and tests read like the
reference's own.  There is no CPU fallback: every compute call goes to the
device library and raises if it is missing.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .scenes import GridSpec

# rh_status values (include/rheon_gpu.h)
OK = 0
UNSUPPORTED = 1
EXECUTION = 2
NONFINITE_OUTPUT = 3
NONZERO_GAUGE = 4
QUALITY_GATE = 5
CANCELLED = 6
INVALID_ARGUMENT = 7
INVALID_SETTINGS = 8
INCOMPATIBLE_RHS = 9
BREAKDOWN = 10
ITERATION_LIMIT = 11
NONFINITE_INPUT = 12
NONZERO_WALL_VELOCITY = 13
ARITHMETIC_FAILURE = 14
INVALID_TIME_STEP = 15
INVALID_DENSITY = 16
TIME_RESOLUTION = 17
DIVERGENCE_LIMIT = 18
COURANT_LIMIT = 19
GEOMETRY = 20
BUFFER_LIMIT = 21
DEVICE_ERROR = 22
REFUSED = 23
INVALID_INITIAL_PRESSURE = 24
INVALID_STATE = 25


STATUS_NAMES = {v: k for k, v in dict(globals()).items() if isinstance(v, int) and k.isupper()}

# PressureInitializationPolicy (rh_init_policy)
INIT_COLD = 0
INIT_ACCEPTED_PREVIOUS = 1

F64, F32 = 0, 1
VARIANT_SCALAR_MG = 1
SMOOTHER_CHEBYSHEV = 0  # multigrid smoothers (rh_solver_config.mg_smoother)
SMOOTHER_JACOBI = 1
JACOBI_PCG, RBSGS_PCG, MG_PCG, MIXED_MG, SGS_PCG = 0, 1, 2, 3, 4
METHOD_IDS = {"jacobi-pcg-v1": JACOBI_PCG, "rbsgs-pcg": RBSGS_PCG, "mg-pcg": MG_PCG,
              "mixed-mg-dc": MIXED_MG, "mg-f64-dc": MIXED_MG, "sgs-pcg-v1": SGS_PCG}


class RheonError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        msg = (N.lib().rh_last_error().decode(errors="replace")
               if status in (DEVICE_ERROR, INVALID_STATE) else "")
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} {msg}".strip())
        self.status = status


def _check(status: int, where: str) -> None:
    if status != OK:
        raise RheonError(status, where)


def lib():
    return N.lib()


def precision_of(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported storage dtype {dt}")


def dtype_of(precision: int):
    return np.float64 if precision == F64 else np.float32


def make_grid(spec: GridSpec) -> N.Grid:
    g = N.Grid()
    for d in range(3):
        g.counts[d] = int(spec.counts[d])
        g.spacing[d] = float(spec.spacing[d])
        g.origin[d] = float(spec.origin[d])
    return g


def device_info():
    n = C.c_int32()
    arch = C.create_string_buffer(64)
    _check(lib().rh_device_info(C.byref(n), arch, 64), "rh_device_info")
    return n.value, arch.value.decode()


class DeviceArray:
    """Device memory owned through rh_device_alloc/rh_device_free."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        ptr = C.c_void_p()
        _check(lib().rh_device_alloc(self.nbytes, C.byref(ptr)), "rh_device_alloc")
        self.ptr = ptr.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.count
        _check(lib().rh_copy_to_device(self.ptr, a.ctypes.data, self.nbytes), "rh_copy_to_device")

    def download(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        _check(lib().rh_copy_to_host(out.ctypes.data, self.ptr, self.nbytes), "rh_copy_to_host")
        return out

    def free(self) -> None:
        if self.ptr:
            lib().rh_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _faces(vel, dtype=None):
    arrs = [x if isinstance(x, DeviceArray) else DeviceArray.from_host(np.ascontiguousarray(x, dtype or x.dtype))
            for x in vel]
    return arrs


# ----------------------------------------------------------- operator calls
def apply_full(spec: GridSpec, density: float, x: np.ndarray) -> np.ndarray:
    """PressureOperator::apply_full (src/operator.rs:378-456) on the device."""
    g = make_grid(spec)
    dx = DeviceArray.from_host(np.ascontiguousarray(x, np.float64))
    dy = DeviceArray(dx.count, np.float64)
    _check(lib().rh_apply_full(C.byref(g), density, dx.ptr, dy.ptr), "rh_apply_full")
    return dy.download()


def advect_velocity(spec: GridSpec, vel, dt: float):
    """advect_velocity (src/advection.rs:28-68); storage = dtype of vel."""
    g = make_grid(spec)
    prec = precision_of(vel[0].dtype)
    src = _faces(vel)
    dst = [DeviceArray(a.count, a.dtype) for a in src]
    _check(lib().rh_advect_velocity(C.byref(g), prec, src[0].ptr, src[1].ptr, src[2].ptr, dt,
                                    dst[0].ptr, dst[1].ptr, dst[2].ptr), "rh_advect_velocity")
    return [d.download() for d in dst]


def advect_tracer(spec: GridSpec, old: np.ndarray, vel, dt: float) -> np.ndarray:
    """advect_tracer (src/advection.rs:72-102), f32."""
    g = make_grid(spec)
    src = _faces(vel, np.float32)
    dold = DeviceArray.from_host(np.ascontiguousarray(old, np.float32))
    out = DeviceArray(dold.count, np.float32)
    _check(lib().rh_advect_tracer(C.byref(g), dold.ptr, src[0].ptr, src[1].ptr, src[2].ptr, dt, out.ptr),
           "rh_advect_tracer")
    return out.download()


def build_rhs(spec: GridSpec, vel, dt: float) -> np.ndarray:
    """PressureOperator::build_rhs (src/operator.rs:460-509)."""
    g = make_grid(spec)
    prec = precision_of(vel[0].dtype)
    src = _faces(vel)
    out = DeviceArray(spec.cell_len, np.float64)
    _check(lib().rh_build_rhs(C.byref(g), prec, src[0].ptr, src[1].ptr, src[2].ptr, dt, out.ptr), "rh_build_rhs")
    return out.download()


def correct_velocity(spec: GridSpec, density: float, pressure: np.ndarray, dt: float, vel):
    """correct_candidate_in_place (src/operator.rs:627-668); returns corrected copies."""
    g = make_grid(spec)
    prec = precision_of(vel[0].dtype)
    faces = _faces(vel)
    dp = DeviceArray.from_host(np.ascontiguousarray(pressure, np.float64))
    _check(lib().rh_correct_velocity(C.byref(g), prec, density, dp.ptr, dt, faces[0].ptr, faces[1].ptr,
                                     faces[2].ptr), "rh_correct_velocity")
    return [f.download() for f in faces]


def divergence_max(spec: GridSpec, vel) -> float:
    """PressureWorkspace::actual_divergence_max (src/pressure.rs:330-342)."""
    g = make_grid(spec)
    prec = precision_of(vel[0].dtype)
    faces = _faces(vel)
    out = C.c_double()
    _check(lib().rh_divergence_max(C.byref(g), prec, faces[0].ptr, faces[1].ptr, faces[2].ptr, C.byref(out)),
           "rh_divergence_max")
    return out.value


def velocity_rate(spec: GridSpec, vel) -> float:
    """velocity_rate (src/simulation.rs:1415-1429)."""
    g = make_grid(spec)
    prec = precision_of(vel[0].dtype)
    faces = _faces(vel)
    out = C.c_double()
    _check(lib().rh_velocity_rate(C.byref(g), prec, faces[0].ptr, faces[1].ptr, faces[2].ptr, C.byref(out)),
           "rh_velocity_rate")
    return out.value


def precondition(spec: GridSpec, density: float, r: np.ndarray, method: str = "jacobi-pcg-v1"):
    """One preconditioner application; returns (r with r[0]=0, z, rz)."""
    g = make_grid(spec)
    dr = DeviceArray.from_host(np.ascontiguousarray(r, np.float64))
    dz = DeviceArray(dr.count, np.float64)
    rz = C.c_double()
    _check(lib().rh_precondition(C.byref(g), density, METHOD_IDS[method], dr.ptr, dz.ptr, C.byref(rz)),
           "rh_precondition")
    return dr.download(), dz.download(), rz.value


# ----------------------------------------------------------- pressure solve
@dataclass
class PressureSettings:
    """PressureSettings (src/pressure.rs:33-39)."""
    relative_residual: float = 1e-9
    absolute_residual: float = 1e-12
    divergence_limit: float = 1e-7
    max_iterations: int = 2000
    poll_interval: int = 0

    def native(self) -> N.PressureSettings:
        return N.PressureSettings(self.relative_residual, self.absolute_residual, self.divergence_limit,
                                  int(self.max_iterations), int(self.poll_interval), 0)


@dataclass
class PressureReport:
    """PressureReport (src/pressure.rs:43-49) + error payload and timing."""
    status: int
    iterations: int
    true_residual_l2: float
    true_residual_max: float
    predicted_divergence_max: float
    error_iteration: int = 0
    error_value: float = 0.0
    threshold: float = 0.0
    seconds: float = 0.0

    @classmethod
    def from_native(cls, status: int, r: N.PressureReport) -> "PressureReport":
        return cls(status, r.iterations, r.true_residual_l2, r.true_residual_max, r.predicted_divergence_max,
                   r.error_iteration, r.error_value, r.threshold, r.seconds)

    @property
    def ok(self) -> bool:
        return self.status == OK


def _cancel_cb(cancelled):
    if cancelled is None:
        return N.CANCEL_FN(0), None
    cb = N.CANCEL_FN(lambda _ctx: 1 if cancelled() else 0)
    return cb, cb


class PressureWorkspace:
    """PressureWorkspace on the device (src/pressure.rs:243-740)."""

    def __init__(self, spec: GridSpec, method: str = "jacobi-pcg-v1", precision: int = F64,
                 memory_limit: int = 0, mg_levels: int = 0, kernel_variant: int = 0, mg_sweeps=(0, 0),
                 mg_omega: float = 0.0, mg_smoother: int = 0, mg_coarse_sweeps=(0, 0)):
        self.spec = spec
        self.grid = make_grid(spec)
        cfg = N.SolverConfig(METHOD_IDS[method], precision, mg_levels, int(mg_sweeps[0]), int(mg_sweeps[1]),
                             float(mg_omega), memory_limit, kernel_variant, int(mg_smoother),
                             int(mg_coarse_sweeps[0]), int(mg_coarse_sweeps[1]))
        h = C.c_void_p()
        _check(lib().rh_solver_create(C.byref(self.grid), C.byref(cfg), C.byref(h)), "rh_solver_create")
        self.handle = h.value
        self.pressure = DeviceArray(spec.cell_len, np.float64)

    def allocated_bytes(self) -> int:
        return int(lib().rh_solver_allocated_bytes(self.handle))

    def set_solid(self, solid) -> None:
        """Solid obstacle cells (one flag per cell, nonzero = solid) or None
        for the sealed box (StaticObstaclePressure's discretisation,
        src/obstacle_pressure.rs:224-309, with all-or-nothing apertures)."""
        if solid is None:
            _check(lib().rh_solver_set_solid(self.handle, None, 0), "rh_solver_set_solid")
            return
        a = np.ascontiguousarray(solid, np.uint8).reshape(-1)
        _check(lib().rh_solver_set_solid(self.handle, a.ctypes.data, a.size), "rh_solver_set_solid")

    def solve_rhs(self, density: float, rhs, dt: float, settings: PressureSettings, cancelled=None,
                  initial=None) -> PressureReport:
        """PressureWorkspace::solve_rhs (src/pressure.rs:365-393), or
        solve_rhs_with_initial_pressure (:413-429) when `initial` is given (a
        host array, a DeviceArray, or "pressure" for this workspace's last
        pressure); returns the report with its status (pressure() holds the
        last iterate)."""
        drhs = rhs if isinstance(rhs, DeviceArray) else DeviceArray.from_host(np.ascontiguousarray(rhs, np.float64))
        rep = N.PressureReport()
        st = settings.native()
        cb, _keep = _cancel_cb(cancelled)
        if initial is None:
            status = lib().rh_solver_solve_rhs(self.handle, density, drhs.ptr, dt, C.byref(st), self.pressure.ptr,
                                               C.byref(rep), cb, None)
        else:
            if isinstance(initial, str) and initial == "pressure":
                seed = self.pressure
            elif isinstance(initial, DeviceArray):
                seed = initial
            else:
                seed = DeviceArray.from_host(np.ascontiguousarray(initial, np.float64))
            status = lib().rh_solver_solve_rhs_initial(self.handle, density, drhs.ptr, seed.ptr, dt, C.byref(st),
                                                       self.pressure.ptr, C.byref(rep), cb, None)
        return PressureReport.from_native(status, rep)

    def solve_velocity(self, density: float, vel, dt: float, settings: PressureSettings, cancelled=None):
        """PressureWorkspace::solve_velocity (src/pressure.rs:344-361)."""
        return self.solve_rhs(density, build_rhs(self.spec, vel, dt), dt, settings, cancelled)

    def get_pressure(self) -> np.ndarray:
        return self.pressure.download()

    def close(self):
        if self.handle:
            lib().rh_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ provider seam
@dataclass
class PressureRequest:
    """PressureRequest (src/solver_provider.rs:47-52)."""
    spec: GridSpec
    density: float
    dt: float
    rhs: np.ndarray

    def native(self):
        g = make_grid(self.spec)
        rhs = np.ascontiguousarray(self.rhs, np.float64)
        req = N.PressureRequest(C.pointer(g), self.density, self.dt, rhs.ctypes.data, rhs.size)
        return req, (g, rhs)


class ProviderRejection(Exception):
    """ProviderRejection (src/solver_provider.rs:54-62)."""

    def __init__(self, status: int, reason: str | None = None):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {reason}")
        self.status = status
        self.reason = reason


def _metadata_dict(m) -> dict:
    return {k: (getattr(m, k).decode() if isinstance(getattr(m, k), bytes) else list(getattr(m, k)))
            for k, _ in N.SolverMetadata._fields_}


class GpuPressureProvider:
    """impl PressureProvider (src/solver_provider.rs:67-76) on the device."""

    def __init__(self, method: str = "jacobi-pcg-v1", precision: int = F64,
                 settings: PressureSettings | None = None, shape=None):
        settings = settings or PressureSettings()
        cfg = N.ProviderConfig(N.SolverConfig(METHOD_IDS[method], precision, 0, 0, 0, 0.0, 0), settings.native(),
                               (C.c_uint64 * 3)(*(shape or (0, 0, 0))))
        h = C.c_void_p()
        _check(lib().rh_provider_create(C.byref(cfg), C.byref(h)), "rh_provider_create")
        self.handle = h.value

    def metadata(self) -> dict:
        m = N.SolverMetadata()
        _check(lib().rh_provider_metadata(self.handle, C.byref(m)), "rh_provider_metadata")
        return _metadata_dict(m)

    def supports(self, request: PressureRequest) -> None:
        req, _keep = request.native()
        reason = C.c_char_p()
        st = lib().rh_provider_supports(self.handle, C.byref(req), C.byref(reason))
        if st != OK:
            raise ProviderRejection(st, reason.value.decode() if reason.value else None)

    def predict(self, request: PressureRequest, pressure: np.ndarray, cancelled=None) -> None:
        req, _keep = request.native()
        reason = C.c_char_p()
        cb, _k2 = _cancel_cb(cancelled)
        assert pressure.dtype == np.float64 and pressure.flags["C_CONTIGUOUS"]
        st = lib().rh_provider_predict(self.handle, C.byref(req), pressure.ctypes.data, pressure.size, cb, None,
                                       C.byref(reason))
        if st != OK:
            raise ProviderRejection(st, reason.value.decode() if reason.value else None)

    def close(self):
        if self.handle:
            lib().rh_provider_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ProviderReport:
    """ProviderReport (src/solver_provider.rs:78-92)."""
    selected: str
    attempted: dict | None  # the offered provider's metadata (Option<SolverMetadata>)
    rejection: int
    candidate: PressureReport | None
    accepted: PressureReport
    total: float
    inference: float
    qualification: float
    fallback: float


class ProviderWorkspace:
    """ProviderWorkspace (src/solver_provider.rs:94-327): qualification on the
    device; native fallback = device jacobi-pcg-v1."""

    def __init__(self, spec: GridSpec, limit: int = 0, fallback: str = "jacobi-pcg-v1"):
        self.spec = spec
        g = make_grid(spec)
        h = C.c_void_p()
        _check(lib().rh_provider_workspace_create(C.byref(g), limit, METHOD_IDS[fallback], C.byref(h)),
               "rh_provider_workspace_create")
        self.handle = h.value

    def solve(self, request: PressureRequest, settings: PressureSettings, preferred: GpuPressureProvider | None,
              cancelled=None) -> ProviderReport:
        req, _keep = request.native()
        st = settings.native()
        rep = N.ProviderReport()
        cb, _k2 = _cancel_cb(cancelled)
        status = lib().rh_provider_workspace_solve(self.handle, C.byref(req), C.byref(st),
                                                   preferred.handle if preferred else None, cb, None, C.byref(rep))
        if status != OK:
            raise RheonError(status, "ProviderWorkspace::solve")
        cand = PressureReport.from_native(OK, rep.candidate) if rep.has_candidate else None
        return ProviderReport(rep.selected.decode(), _metadata_dict(rep.attempted) if rep.has_attempted else None,
                              rep.rejection, cand,
                              PressureReport.from_native(OK, rep.accepted), rep.total, rep.inference,
                              rep.qualification, rep.fallback)

    def pressure(self) -> np.ndarray:
        out = np.empty(self.spec.cell_len, np.float64)
        _check(lib().rh_provider_workspace_pressure(self.handle, out.ctypes.data, out.size),
               "rh_provider_workspace_pressure")
        return out

    def close(self):
        if self.handle:
            lib().rh_provider_workspace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ scene
@dataclass
class StepReport:
    status: int
    dt: float
    pressure: PressureReport
    actual_divergence_max: float
    courant: float
    seconds: float
    stage_seconds: dict = field(default_factory=dict)
    time: float = 0.0
    generation: int = 0
    tracer_integral: float = 0.0
    kinetic_energy: float = 0.0


class Scene:
    """Device-resident fixed-box carrier: accepted and candidate faces stay in
    HBM; step() runs advect -> RHS -> solve -> correct -> gates -> commit
    (src/simulation.rs:988-1277, commit :805-811)."""

    def __init__(self, spec: GridSpec, density: float = 1.0, precision: int = F32, method: str = "jacobi-pcg-v1",
                 solver_precision: int | None = None, settings: PressureSettings | None = None,
                 actual_divergence_limit: float = 1e-5, max_courant: float = 0.0, mg_levels: int = 0,
                 kernel_variant: int = 0, mg_sweeps=(0, 0), mg_smoother: int = 0, mg_coarse_sweeps=(0, 0)):
        self.spec = spec
        self.precision = precision
        self.grid = make_grid(spec)
        settings = settings or PressureSettings()
        sp = precision if solver_precision is None else solver_precision
        cfg = N.SceneConfig(density, precision,
                            N.SolverConfig(METHOD_IDS[method], sp, mg_levels, int(mg_sweeps[0]), int(mg_sweeps[1]),
                                           0.0, 0, kernel_variant, int(mg_smoother),
                                           int(mg_coarse_sweeps[0]), int(mg_coarse_sweeps[1])),
                            settings.native(), actual_divergence_limit, max_courant)
        h = C.c_void_p()
        _check(lib().rh_scene_create(C.byref(self.grid), C.byref(cfg), C.byref(h)), "rh_scene_create")
        self.handle = h.value

    def allocated_bytes(self) -> int:
        return int(lib().rh_scene_allocated_bytes(self.handle))

    def set_solid(self, solid) -> None:
        """Solid obstacle cells (see PressureWorkspace.set_solid); obstacle
        faces of the velocity are closed from now on."""
        if solid is None:
            _check(lib().rh_scene_set_solid(self.handle, None, 0), "rh_scene_set_solid")
            return
        a = np.ascontiguousarray(solid, np.uint8).reshape(-1)
        _check(lib().rh_scene_set_solid(self.handle, a.ctypes.data, a.size), "rh_scene_set_solid")

    def set_velocity(self, vel) -> None:
        dt = dtype_of(self.precision)
        arrs = [np.ascontiguousarray(v, dt) for v in vel]
        for a in range(3):
            assert arrs[a].size == self.spec.face_len(a)
        _check(lib().rh_scene_set_velocity(self.handle, arrs[0].ctypes.data, arrs[1].ctypes.data,
                                           arrs[2].ctypes.data), "rh_scene_set_velocity")

    def get_velocity(self):
        dt = dtype_of(self.precision)
        out = [np.empty(self.spec.face_len(a), dt) for a in range(3)]
        _check(lib().rh_scene_get_velocity(self.handle, out[0].ctypes.data, out[1].ctypes.data,
                                           out[2].ctypes.data), "rh_scene_get_velocity")
        return out

    def get_pressure(self) -> np.ndarray:
        out = np.empty(self.spec.cell_len, np.float64)
        _check(lib().rh_scene_get_pressure(self.handle, out.ctypes.data, out.size), "rh_scene_get_pressure")
        return out

    def set_initialization(self, policy: int) -> None:
        """PressureInitializationPolicy of every later step's solve: INIT_COLD
        (p = 0) or INIT_ACCEPTED_PREVIOUS (seeded with the last committed
        step's pressure, feature accepted-pressure-warm-start)."""
        _check(lib().rh_scene_set_initialization(self.handle, policy), "rh_scene_set_initialization")

    def set_tracer(self, tracer) -> None:
        """Passive tracer (one f32 per cell) advected and committed with every
        step, or None to drop it."""
        if tracer is None:
            _check(lib().rh_scene_set_tracer(self.handle, None, 0), "rh_scene_set_tracer")
            return
        a = np.ascontiguousarray(tracer, np.float32).reshape(-1)
        _check(lib().rh_scene_set_tracer(self.handle, a.ctypes.data, a.size), "rh_scene_set_tracer")

    def get_tracer(self) -> np.ndarray:
        out = np.empty(self.spec.cell_len, np.float32)
        _check(lib().rh_scene_get_tracer(self.handle, out.ctypes.data, out.size), "rh_scene_get_tracer")
        return out

    @staticmethod
    def _report(status: int, rep) -> StepReport:
        return StepReport(status, rep.dt, PressureReport.from_native(status, rep.pressure),
                          rep.actual_divergence_max, rep.courant, rep.seconds,
                          {"advect": rep.advect_seconds, "rhs": rep.rhs_seconds, "solve": rep.solve_seconds,
                           "correct": rep.correct_seconds},
                          rep.time, int(rep.generation), rep.tracer_integral, rep.kinetic_energy)

    def step(self, dt: float, cancelled=None) -> StepReport:
        rep = N.StepReport()
        cb, _keep = _cancel_cb(cancelled)
        status = lib().rh_scene_step(self.handle, dt, C.byref(rep), cb, None)
        return self._report(status, rep)

    def step_requested(self, requested_dt: float, cancelled=None) -> StepReport:
        """Simulation::step's dt policy: min(requested, 0.5 max_courant / rate)."""
        rep = N.StepReport()
        cb, _keep = _cancel_cb(cancelled)
        status = lib().rh_scene_step_requested(self.handle, requested_dt, C.byref(rep), cb, None)
        return self._report(status, rep)

    @classmethod
    def _borrowed(cls, handle: int, spec: GridSpec, precision: int) -> "Scene":
        """A scene owned by a SceneBatch (not destroyed by this object)."""
        obj = cls.__new__(cls)
        obj.spec = spec
        obj.precision = precision
        obj.grid = make_grid(spec)
        obj.handle = handle
        obj._owned = False
        return obj

    def close(self):
        if self.handle and getattr(self, "_owned", True):
            lib().rh_scene_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = [n for n in dir() if not n.startswith("_")]


class SceneBatch:
    """Many device-resident scenes advanced together: one launch per stage
    over every scene, optionally one hipGraph replay per step (f32 faces,
    mixed-precision multigrid).  Per scene the step is Scene.step's."""

    def __init__(self, specs, density: float = 1.0, settings: PressureSettings | None = None,
                 actual_divergence_limit: float = 1e-3, max_courant: float = 0.0, max_cycles: int = 0,
                 use_graph: bool = True, kernel_variant: int = 0, mg_sweeps=(0, 0), mg_smoother: int = 0,
                 mg_coarse_sweeps=(0, 0), adaptive: bool = False):
        """use_graph: one hipGraph replay per step with max_cycles V-cycles;
        with adaptive, graph segments that run only the cycles the scenes
        need (the same results); without a graph, direct launches."""
        self.specs = list(specs)
        settings = settings or PressureSettings(1e-8, 1e-14, 1e-3, 40)
        cfg = N.BatchConfig(
            N.SceneConfig(density, F32, N.SolverConfig(MIXED_MG, F32, 0, int(mg_sweeps[0]), int(mg_sweeps[1]), 0.0,
                                                       0, kernel_variant, int(mg_smoother),
                                                       int(mg_coarse_sweeps[0]), int(mg_coarse_sweeps[1])),
                          settings.native(), actual_divergence_limit, max_courant),
            int(max_cycles), (2 if adaptive else 1) if use_graph else 0)
        grids = (N.Grid * len(self.specs))(*[make_grid(sp) for sp in self.specs])
        h = C.c_void_p()
        _check(lib().rh_batch_create(grids, len(self.specs), C.byref(cfg), C.byref(h)), "rh_batch_create")
        self.handle = h.value
        self._scenes = []
        for i, sp in enumerate(self.specs):
            sh = C.c_void_p()
            _check(lib().rh_batch_scene(self.handle, i, C.byref(sh)), "rh_batch_scene")
            self._scenes.append(Scene._borrowed(sh.value, sp, F32))

    def __len__(self) -> int:
        return len(self.specs)

    def scene(self, i: int) -> Scene:
        return self._scenes[i]

    def allocated_bytes(self) -> int:
        return int(lib().rh_batch_allocated_bytes(self.handle))

    def set_dt(self, dts) -> None:
        a = np.ascontiguousarray(dts, np.float64)
        _check(lib().rh_batch_set_dt(self.handle, a.ctypes.data, a.size), "rh_batch_set_dt")

    def set_variant(self, kernel_variant: int) -> None:
        _check(lib().rh_batch_set_variant(self.handle, int(kernel_variant)), "rh_batch_set_variant")

    def set_tuning(self, kernel_variant: int, block_divisor: int) -> None:
        t = N.BatchTuning(int(kernel_variant), int(block_divisor), 0.0)
        _check(lib().rh_batch_set_tuning(self.handle, C.byref(t)), "rh_batch_set_tuning")

    def autotune(self, steps: int = 1):
        """Time every candidate launch shape on these scenes; applies and
        returns the fastest, with the table of all candidates."""
        table = (N.BatchTuning * 32)()
        tried = C.c_size_t()
        best = N.BatchTuning()
        _check(lib().rh_batch_autotune(self.handle, int(steps), table, 32, C.byref(tried), C.byref(best)),
               "rh_batch_autotune")
        rows = [{"kernel_variant": t.kernel_variant, "block_divisor": t.block_divisor,
                 "seconds_per_step": t.seconds_per_step} for t in table[:tried.value]]
        return {"kernel_variant": best.kernel_variant, "block_divisor": best.block_divisor,
                "seconds_per_step": best.seconds_per_step}, rows

    def step(self, steps: int = 1):
        """Run `steps` steps; returns the StepReports of the last step."""
        n = len(self.specs)
        reps = (N.StepReport * n)()
        sts = (C.c_int * n)()
        _check(lib().rh_batch_step(self.handle, int(steps), reps, sts), "rh_batch_step")
        out = []
        for i in range(n):
            r = reps[i]
            out.append(StepReport(sts[i], r.dt, PressureReport.from_native(sts[i], r.pressure),
                                  r.actual_divergence_max, r.courant, r.seconds, {}))
        return out

    def close(self):
        for sc in self._scenes:
            sc.handle = None
        if self.handle:
            lib().rh_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


REQUEST_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class ProjectionServer:
    """Continuous-batching serving loop over device-resident scenes (see
    rh_server_* in include/rheon_gpu.h): submit() queues a projection run,
    tick() runs one batched step of every running request, admits queued
    ones by planned device bytes and retires finished ones."""

    def __init__(self, density: float = 1.0, settings: PressureSettings | None = None,
                 actual_divergence_limit: float = 1e-6, max_courant: float = 0.0, hbm_budget: int = 0,
                 max_active: int = 0, block_divisor: int = 1, kernel_variant: int = 0):
        settings = settings or PressureSettings(1e-8, 1e-14, 5e-7, 40)
        self.scene_config = N.SceneConfig(
            density, F32, N.SolverConfig(MIXED_MG, F32, 0, 0, 0, 0.0, 0, kernel_variant, 0), settings.native(),
            actual_divergence_limit, max_courant)
        cfg = N.ServerConfig(self.scene_config, int(hbm_budget), int(max_active), int(block_divisor),
                             int(kernel_variant), 0)
        h = C.c_void_p()
        _check(lib().rh_server_create(C.byref(cfg), C.byref(h)), "rh_server_create")
        self.handle = h.value
        self.specs = {}

    def plan_bytes(self, spec: GridSpec) -> int:
        g = make_grid(spec)
        return int(lib().rh_server_plan_bytes(C.byref(g), C.byref(self.scene_config)))

    def submit(self, spec: GridSpec, vel, dt: float, steps: int, solid=None) -> int:
        arrs = [np.ascontiguousarray(v, np.float32) for v in vel]
        for a in range(3):
            assert arrs[a].size == spec.face_len(a)
        sol = None if solid is None else np.ascontiguousarray(solid, np.uint8).reshape(-1)
        req = N.Request(make_grid(spec), arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data,
                        None if sol is None else sol.ctypes.data, float(dt), int(steps), 0)
        rid = C.c_int64()
        _check(lib().rh_server_submit(self.handle, C.byref(req), C.byref(rid)), "rh_server_submit")
        self.specs[rid.value] = spec
        return rid.value

    def tick(self) -> dict:
        t = N.TickReport()
        _check(lib().rh_server_tick(self.handle, C.byref(t)), "rh_server_tick")
        return {k: getattr(t, k) for k, _ in N.TickReport._fields_ if k != "reserved"}

    def state(self, rid: int) -> dict:
        s = N.RequestState()
        _check(lib().rh_server_state(self.handle, int(rid), C.byref(s)), "rh_server_state")
        return {"state": REQUEST_STATES[s.state], "status": s.status, "steps_done": s.steps_done,
                "submitted": s.submitted, "admitted": s.admitted, "finished": s.finished, "bytes": s.bytes,
                "actual_divergence_max": s.last.actual_divergence_max,
                "iterations": s.last.pressure.iterations}

    def result(self, rid: int):
        spec = self.specs[rid]
        out = [np.empty(spec.face_len(a), np.float32) for a in range(3)]
        p = np.empty(spec.cell_len, np.float64)
        _check(lib().rh_server_result(self.handle, int(rid), out[0].ctypes.data, out[1].ctypes.data,
                                      out[2].ctypes.data, p.ctypes.data), "rh_server_result")
        return out, p

    def close(self):
        if self.handle:
            lib().rh_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------- supplied-projection qualification
DOMAIN_SEALED_BOX, DOMAIN_FREE_SURFACE, DOMAIN_EMBEDDED_OBSTACLE, DOMAIN_MOVING_WALL = 0, 1, 2, 3
SELECTIONS = {0: "supplied", 1: "native_baseline", 2: "native_fallback"}
REJECTIONS = {0: None, 1: "field", 2: "momentum_witness", 3: "pressure", 4: "actual_divergence"}
REFUSALS = {0: None, 1: "domain", 2: "workspace_grid", 3: "input_binding", 4: "representation", 5: "identity"}


@dataclass
class ProjectionInput:
    """ProjectionInput (src/solver_provider/composition/projection.rs:31-42)."""
    spec: GridSpec
    velocity: list
    density: float = 1.25
    interval_start: float = 0.0
    dt: float = 0.01
    run_id: int = 1
    source_revision: str = "rheon-amd"
    generation: int = 0
    settings: PressureSettings = field(default_factory=lambda: PressureSettings(1e-10, 1e-12, 1e-8, 1000))
    actual_divergence_limit: float = 1e-5
    domain: int = DOMAIN_SEALED_BOX

    def native(self):
        g = make_grid(self.spec)
        vel = [np.ascontiguousarray(v, np.float32) for v in self.velocity]
        keep = [g, vel, self.source_revision.encode()]
        n = N.ProjectionInput(C.pointer(g), self.density, self.interval_start, self.dt, int(self.run_id), keep[2],
                              int(self.generation), (C.c_void_p * 3)(*[v.ctypes.data for v in vel]),
                              (C.c_size_t * 3)(*[v.size for v in vel]), self.settings.native(),
                              self.actual_divergence_limit, int(self.domain))
        return n, keep


def projection_input_id(inp: ProjectionInput) -> bytes:
    """ProjectionRequest::new: the input id (raises RheonError on an invalid input)."""
    n, _keep = inp.native()
    out = (C.c_uint8 * 32)()
    _check(lib().rh_projection_input_id(C.byref(n), out), "rh_projection_input_id")
    return bytes(out)


def projection_representation(spec: GridSpec):
    r = N.ProjectionRepresentation()
    g = make_grid(spec)
    _check(lib().rh_projection_representation_of(C.byref(g), C.byref(r)), "rh_projection_representation_of")
    return r


@dataclass
class ProjectionCandidate:
    """SuppliedProjection (projection.rs:187-193)."""
    input_id: bytes
    pressure: np.ndarray
    velocity: list
    representation: object = None
    algorithm: str = "classical-supplied-projection"
    version: str = "1"
    source_revision: str = "rheon-amd-supplied"

    def native(self, spec: GridSpec):
        p = np.ascontiguousarray(self.pressure, np.float64)
        vel = [np.ascontiguousarray(v, np.float32) for v in self.velocity]
        rep = self.representation if self.representation is not None else projection_representation(spec)
        texts = [self.algorithm.encode(), self.version.encode(), self.source_revision.encode()]
        n = N.ProjectionCandidate((C.c_uint8 * 32)(*self.input_id), rep, texts[0], texts[1], texts[2],
                                  p.ctypes.data, p.size, (C.c_void_p * 3)(*[v.ctypes.data for v in vel]),
                                  (C.c_size_t * 3)(*[v.size for v in vel]))
        return n, (p, vel, texts, rep)


def _provider_report(r) -> ProviderReport:
    cand = PressureReport.from_native(OK, r.candidate) if r.has_candidate else None
    return ProviderReport(r.selected.decode(), _metadata_dict(r.attempted) if r.has_attempted else None, r.rejection,
                          cand, PressureReport.from_native(OK, r.accepted), r.total, r.inference, r.qualification,
                          r.fallback)


class ProjectionWorkspace:
    """ProjectionWorkspace (projection.rs:275-479) on the device: qualify()
    accepts a supplied projection only through the momentum witness, the
    pressure requalification and the actual-divergence gate, else answers
    with the native device solve."""

    def __init__(self, spec: GridSpec, limit: int = 0, fallback: str = "jacobi-pcg-v1"):
        self.spec = spec
        g = make_grid(spec)
        h = C.c_void_p()
        _check(lib().rh_projection_workspace_create(C.byref(g), int(limit), METHOD_IDS[fallback], C.byref(h)),
               "rh_projection_workspace_create")
        self.handle = h.value

    def qualify(self, inp: ProjectionInput, candidate: ProjectionCandidate | None = None, cancelled=None):
        """Returns (status, report dict)."""
        n, _k1 = inp.native()
        if candidate is not None:
            c, _k2 = candidate.native(self.spec)
            cp = C.byref(c)
        else:
            cp = None
        rep = N.ProjectionReport()
        cb, _k3 = _cancel_cb(cancelled)
        st = lib().rh_projection_qualify(self.handle, C.byref(n), cp, cb, None, C.byref(rep))
        out = {"input_id": bytes(rep.input_id), "selection": SELECTIONS.get(rep.selection),
               "refusal": REFUSALS.get(rep.refusal), "rejection": REJECTIONS.get(rep.rejection),
               "rejection_status": rep.rejection_status, "witness": (rep.witness_axis, rep.witness_index),
               "wall": (rep.wall_axis, rep.wall_index), "rejected_divergence": rep.rejected_divergence,
               "supplied_identity": bool(rep.has_supplied_identity),
               "pressure": _provider_report(rep.pressure),
               "discarded": _provider_report(rep.discarded_candidate_pressure) if rep.has_discarded else None,
               "actual_divergence_max": rep.actual_divergence_max}
        return st, out

    def result(self):
        p = np.empty(self.spec.cell_len, np.float64)
        vel = [np.empty(self.spec.face_len(a), np.float32) for a in range(3)]
        _check(lib().rh_projection_result(self.handle, p.ctypes.data, p.size, vel[0].ctypes.data,
                                          vel[1].ctypes.data, vel[2].ctypes.data), "rh_projection_result")
        return p, vel

    def close(self):
        if self.handle:
            lib().rh_projection_workspace_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ dense3d frames
@dataclass
class Dense3dState:
    """The LiquidTransportView fields a dense3d frame serializes
    (src/dense3d_sequence.rs:415-495): the carrier faces and tracer (f32),
    the liquid fraction and pressure (f64) as host arrays or DeviceArrays,
    the clocks and the stamps ((id, version) pairs)."""
    velocity: tuple
    tracer: object
    fraction: object
    pressure: object
    carrier_time: float = 0.0
    liquid_time: float = 0.0
    carrier_generation: int = 0
    carrier_stamp: tuple = (0, 0)
    liquid_stamp: tuple = (0, 0)
    transport_only: bool = True


@dataclass
class Dense3dReport:
    """The LiquidStepReport fields a frame checks and prints (carrier.step and
    liquid of src/liquid_step.rs:112-131)."""
    carrier_time: float
    carrier_dt: float
    carrier_generation: int
    pressure_iterations: int = 0
    pressure_true_residual_max: float = 0.0
    carrier_actual_divergence_max: float = 0.0
    liquid_time: float = 0.0
    liquid_dt: float = 0.0
    liquid_stamp: tuple = (0, 0)
    liquid_flow: tuple = (0, 0)
    liquid_actual_divergence_max: float = 0.0
    liquid_volume_before: float = 0.0
    liquid_volume_after: float = 0.0
    liquid_mass_after: float = 0.0
    inward_boundary_volume: float = 0.0
    outward_boundary_volume: float = 0.0
    source_volume: float = 0.0
    volume_balance_error: float = 0.0
    volume_rounding_budget: float = 0.0


class Dense3dShapeError(RheonError):
    """A field array whose length is not the grid's (Dense3dError::InvalidState
    "face shape" / "cell shape"); raised before anything reaches the device."""

    def __init__(self, msg: str):
        RuntimeError.__init__(self, f"write_dense3d_frame: INVALID_STATE {msg}")
        self.status = INVALID_STATE


def format_sci(v: float, digits: int) -> str:
    """Rust's `{v:.(digits-1)e}` on the host (the frame's number format)."""
    buf = C.create_string_buffer(32)
    n = lib().rh_format_sci(float(v), int(digits), buf)
    if n < 0:
        raise ValueError("finite value and 10 or 18 digits required")
    return buf.value.decode()


def _dense3d_structs(spec: GridSpec, state: Dense3dState, report):
    """The C view and report; every array's length is checked against the
    grid first (the reference's "face shape" / "cell shape" refusals,
    src/dense3d_sequence.rs:431-442), since the device reads each array to
    the grid's length."""
    keep = []
    n = [int(c) for c in spec.counts]

    def length(a):
        return a.count if isinstance(a, DeviceArray) else int(np.asarray(a).size)

    def refuse(msg):
        raise Dense3dShapeError(msg)

    for axis in range(3):
        if length(state.velocity[axis]) != (n[0] + (axis == 0)) * (n[1] + (axis == 1)) * (n[2] + (axis == 2)):
            refuse("face shape")
    cells = n[0] * n[1] * n[2]
    if not all(length(a) == cells for a in (state.tracer, state.fraction, state.pressure)):
        refuse("cell shape")
    for a, dt in ((state.velocity[0], np.float32), (state.velocity[1], np.float32), (state.velocity[2], np.float32),
                  (state.tracer, np.float32), (state.fraction, np.float64), (state.pressure, np.float64)):
        if isinstance(a, DeviceArray) and a.dtype != np.dtype(dt):
            raise TypeError(f"dense3d field arrays: expected {np.dtype(dt)}, got {a.dtype}")

    def dev(a, dtype):
        if isinstance(a, DeviceArray):
            return a
        d = DeviceArray.from_host(np.ascontiguousarray(a, dtype))
        keep.append(d)
        return d

    view = N.Dense3dView()
    for d in range(3):
        view.velocity[d] = dev(state.velocity[d], np.float32).ptr
    view.tracer = dev(state.tracer, np.float32).ptr
    view.fraction = dev(state.fraction, np.float64).ptr
    view.pressure = dev(state.pressure, np.float64).ptr
    view.carrier_time = float(state.carrier_time)
    view.liquid_time = float(state.liquid_time)
    view.carrier_generation = int(state.carrier_generation)
    view.carrier_stamp.id, view.carrier_stamp.version = (int(x) for x in state.carrier_stamp)
    view.liquid_stamp.id, view.liquid_stamp.version = (int(x) for x in state.liquid_stamp)
    view.transport_only = 1 if state.transport_only else 0
    rep = None
    if report is not None:
        rep = N.Dense3dReport()
        for name, _ in N.Dense3dReport._fields_:
            val = getattr(report, name)
            if name in ("liquid_stamp", "liquid_flow"):
                getattr(rep, name).id, getattr(rep, name).version = (int(x) for x in val)
            else:
                setattr(rep, name, val)
    return view, rep, keep


def dense3d_frame_device(spec: GridSpec, state: Dense3dState, frame: int, report, out: "DeviceArray | None"):
    """write_dense3d_frame into device memory (rh_dense3d_frame).  Returns
    the line's length; with ``out`` None or too small raises BUFFER_LIMIT
    after setting ``needed`` on the error."""
    g = make_grid(spec)
    view, rep, keep = _dense3d_structs(spec, state, report)
    n = C.c_uint64()
    st = lib().rh_dense3d_frame(C.byref(g), C.byref(view), int(frame), C.byref(rep) if rep is not None else None,
                                out.ptr if out is not None else None, out.nbytes if out is not None else 0,
                                C.byref(n))
    if st != OK:
        err = RheonError(st, "rh_dense3d_frame")
        err.needed = n.value
        raise err
    return n.value


def write_dense3d_frame(spec: GridSpec, state: Dense3dState, frame: int, report=None) -> bytes:
    """write_dense3d_frame (src/dense3d_sequence.rs:415-564): the frame's
    JSONL line, formatted on the device and copied back.  A failed preflight
    raises RheonError(INVALID_STATE) with the reference's message."""
    try:
        dense3d_frame_device(spec, state, frame, report, None)
        raise AssertionError("rh_dense3d_frame wrote into a null buffer")
    except RheonError as e:
        if e.status != BUFFER_LIMIT:
            raise
        needed = e.needed
    out = DeviceArray(needed, np.uint8)
    n = dense3d_frame_device(spec, state, frame, report, out)
    return out.download()[:n].tobytes()
