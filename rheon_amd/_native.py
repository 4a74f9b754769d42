"""ctypes binding of include/rheon_gpu.h (librheon_amd.so).

The library is loaded lazily on first use and never replaced by anything
else: if it is missing or fails to load, every call raises NativeMissing.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librheon_amd.so")


class NativeMissing(RuntimeError):
    pass


# ----------------------------------------------------------------- structs
class Grid(C.Structure):
    _fields_ = [("counts", C.c_uint64 * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3)]


class PressureSettings(C.Structure):
    _fields_ = [
        ("relative_residual", C.c_double),
        ("absolute_residual", C.c_double),
        ("divergence_limit", C.c_double),
        ("max_iterations", C.c_int64),
        ("poll_interval", C.c_int32),
        ("reserved", C.c_int32),
    ]


class PressureReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("true_residual_l2", C.c_double),
        ("true_residual_max", C.c_double),
        ("predicted_divergence_max", C.c_double),
        ("error_iteration", C.c_int64),
        ("error_value", C.c_double),
        ("threshold", C.c_double),
        ("seconds", C.c_double),
    ]


class SolverConfig(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("precision", C.c_int),
        ("mg_levels", C.c_int32),
        ("mg_pre_sweeps", C.c_int32),
        ("mg_post_sweeps", C.c_int32),
        ("mg_omega", C.c_double),
        ("memory_limit", C.c_uint64),
        ("kernel_variant", C.c_uint32),
        ("mg_smoother", C.c_uint32),
        ("mg_coarse_pre_sweeps", C.c_int32),
        ("mg_coarse_post_sweeps", C.c_int32),
    ]


class PressureRequest(C.Structure):
    _fields_ = [
        ("grid", C.POINTER(Grid)),
        ("density", C.c_double),
        ("dt", C.c_double),
        ("rhs", C.c_void_p),
        ("rhs_len", C.c_size_t),
    ]


class SolverMetadata(C.Structure):
    _fields_ = [
        ("id", C.c_char * 48),
        ("version", C.c_char * 16),
        ("stage", C.c_char * 16),
        ("shape_xyz", C.c_uint64 * 3),
        ("layout", C.c_char * 32),
        ("scalar", C.c_char * 8),
        ("units_in", C.c_char * 16),
        ("units_out", C.c_char * 16),
        ("domain", C.c_char * 48),
    ]


class ProviderConfig(C.Structure):
    _fields_ = [("solver", SolverConfig), ("settings", PressureSettings), ("shape_xyz", C.c_uint64 * 3)]


class ProviderReport(C.Structure):
    _fields_ = [
        ("selected", C.c_char * 48),
        ("has_attempted", C.c_int32),
        ("attempted", SolverMetadata),
        ("rejection", C.c_int),
        ("has_candidate", C.c_int32),
        ("candidate", PressureReport),
        ("accepted", PressureReport),
        ("total", C.c_double),
        ("inference", C.c_double),
        ("qualification", C.c_double),
        ("fallback", C.c_double),
    ]


class SceneConfig(C.Structure):
    _fields_ = [
        ("density", C.c_double),
        ("precision", C.c_int),
        ("solver", SolverConfig),
        ("pressure", PressureSettings),
        ("actual_divergence_limit", C.c_double),
        ("max_courant", C.c_double),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("scene", SceneConfig),
        ("max_cycles", C.c_int32),
        ("use_graph", C.c_int32),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("kernel_variant", C.c_uint32), ("block_divisor", C.c_int32), ("seconds_per_step", C.c_double)]


class StepReport(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("pressure", PressureReport),
        ("actual_divergence_max", C.c_double),
        ("courant", C.c_double),
        ("seconds", C.c_double),
        ("advect_seconds", C.c_double),
        ("rhs_seconds", C.c_double),
        ("solve_seconds", C.c_double),
        ("correct_seconds", C.c_double),
        ("time", C.c_double),
        ("generation", C.c_uint64),
        ("tracer_integral", C.c_double),
        ("kinetic_energy", C.c_double),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("scene", SceneConfig),
        ("hbm_budget", C.c_uint64),
        ("max_active", C.c_int32),
        ("block_divisor", C.c_int32),
        ("kernel_variant", C.c_uint32),
        ("reserved", C.c_int32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("grid", Grid),
        ("u", C.c_void_p),
        ("v", C.c_void_p),
        ("w", C.c_void_p),
        ("solid", C.c_void_p),
        ("dt", C.c_double),
        ("steps", C.c_int32),
        ("reserved", C.c_int32),
    ]


class RequestState(C.Structure):
    _fields_ = [
        ("state", C.c_int32),
        ("status", C.c_int),
        ("steps_done", C.c_int32),
        ("reserved", C.c_int32),
        ("submitted", C.c_double),
        ("admitted", C.c_double),
        ("finished", C.c_double),
        ("bytes", C.c_uint64),
        ("last", StepReport),
    ]


class TickReport(C.Structure):
    _fields_ = [
        ("stepped", C.c_int32),
        ("admitted", C.c_int32),
        ("finished", C.c_int32),
        ("active", C.c_int32),
        ("queued", C.c_int32),
        ("reserved", C.c_int32),
        ("cells", C.c_int64),
        ("bytes_in_use", C.c_uint64),
        ("seconds", C.c_double),
        ("device_seconds", C.c_double),
        ("admit_seconds", C.c_double),
        ("build_seconds", C.c_double),
        ("wait_seconds", C.c_double),
    ]



class ProjectionInput(C.Structure):
    _fields_ = [
        ("grid", C.POINTER(Grid)),
        ("density", C.c_double),
        ("interval_start", C.c_double),
        ("dt", C.c_double),
        ("run_id", C.c_uint64),
        ("source_revision", C.c_char_p),
        ("generation", C.c_uint64),
        ("velocity", C.c_void_p * 3),
        ("velocity_len", C.c_size_t * 3),
        ("settings", PressureSettings),
        ("actual_divergence_limit", C.c_double),
        ("domain", C.c_int),
    ]


class TensorMetadata(C.Structure):
    _fields_ = [("shape_xyz", C.c_uint64 * 3), ("layout", C.c_char * 32), ("scalar", C.c_char * 8),
                ("units", C.c_char * 16)]


class ProjectionRepresentation(C.Structure):
    _fields_ = [("version", C.c_uint32), ("operator_id", C.c_char * 64), ("pressure", TensorMetadata),
                ("velocity", TensorMetadata * 3)]


class ProjectionCandidate(C.Structure):
    _fields_ = [
        ("input_id", C.c_uint8 * 32),
        ("representation", ProjectionRepresentation),
        ("algorithm", C.c_char_p),
        ("version", C.c_char_p),
        ("source_revision", C.c_char_p),
        ("pressure", C.c_void_p),
        ("pressure_len", C.c_size_t),
        ("velocity", C.c_void_p * 3),
        ("velocity_len", C.c_size_t * 3),
    ]


class ProjectionReport(C.Structure):
    _fields_ = [
        ("input_id", C.c_uint8 * 32),
        ("selection", C.c_int32),
        ("refusal", C.c_int32),
        ("rejection", C.c_int32),
        ("rejection_status", C.c_int),
        ("witness_axis", C.c_int32),
        ("wall_axis", C.c_int32),
        ("witness_index", C.c_int64),
        ("wall_index", C.c_int64),
        ("rejected_divergence", C.c_double),
        ("has_supplied_identity", C.c_int32),
        ("has_discarded", C.c_int32),
        ("pressure", ProviderReport),
        ("discarded_candidate_pressure", ProviderReport),
        ("actual_divergence_max", C.c_double),
    ]


CANCEL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p)

_P = C.c_void_p
_G = C.POINTER(Grid)
_S = C.c_int

# name -> (restype, argtypes); every function declared in include/rheon_gpu.h

class VolumeStamp(C.Structure):
    _fields_ = [("id", C.c_uint64), ("version", C.c_uint64)]


class Dense3dView(C.Structure):
    _fields_ = [
        ("velocity", C.c_void_p * 3),
        ("tracer", C.c_void_p),
        ("fraction", C.c_void_p),
        ("pressure", C.c_void_p),
        ("carrier_time", C.c_double),
        ("liquid_time", C.c_double),
        ("carrier_generation", C.c_uint64),
        ("carrier_stamp", VolumeStamp),
        ("liquid_stamp", VolumeStamp),
        ("transport_only", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Dense3dReport(C.Structure):
    _fields_ = [
        ("carrier_time", C.c_double),
        ("carrier_dt", C.c_double),
        ("carrier_generation", C.c_uint64),
        ("pressure_iterations", C.c_uint64),
        ("pressure_true_residual_max", C.c_double),
        ("carrier_actual_divergence_max", C.c_double),
        ("liquid_time", C.c_double),
        ("liquid_dt", C.c_double),
        ("liquid_stamp", VolumeStamp),
        ("liquid_flow", VolumeStamp),
        ("liquid_actual_divergence_max", C.c_double),
        ("liquid_volume_before", C.c_double),
        ("liquid_volume_after", C.c_double),
        ("liquid_mass_after", C.c_double),
        ("inward_boundary_volume", C.c_double),
        ("outward_boundary_volume", C.c_double),
        ("source_volume", C.c_double),
        ("volume_balance_error", C.c_double),
        ("volume_rounding_budget", C.c_double),
    ]


SIGNATURES = {
    "rh_abi_version": (C.c_int32, []),
    "rh_last_error": (C.c_char_p, []),
    "rh_device_info": (_S, [C.POINTER(C.c_int32), C.c_char_p, C.c_size_t]),
    "rh_device_alloc": (_S, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "rh_device_free": (_S, [_P]),
    "rh_copy_to_device": (_S, [_P, _P, C.c_size_t]),
    "rh_copy_to_host": (_S, [_P, _P, C.c_size_t]),
    "rh_device_synchronize": (_S, []),
    "rh_copy_device": (_S, [_P, _P, C.c_size_t]),
    "rh_grid_validate": (_S, [_G, C.POINTER(C.c_uint64)]),
    "rh_buffer_plan": (_S, [_G] + [C.POINTER(C.c_uint64)] * 4),
    "rh_profile_enable": (_S, [C.c_int32]),
    "rh_profile_query": (_S, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "rh_profile_focus": (_S, [C.c_char_p]),
    "rh_apply_full": (_S, [_G, C.c_double, _P, _P]),
    "rh_advect_velocity": (_S, [_G, C.c_int, _P, _P, _P, C.c_double, _P, _P, _P]),
    "rh_advect_tracer": (_S, [_G, _P, _P, _P, _P, C.c_double, _P]),
    "rh_build_rhs": (_S, [_G, C.c_int, _P, _P, _P, C.c_double, _P]),
    "rh_correct_velocity": (_S, [_G, C.c_int, C.c_double, _P, C.c_double, _P, _P, _P]),
    "rh_divergence_max": (_S, [_G, C.c_int, _P, _P, _P, C.POINTER(C.c_double)]),
    "rh_divergence_max_noalloc": (_S, [_G, C.c_int, _P, _P, _P, C.POINTER(C.c_double)]),
    "rh_velocity_rate": (_S, [_G, C.c_int, _P, _P, _P, C.POINTER(C.c_double)]),
    "rh_precondition": (_S, [_G, C.c_double, C.c_int, _P, _P, C.POINTER(C.c_double)]),
    "rh_solver_create": (_S, [_G, C.POINTER(SolverConfig), C.POINTER(C.c_void_p)]),
    "rh_solver_destroy": (None, [_P]),
    "rh_solver_allocated_bytes": (C.c_uint64, [_P]),
    "rh_solver_set_solid": (_S, [_P, _P, C.c_size_t]),
    "rh_solver_solve_rhs": (_S, [_P, C.c_double, _P, C.c_double, C.POINTER(PressureSettings), _P,
                                 C.POINTER(PressureReport), CANCEL_FN, _P]),
    "rh_solver_solve_rhs_initial": (_S, [_P, C.c_double, _P, _P, C.c_double, C.POINTER(PressureSettings), _P,
                                         C.POINTER(PressureReport), CANCEL_FN, _P]),
    "rh_provider_create": (_S, [C.POINTER(ProviderConfig), C.POINTER(C.c_void_p)]),
    "rh_provider_destroy": (None, [_P]),
    "rh_provider_metadata": (_S, [_P, C.POINTER(SolverMetadata)]),
    "rh_provider_supports": (_S, [_P, C.POINTER(PressureRequest), C.POINTER(C.c_char_p)]),
    "rh_provider_predict": (_S, [_P, C.POINTER(PressureRequest), _P, C.c_size_t, CANCEL_FN, _P,
                                 C.POINTER(C.c_char_p)]),
    "rh_provider_workspace_create": (_S, [_G, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "rh_provider_workspace_destroy": (None, [_P]),
    "rh_provider_workspace_solve": (_S, [_P, C.POINTER(PressureRequest), C.POINTER(PressureSettings), _P,
                                         CANCEL_FN, _P, C.POINTER(ProviderReport)]),
    "rh_provider_workspace_pressure": (_S, [_P, _P, C.c_size_t]),
    "rh_scene_create": (_S, [_G, C.POINTER(SceneConfig), C.POINTER(C.c_void_p)]),
    "rh_scene_destroy": (None, [_P]),
    "rh_scene_allocated_bytes": (C.c_uint64, [_P]),
    "rh_scene_set_velocity": (_S, [_P, _P, _P, _P]),
    "rh_scene_set_solid": (_S, [_P, _P, C.c_size_t]),
    "rh_scene_get_velocity": (_S, [_P, _P, _P, _P]),
    "rh_scene_get_pressure": (_S, [_P, _P, C.c_size_t]),
    "rh_scene_device_velocity": (_S, [_P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "rh_scene_step": (_S, [_P, C.c_double, C.POINTER(StepReport), CANCEL_FN, _P]),
    "rh_projection_input_id": (_S, [C.POINTER(ProjectionInput), C.c_uint8 * 32]),
    "rh_projection_representation_of": (_S, [_G, C.POINTER(ProjectionRepresentation)]),
    "rh_projection_workspace_create": (_S, [_G, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "rh_projection_workspace_destroy": (None, [_P]),
    "rh_projection_qualify": (_S, [_P, C.POINTER(ProjectionInput), C.POINTER(ProjectionCandidate), CANCEL_FN, _P,
                                   C.POINTER(ProjectionReport)]),
    "rh_projection_result": (_S, [_P, _P, C.c_size_t, _P, _P, _P]),
    "rh_scene_step_requested": (_S, [_P, C.c_double, C.POINTER(StepReport), CANCEL_FN, _P]),
    "rh_scene_set_tracer": (_S, [_P, _P, C.c_size_t]),
    "rh_scene_set_initialization": (_S, [_P, C.c_int]),
    "rh_scene_get_tracer": (_S, [_P, _P, C.c_size_t]),
    "rh_batch_create": (_S, [C.POINTER(Grid), C.c_size_t, C.POINTER(BatchConfig), C.POINTER(C.c_void_p)]),
    "rh_batch_destroy": (None, [_P]),
    "rh_batch_allocated_bytes": (C.c_uint64, [_P]),
    "rh_batch_scene": (_S, [_P, C.c_size_t, C.POINTER(C.c_void_p)]),
    "rh_batch_set_dt": (_S, [_P, _P, C.c_size_t]),
    "rh_batch_set_variant": (_S, [_P, C.c_uint32]),
    "rh_batch_step": (_S, [_P, C.c_int32, C.POINTER(StepReport), C.POINTER(C.c_int)]),
    "rh_batch_set_tuning": (_S, [_P, C.POINTER(BatchTuning)]),
    "rh_server_create": (_S, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "rh_server_destroy": (None, [_P]),
    "rh_server_submit": (_S, [_P, C.POINTER(Request), C.POINTER(C.c_int64)]),
    "rh_server_tick": (_S, [_P, C.POINTER(TickReport)]),
    "rh_server_state": (_S, [_P, C.c_int64, C.POINTER(RequestState)]),
    "rh_server_result": (_S, [_P, C.c_int64, _P, _P, _P, _P]),
    "rh_server_plan_bytes": (C.c_uint64, [C.POINTER(Grid), C.POINTER(SceneConfig)]),
    "rh_batch_autotune": (_S, [_P, C.c_int32, C.POINTER(BatchTuning), C.c_size_t, C.POINTER(C.c_size_t),
                               C.POINTER(BatchTuning)]),
    "rh_dense3d_frame": (_S, [C.POINTER(Grid), C.POINTER(Dense3dView), C.c_uint64, C.POINTER(Dense3dReport),
                              _P, C.c_uint64, C.POINTER(C.c_uint64)]),
    "rh_format_sci": (C.c_int32, [C.c_double, C.c_int32, C.c_char_p]),
}

_lib = None


def lib():
    """Load librheon_amd.so (built by rheon_amd/csrc/Makefile)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeMissing(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - depends on the host
            raise NativeMissing(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def header_symbols() -> list[str]:
    """Function names declared in include/rheon_gpu.h."""
    import re
    path = os.path.join(os.path.dirname(_HERE), "include", "rheon_gpu.h")
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[a-z0-9_]+)\s*\(", text)) - {"rh_cancel_fn"})
