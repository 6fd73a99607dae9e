"""ctypes binding of the C ABI in ``include/neon_hip.h`` (library ``neon_amd/libneon_hip.so``).

The product path is the HIP library; there is no CPU fallback.  Loading fails loudly when
the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libneon_hip.so"

NEON_OK = 0
NEON_NONCONVERGED = 1
NEON_NAN_OR_INF = 2
NEON_NEGATIVE_JACOBIAN = 3
NEON_RETURN_MAP_FAILED = 4
NEON_BREAKDOWN = 5
NEON_FATAL = -1
NEON_BAD_ARGUMENT = -2
NEON_HIP_ERROR = -3
NEON_OOM = -4
NEON_NO_DEVICE = -5

MODEL_NEOHOOKE = 0
MODEL_J2 = 1
MODEL_LINEAR_ELASTIC = 2
MODEL_MICROSPHERE_GAUSSIAN = 3
MODEL_MICROSPHERE_LANGEVIN = 4
MODEL_MICROSPHERE_NONAFFINE = 5
MODEL_MICROSPHERE_AGEING = 6
MODEL_J2_DAMAGE = 7
SPHERE_BO33 = 0x100  # microsphere rule flag: 33 directions instead of 21
TANGENT_REFERENCE = 0
TANGENT_CONSISTENT = 1
SPECTRUM_LOWER = 0
SPECTRUM_UPPER = 1
SOLVER_CG = 0
SOLVER_BICGSTAB = 1
STOP_RELATIVE = 0
STOP_ABSOLUTE_PRECONDITIONED = 1
PRECISION_FP64 = 0
PRECISION_MIXED = 1
KERNEL_SPMV_FP64 = 0
KERNEL_SPMV_FP32 = 1
KERNEL_RESIDUAL_FP64 = 2
KERNEL_ASSEMBLY = 3
SPMV_NONTEMPORAL = 1
SPMV_SLICED = 256
SPMV_INDEX32 = 1024
SPMV_SYMMETRIC = 2048
# flags of the measurement library's neon_mesh_spmv_bench (include/neon_hip_measure.h)
MEASURE_DOT = 4096
MEASURE_DIRTY = 8192

STATUS_NAMES = {
    0: "OK",
    1: "NONCONVERGED",
    2: "NAN_OR_INF",
    3: "NEGATIVE_JACOBIAN",
    4: "RETURN_MAP_FAILED",
    5: "BREAKDOWN",
    -1: "FATAL",
    -2: "BAD_ARGUMENT",
    -3: "HIP_ERROR",
    -4: "OOM",
    -5: "NO_DEVICE",
}

P = C.POINTER
i32p = P(C.c_int32)
f64p = P(C.c_double)


class Dirichlet(C.Structure):
    _fields_ = [
        ("n_dofs", C.c_int32),
        ("dofs", i32p),
        ("n_points", C.c_int32),
        ("times", f64p),
        ("values", f64p),
    ]


class Neumann(C.Structure):
    _fields_ = [
        ("n_entries", C.c_int32),
        ("dofs", i32p),
        ("values", f64p),
        ("n_points", C.c_int32),
        ("times", f64p),
        ("loads", f64p),
        ("fixed_time", C.c_double),
    ]


# void (*)(void* user, int32_t step, double time)
StepCallback = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_double)


class NewtonOptions(C.Structure):
    _fields_ = [
        ("period", C.c_double),
        ("initial", C.c_double),
        ("minimum", C.c_double),
        ("maximum", C.c_double),
        ("adaptive", C.c_int32),
        ("residual_tolerance", C.c_double),
        ("displacement_tolerance", C.c_double),
        ("max_newton_iterations", C.c_int32),
        ("absolute_norm", C.c_int32),
        ("norm_mode", C.c_int32),
        ("solver_tolerance", C.c_double),
        ("solver_max_iterations", C.c_int32),
        ("solver_criterion", C.c_int32),
        ("solver", C.c_int32),
        ("precision", C.c_int32),
        ("max_load_steps", C.c_int32),
        ("max_total_iterations", C.c_int32),
        ("record_trajectory", C.c_int32),
        ("on_step", StepCallback),
        ("callback_user", C.c_void_p),
    ]


class NewtonRecord(C.Structure):
    _fields_ = [
        ("step", C.c_int32),
        ("iteration", C.c_int32),
        ("solver_iterations", C.c_int32),
        ("status", C.c_int32),
        ("step_time", C.c_double),
        ("displacement_norm", C.c_double),
        ("force_norm", C.c_double),
    ]


class NewtonTiming(C.Structure):
    _fields_ = [
        ("assemble_ms", C.c_double),
        ("dirichlet_ms", C.c_double),
        ("solve_ms", C.c_double),
        ("update_ms", C.c_double),
        ("total_ms", C.c_double),
        ("spmv_count", C.c_int64),
        ("refinement_count", C.c_int64),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "neon_device_count": (C.c_int, [P(C.c_int)]),
    "neon_context_create": (C.c_int, [C.c_int, P(C.c_void_p)]),
    "neon_context_destroy": (C.c_int, [C.c_void_p]),
    "neon_context_last_error": (C.c_char_p, [C.c_void_p]),
    "neon_context_synchronize": (C.c_int, [C.c_void_p]),
    "neon_context_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "neon_context_profile_read": (C.c_int, [C.c_void_p, C.c_int, C.c_int, P(C.c_int64), P(C.c_double)]),
    "neon_krylov_solve_csr": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int32, C.c_int64, i32p, i32p, f64p, f64p, f64p,
         C.c_double, C.c_int32, i32p, f64p],
    ),
    "neon_krylov_solve_csr_noalloc": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_int, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_double, C.c_int32, i32p, f64p],
    ),
    "neon_mesh_create": (C.c_int, [C.c_void_p, C.c_int32, f64p, C.c_int32, i32p, C.c_int32, f64p, C.c_int32]),
    "neon_model_traits": (C.c_int, [C.c_int32, i32p, i32p]),
    "neon_mesh_set_damage": (C.c_int, [C.c_void_p, f64p]),
    "neon_mesh_damage_state": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    "neon_model_variable_names": (C.c_char_p, [C.c_int32]),
    "neon_mesh_model": (C.c_int, [C.c_void_p, i32p]),
    "neon_mesh_set_parameters": (C.c_int, [C.c_void_p, f64p]),
    "neon_mesh_pattern": (C.c_int, [C.c_void_p, P(C.c_int64), i32p, i32p]),
    "neon_mesh_set_displacement": (C.c_int, [C.c_void_p, f64p]),
    "neon_mesh_update_internal_variables": (C.c_int, [C.c_void_p, C.c_double]),
    "neon_mesh_set_ageing": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "neon_mesh_ageing_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "neon_mesh_save_internal_variables": (C.c_int, [C.c_void_p, C.c_int32]),
    "neon_mesh_qp_state": (C.c_int, [C.c_void_p, f64p, f64p, f64p, f64p, f64p, f64p]),
    "neon_mesh_assemble": (C.c_int, [C.c_void_p, C.c_int32]),
    "neon_mesh_download": (C.c_int, [C.c_void_p, f64p, f64p]),
    "neon_mesh_enforce_dirichlet": (C.c_int, [C.c_void_p, C.c_int32, i32p, f64p]),
    "neon_mesh_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int32, f64p, f64p, i32p, f64p]),
    "neon_mesh_spmv_config": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "neon_newton_solve": (
        C.c_int,
        [C.c_void_p, P(NewtonOptions), C.c_int32, P(Dirichlet), f64p, P(NewtonRecord), C.c_int32, i32p,
         P(NewtonTiming)],
    ),
    "neon_adjoint_gradient": (C.c_int, [C.c_void_p, f64p, C.c_int32, i32p, C.c_double, C.c_int32, f64p, i32p]),
    "neon_natural_frequency": (C.c_int, [C.c_void_p, f64p, C.c_int32, i32p, C.c_int32, C.c_int32, C.c_double,
                                         C.c_int32, C.c_int32, f64p, f64p, i32p]),
    "neon_mesh_mass": (C.c_int, [C.c_void_p, f64p, f64p]),
    "neon_mesh_box_size": (
        C.c_int,
        [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, i32p, i32p],
    ),
    "neon_mesh_displacement": (C.c_int, [C.c_void_p, f64p]),
    "neon_mesh_sliced_info": (C.c_int, [C.c_void_p, i32p, P(C.c_int64), P(C.c_int64), i32p]),
    "neon_mesh_symmetric_info": (C.c_int, [C.c_void_p, i32p, i32p, i32p]),
    "neon_newton_solve_loaded": (
        C.c_int,
        [C.c_void_p, P(NewtonOptions), C.c_int32, P(Dirichlet), C.c_int32, P(Neumann), f64p, P(NewtonRecord),
         C.c_int32, i32p, P(NewtonTiming)],
    ),
    "neon_latin_solve": (
        C.c_int,
        [C.c_void_p, P(NewtonOptions), C.c_double, C.c_int32, P(Dirichlet), C.c_int32, P(Neumann), f64p,
         P(NewtonRecord), C.c_int32, i32p, P(NewtonTiming)],
    ),
    "neon_surface_load_vector": (C.c_int, [C.c_int32, f64p, C.c_int32, i32p, C.c_int32, C.c_int32, f64p]),
    "neon_volume_load_vector": (C.c_int, [C.c_int32, f64p, C.c_int32, i32p, C.c_int32, C.c_int32, f64p]),
    "neon_mesh_box": (
        C.c_int,
        [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_double,
         f64p, i32p, i32p],
    ),
    "neon_random_uniform": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_double, f64p]),
    "neon_random_normal": (C.c_int, [C.c_uint64, C.c_int64, f64p]),
}

# the measurement library (include/neon_hip_measure.h): timing probes over this library's contexts
MEASURE_PATH = pathlib.Path(__file__).resolve().parent / "libneon_hip_measure.so"
MEASURE_SIGNATURES = {
    "neon_mesh_spmv_bench": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p]),
    "neon_mesh_spmv_apply": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, f64p, f64p]),
    "neon_mesh_assembly_bench": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, f64p]),
    "neon_stream_probe": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, f64p, f64p]),
}

_lib = None
_measure = None


def lib() -> C.CDLL:
    """Load ``libneon_hip.so``; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -c 'from neon_amd import build; build.build()')"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def measure_lib() -> C.CDLL:
    """Load ``libneon_hip_measure.so`` (after ``libneon_hip.so``, whose contexts it takes)."""
    global _measure
    if _measure is None:
        lib()
        if not MEASURE_PATH.exists():
            raise RuntimeError(f"{MEASURE_PATH} is missing: build the HIP libraries first")
        handle = C.CDLL(str(MEASURE_PATH))
        for name, (res, args) in MEASURE_SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _measure = handle
    return _measure


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(P(ctype))


def f64(a):
    return ptr(a, C.c_double)


def i32(a):
    return ptr(a, C.c_int32)


class NeonError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


def device_count() -> int:
    c = C.c_int(0)
    lib().neon_device_count(C.byref(c))
    return c.value
