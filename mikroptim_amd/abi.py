"""ctypes mirror of include/mikroptim_types.h (the plain-C layout of a system batch, residual
settings, variable map and LM options shared by libmikroptim_hip.so and its host layer)."""

import ctypes as C

MAX_SURFACES = 64
MAX_RINGS = 16

SURFACE_STOP, SURFACE_PLANE, SURFACE_SPHERE = 0, 1, 2
PUPIL_HALF, PUPIL_FULL, PUPIL_AXIS = 0, 1, 2
SOURCE_COLLIMATED, SOURCE_POINT = 0, 1
REF_LAST_SURFACE, REF_IMAGE_EXIT_PUPIL, REF_IMAGE_INFINITE = 0, 1, 2
FAILURE_SOFT, FAILURE_HARD = 0, 1
(STATUS_OK, STATUS_MISSED, STATUS_TIR, STATUS_BACKWARD, STATUS_UNSUPPORTED, STATUS_SOFT_MISSED,
 STATUS_SOFT_TIR) = range(7)
VAR_CURVATURE, VAR_THICKNESS, VAR_WORKING_DISTANCE, VAR_GLASS = 0, 1, 2, 3
STENCIL_CENTRAL, STENCIL_ONE_SIDED, STENCIL_TRUNCATED = 0, 1, 2
JACOBIAN_FD, JACOBIAN_ADJOINT = 0, 1
(LM_RUNNING, LM_CONVERGED_MERIT, LM_CONVERGED_STEP, LM_CONVERGED_GRADIENT, LM_CONVERGED_STALL,
 LM_MAX_ITERATIONS, LM_DAMPING_LIMIT, LM_NUMERICAL_ERROR) = range(8)
PRECISION_FP64, PRECISION_MIXED, PRECISION_DF64, PRECISION_FP64_MFMA = 0, 1, 2, 3
# ConstraintType / ConstraintMode / EdgeHeight (include/mikroptim/model/constraints.hpp:28,76,69)
(CON_FOCAL_LENGTH, CON_PARFOCAL_LENGTH, CON_WORKING_DISTANCE, CON_PARAXIAL_FOCUS,
 CON_PETZVAL) = range(5)
(CON_MAX_SEMI_DIAMETER, CON_MIN_CENTER_THICKNESS, CON_MAX_CENTER_THICKNESS,
 CON_MIN_EDGE_THICKNESS, CON_MAX_INCIDENCE_ANGLE) = range(7, 12)
CON_SINE_CONDITION, CON_DISTORTION = 12, 13
CON_GLASS_HULL = 6
MODE_EQUAL, MODE_MIN, MODE_MAX = 0, 1, 2
EDGE_SEMI_DIAMETER, EDGE_RAYS = 0, 1

LM_STATUS_NAMES = ["running", "converged_merit", "converged_step", "converged_gradient",
                   "converged_stall", "max_iterations", "damping_limit", "numerical_error"]

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)


class Batch(C.Structure):
    _fields_ = [
        ("variants", C.c_int32), ("surfaces", C.c_int32), ("wavelengths", C.c_int32),
        ("fields", C.c_int32), ("source", C.c_int32), ("reference_sphere", C.c_int32),
        ("reference_wavelength_nm", C.c_double), ("source_point", C.c_double * 3),
        ("source_na", C.c_double),
        ("type", _u8p), ("pupil_mode", _u8p), ("wavelength_weight", _dp), ("field_weight", _dp),
        ("curvature", _dp), ("thickness", _dp), ("index", _dp), ("pupil_radius", _dp),
        ("field_angle", _dp), ("field_target", _dp), ("reference_curvature", _dp),
        ("paraxial_index", _dp), ("wavelength_nm", _dp), ("buchdahl", _dp),
    ]


class Constraint(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("mode", C.c_int32), ("weight", C.c_double), ("target", C.c_double),
        ("surface", C.c_int32), ("next_surface", C.c_int32), ("offset", C.c_double),
        ("semi_diameter", C.c_double), ("next_semi_diameter", C.c_double),
        ("edge_height", C.c_int32), ("shift", C.c_double), ("zone", C.c_double),
        ("field", C.c_int32),
    ]


class ExtraRows(C.Structure):
    _fields_ = [("count", C.c_int32), ("items", C.POINTER(Constraint)),
                ("glass_hull", C.POINTER(C.c_double)), ("glass_hull_size", C.c_int32)]


class ResidualSettings(C.Structure):
    _fields_ = [("rings", C.c_int32), ("spokes", C.c_int32), ("failure", C.c_int32),
                ("fail_weight", C.c_double), ("precision", C.c_int32),
                ("extra", C.POINTER(ExtraRows))]


class Variables(C.Structure):
    _fields_ = [("count", C.c_int32), ("type", _i32p), ("batch_surface", _i32p),
                ("lower", _dp), ("upper", _dp), ("typical", _dp), ("component", _i32p)]


class Tuning(C.Structure):
    _fields_ = [("pack_rays", C.c_int32), ("fused_waves", C.c_int32),
                ("full_row_traces", C.c_int32)]


class LmOptions(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32), ("ftol", C.c_double), ("xtol", C.c_double),
        ("gtol", C.c_double), ("stall_window", C.c_int32), ("stall_tolerance", C.c_double),
        ("initial_damping", C.c_double), ("max_damping", C.c_double), ("diag_floor", C.c_double),
        ("reject_new_failures", C.c_int32), ("threads", C.c_int32),
        ("jacobian_method", C.c_int32), ("fail_weight_factor", C.c_double),
        ("max_fail_weight", C.c_double), ("host_control", C.c_int32),
    ]


PREFILTER_KEEP, PREFILTER_FAILED_RAYS, PREFILTER_MERIT_ABSOLUTE, PREFILTER_MERIT_RELATIVE = (
    0, 1, 2, 4)


class PrefilterOptions(C.Structure):
    _fields_ = [("max_failed_fraction", C.c_double), ("max_merit", C.c_double),
                ("max_merit_factor", C.c_double), ("boundary", C.c_double),
                ("opl_error_factor", C.c_double), ("merit_relative", C.c_double)]


def prefilter_options(max_failed_fraction=0.5, max_merit=float("inf"),
                      max_merit_factor=float("inf"), boundary=0.15, opl_error_factor=32.0,
                      merit_relative=0.01):
    """PrefilterCriteria + PrefilterMargins with the reference defaults
    (include/mikroptim/batch/batch_prefilter.hpp:32-56)."""
    return PrefilterOptions(max_failed_fraction, max_merit, max_merit_factor, boundary,
                            opl_error_factor, merit_relative)


class PrefilterResult(C.Structure):
    _fields_ = [("merit", _dp), ("optical_length", _dp), ("failed", C.POINTER(C.c_int64)),
                ("uncertain_ok", C.POINTER(C.c_int64)), ("uncertain_failed", C.POINTER(C.c_int64)),
                ("merit_lower", _dp), ("merit_upper", _dp), ("reject", _u8p),
                ("reference_merit", C.c_double), ("reference_variant", C.c_int32),
                ("kept", C.c_int32)]


class LmRun(C.Structure):
    _fields_ = [
        ("merit", C.c_double), ("failed_rays", C.c_int64), ("hard_failed_rays", C.c_int64),
        ("status", C.c_int32), ("iterations", C.c_int32), ("accepted", C.c_int32),
        ("jacobians", C.c_int32), ("damping", C.c_double), ("gradient_norm", C.c_double),
        ("model_decrease", C.c_double), ("fail_weight", C.c_double),
        ("penalty_rounds", C.c_int32),
    ]


def lm_options(max_iterations=100, ftol=1e-10, xtol=1e-10, gtol=1e-10, stall_window=0,
               stall_tolerance=1e-3, initial_damping=1e-3, max_damping=1e16, diag_floor=1e-12,
               reject_new_failures=True, threads=0, jacobian_method=JACOBIAN_FD,
               fail_weight_factor=10.0, max_fail_weight=1e14, host_control=False, graph=True):
    """LmOptions with the reference defaults (include/mikroptim/optim/levenberg_marquardt.hpp:85);
    host_control selects the host-side LM control of this library (default: device); graph=False
    launches the device cycle kernel by kernel instead of replaying a captured hipGraph."""
    mode = 1 if host_control else (0 if graph else 2)
    return LmOptions(max_iterations, ftol, xtol, gtol, stall_window, stall_tolerance,
                     initial_damping, max_damping, diag_floor, int(reject_new_failures), threads,
                     jacobian_method, fail_weight_factor, max_fail_weight, mode)


def ptr(array, ctype=C.c_double):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if array is None:
        return None
    return array.ctypes.data_as(C.POINTER(ctype))
