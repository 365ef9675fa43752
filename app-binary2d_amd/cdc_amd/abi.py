"""ctypes view of the C ABI declared in include/cdc_solver.h.

The structs mirror the header field for field; `check_layout()` compares
their sizes with the library's own `cdc_struct_size`.  Loading fails loudly
when the built library is missing: there is no CPU fallback.
"""
import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int32, c_int64, c_uint64, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcdc_hip.so")

CDC_OK = 0
CDC_ERR_NEGATIVE_DENSITY = 1
CDC_ERR_NEGATIVE_PRESSURE = 2
CDC_ERR_UNBOUND_ORBIT = 3
CDC_ERR_INVALID_ARGUMENT = 100
CDC_ERR_HIP = 101
CDC_ERR_OUT_OF_MEMORY = 102
CDC_ERR_NO_DEVICE = 103

CDC_ISOTHERMAL = 0
CDC_EULER = 1
CDC_VISC_CONSTANT_NU = 0
CDC_VISC_ALPHA = 1
CDC_SINK_IMPARTIAL = 0
CDC_SINK_TORQUE_FREE = 1
CDC_COOLING_T4 = 0
CDC_COOLING_BETA = 1

ITEMS = ("sink1", "sink2", "grav1", "grav2", "buffer", "cooling", "fake_mass")


class CMesh(ctypes.Structure):
    _fields_ = [("num_blocks", c_int32), ("block_size", c_int32), ("domain_radius", c_double)]


class CHydro(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("density_index", c_int32),
        ("mach_number", c_double),
        ("gamma_law_index", c_double),
        ("cooling_coefficient", c_double),
        ("cooling_slow_start", c_int32),
        ("has_density_floor", c_int32),
        ("density_floor", c_double),
        ("has_pressure_floor", c_int32),
        ("cooling_model", c_int32),
        ("pressure_floor", c_double),
        ("cooling_beta", c_double),
    ]


class CPhysics(ctypes.Structure):
    _fields_ = [
        ("buffer_rate", c_double),
        ("buffer_scale", c_double),
        ("binary_eccentricity", c_double),
        ("binary_mass_ratio", c_double),
        ("lambda_multiplier", c_double),
        ("viscosity_model", c_int32),
        ("sink_model", c_int32),
        ("viscosity", c_double),
        ("cfl", c_double),
        ("plm", c_double),
        ("has_plm_density_trigger", c_int32),
        ("one_body", c_int32),
        ("plm_density_trigger", c_double),
        ("fake_mass_threshold", c_double),
        ("fake_mass_rate", c_double),
        ("has_mach_ceiling", c_int32),
        ("rk_order", c_int32),
        ("mach_ceiling", c_double),
        ("sink_radius", c_double),
        ("sink_rate", c_double),
        ("has_softening_length", c_int32),
        ("has_dt_min", c_int32),
        ("softening_length", c_double),
        ("dt_min", c_double),
    ]


class CPointMass(ctypes.Structure):
    _fields_ = [
        ("mass", c_double),
        ("position_x", c_double),
        ("position_y", c_double),
        ("velocity_x", c_double),
        ("velocity_y", c_double),
    ]


class CError(ctypes.Structure):
    _fields_ = [
        ("code", c_int32),
        ("has_binary", c_int32),
        ("value", c_double),
        ("x", c_double),
        ("y", c_double),
        ("time", c_double),
        ("binary", CPointMass * 2),
        ("message", c_char * 256),
    ]


class CCavityModel(ctypes.Structure):
    _fields_ = [
        ("cavity_radius", c_double),
        ("cavity_exponent", c_double),
        ("density_floor", c_double),
        ("noise_amplitude", c_double),
        ("softening_length", c_double),
        ("mach_number", c_double),
        ("taper_radius", c_double),
        ("taper_width", c_double),
        ("seed", c_uint64),
    ]


class CClosureConfig(ctypes.Structure):
    _fields_ = [
        ("grid", c_int32),
        ("max_samples", c_int32),
        ("seed", c_uint64),
        ("domain_radius", c_double),
        ("feature_scale", c_double * 3),
        ("output_scale", c_double * 3),
        ("loss_weight", c_double * 3),
        ("init_output_gain", c_double),
        ("limiter", c_double),
        ("mask_density", c_double),
    ]


# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "cdc_abi_version": (c_int32, []),
    "cdc_struct_size": (c_int32, [c_int32]),
    "cdc_device_count": (c_int32, []),
    "cdc_math_probe": (c_int32, [c_void_p, c_void_p, c_int32]),
    "cdc_error_string": (c_char_p, [c_int32]),
    "cdc_last_error_message": (c_char_p, []),
    "cdc_solver_create": (c_int32, [POINTER(CMesh), POINTER(CHydro), POINTER(CPhysics),
                                     c_void_p, POINTER(c_void_p)]),
    "cdc_solver_destroy": (None, [c_void_p]),
    "cdc_solver_set_physics": (c_int32, [c_void_p, POINTER(CPhysics)]),
    "cdc_solver_set_state": (c_int32, [c_void_p, c_double, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "cdc_solver_get_state": (c_int32, [c_void_p, POINTER(c_double), POINTER(c_int64), POINTER(c_int64),
                                        c_void_p, c_void_p, c_void_p]),
    "cdc_solver_advance": (c_int32, [c_void_p, c_int32, POINTER(c_double), POINTER(CError)]),
    "cdc_solver_advance_dt": (c_int32, [c_void_p, c_int32, c_double, POINTER(CError)]),
    "cdc_solver_resample_from": (c_int32, [c_void_p, c_void_p]),
    "cdc_resample_state": (c_int32, [POINTER(CMesh), POINTER(CMesh), c_int32, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]),
    "cdc_solver_max_signal_speed": (c_int32, [c_void_p, POINTER(c_double)]),
    "cdc_solver_get_torque": (c_int32, [c_void_p, c_void_p]),
    "cdc_solver_set_torque": (c_int32, [c_void_p, c_void_p]),
    "cdc_max_signal_speed_noalloc": (c_int32, [c_int32, c_double, c_void_p, c_int64, c_void_p, c_void_p]),
    "cdc_snapshot_create": (c_int32, [c_void_p, POINTER(c_void_p)]),
    "cdc_snapshot_destroy": (None, [c_void_p]),
    "cdc_snapshot_time": (c_int32, [c_void_p, POINTER(c_double)]),
    "cdc_solver_save": (c_int32, [c_void_p, c_void_p]),
    "cdc_solver_restore": (c_int32, [c_void_p, c_void_p]),
    "cdc_solver_time_series_sample": (c_int32, [c_void_p, POINTER(c_double), c_void_p, c_void_p, c_void_p]),
    "cdc_solver_set_timing": (c_int32, [c_void_p, c_int32]),
    "cdc_solver_get_timing": (c_int32, [c_void_p, POINTER(c_int64), POINTER(c_double), POINTER(c_double)]),
    "cdc_solver_set_tile": (c_int32, [c_void_p, c_int32, c_int32]),
    "cdc_solver_get_tile": (c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "cdc_solver_set_kernel": (c_int32, [c_void_p, c_int32, c_int32, c_int32]),
    "cdc_solver_get_kernel": (c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    "cdc_solver_autotune": (c_int32, [c_void_p, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_double)]),
    "cdc_ensemble_create": (c_int32, [POINTER(c_void_p), c_int32, POINTER(c_void_p)]),
    "cdc_ensemble_destroy": (None, [c_void_p]),
    "cdc_ensemble_advance": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "cdc_ensemble_set_graph": (c_int32, [c_void_p, c_int32]),
    "cdc_ensemble_get_tile": (c_int32, [c_void_p, POINTER(c_int32), POINTER(c_int32)]),
    "cdc_ensemble_get_stats": (c_int32, [c_void_p, POINTER(c_int64), POINTER(c_double)]),
    "cdc_closure_create": (c_int32, [POINTER(CClosureConfig), POINTER(c_void_p)]),
    "cdc_closure_destroy": (None, [c_void_p]),
    "cdc_closure_num_params": (c_int64, []),
    "cdc_closure_get_params": (c_int32, [c_void_p, c_void_p]),
    "cdc_closure_set_params": (c_int32, [c_void_p, c_void_p]),
    "cdc_closure_get_config": (c_int32, [c_void_p, POINTER(CClosureConfig)]),
    "cdc_closure_set_scales": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "cdc_closure_set_interval": (c_int32, [c_void_p, c_int32]),
    "cdc_closure_features": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "cdc_closure_forward": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "cdc_closure_add_sample": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_double]),
    "cdc_closure_num_samples": (c_int32, [c_void_p, POINTER(c_int32), c_void_p]),
    "cdc_closure_clear_samples": (c_int32, [c_void_p]),
    "cdc_closure_sample_grad": (c_int32, [c_void_p, c_int32, POINTER(c_double), c_void_p]),
    "cdc_closure_train": (c_int32, [c_void_p, c_int32, c_double, c_uint64, c_void_p]),
    "cdc_closure_advance": (c_int32, [c_void_p, c_void_p, POINTER(c_double), POINTER(CError)]),
    "cdc_closure_advance_fold": (c_int32, [c_void_p, c_void_p, c_int32, POINTER(c_double), POINTER(CError)]),
    "cdc_orbital_state_from_time": (c_int32, [POINTER(CPhysics), c_double, POINTER(CPointMass)]),
    "cdc_model_seeded_cavity": (c_int32, [POINTER(CMesh), POINTER(CHydro), POINTER(CCavityModel), c_void_p]),
}

_lib = None


def load(path=None):
    """Load libcdc_hip.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP library {p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check_layout(lib=None):
    lib = lib or load()
    sizes = {0: CMesh, 1: CHydro, 2: CPhysics, 3: CError, 4: CCavityModel, 5: CClosureConfig}
    for k, cls in sizes.items():
        got = lib.cdc_struct_size(k)
        if got != ctypes.sizeof(cls):
            raise RuntimeError(f"struct {cls.__name__}: library {got} bytes, binding {ctypes.sizeof(cls)}")
    return True
