"""ctypes declarations for liboe.so (include/oe/oe.h).

The library is the product: there is no Python or CPU compute path behind
these bindings.  If liboe.so is missing, or no HIP device is present when an
engine is created, the calls raise instead of falling back.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboe.so")

OE_OK = 0
STATUS_NAMES = {0: "OK", 1: "CONFIG", 2: "IO", 3: "GRID", 4: "MATERIAL", 5: "NUMERICAL", 6: "XML", 7: "DEVICE"}


class OEError(RuntimeError):
    """Raised for a non-OK oe_status (mirrors crate::Error, src/lib.rs:61-88)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Excitation(C.Structure):
    _fields_ = [("i", C.c_uint32), ("j", C.c_uint32), ("k", C.c_uint32), ("direction", C.c_uint32),
                ("soft", C.c_uint32), ("samples", C.POINTER(C.c_float)), ("n_samples", C.c_uint64)]


class EnergySample(C.Structure):
    _fields_ = [("timestep", C.c_uint64), ("e_energy", C.c_double), ("h_energy", C.c_double),
                ("total_energy", C.c_double)]


class BatchDesc(C.Structure):
    _fields_ = [
        ("has_num_steps", C.c_int), ("num_steps", C.c_uint64),
        ("excitations", C.POINTER(Excitation)), ("n_excitations", C.c_size_t),
        ("extensions", C.POINTER(C.c_void_p)), ("n_extensions", C.c_size_t),
        ("has_max_timesteps", C.c_int), ("max_timesteps", C.c_uint64),
        ("has_energy_decay_db", C.c_int), ("energy_decay_db", C.c_double),
        ("check_interval", C.c_uint64),
        ("sample_interval", C.c_uint64), ("track_peak", C.c_int),
        ("persistent_peak", C.c_int),
        ("energy_samples", C.POINTER(EnergySample)), ("energy_samples_capacity", C.c_size_t),
    ]


class BatchResultC(C.Structure):
    _fields_ = [("timesteps_executed", C.c_uint64), ("termination_reason", C.c_int),
                ("final_decay_db", C.c_double), ("n_energy_samples", C.c_size_t),
                ("elapsed_seconds", C.c_double)]


class Waveform(C.Structure):
    _fields_ = [("type", C.c_int), ("t0", C.c_double), ("tau", C.c_double), ("frequency", C.c_double),
                ("times", C.POINTER(C.c_double)), ("values", C.POINTER(C.c_double)), ("n_points", C.c_size_t),
                ("amplitude", C.c_double)]


class DftConfig(C.Structure):
    _fields_ = [("n_signals", C.c_size_t), ("offsets", C.POINTER(C.c_uint64)), ("component", C.POINTER(C.c_uint32)),
                ("cell", C.POINTER(C.c_uint64)), ("weight", C.POINTER(C.c_double)), ("freqs", C.POINTER(C.c_double)),
                ("n_freqs", C.c_size_t), ("precision", C.c_int), ("record_steps", C.c_uint64),
                ("time_stride", C.c_uint32)]


class FarField(C.Structure):
    _fields_ = [("e_theta_re", C.POINTER(C.c_double)), ("e_theta_im", C.POINTER(C.c_double)),
                ("e_phi_re", C.POINTER(C.c_double)), ("e_phi_im", C.POINTER(C.c_double)),
                ("prad", C.c_double), ("dmax", C.c_double), ("gemm_ms", C.c_double), ("total_ms", C.c_double),
                ("gemm_flop", C.c_double), ("e_theta", C.POINTER(C.c_double)), ("e_phi", C.POINTER(C.c_double)),
                ("gemm_split_ms", C.c_double), ("gemm_split_flop", C.c_double)]


class SweepConfig(C.Structure):
    _fields_ = [("max_steps", C.c_uint64), ("has_energy_decay_db", C.c_int), ("energy_decay_db", C.c_double),
                ("check_interval", C.c_uint32), ("use_graph", C.c_int), ("autotune", C.c_int)]


class SweepEntry(C.Structure):
    _fields_ = [("timesteps_executed", C.c_uint64), ("termination_reason", C.c_int), ("final_decay_db", C.c_double),
                ("peak_energy", C.c_double), ("final_energy", C.c_double), ("elapsed_seconds", C.c_double)]


class ExtrapDims(C.Structure):
    _fields_ = [("L0", C.c_uint32), ("c1", C.c_uint32), ("c2", C.c_uint32), ("c3", C.c_uint32),
                ("hidden", C.c_uint32), ("modes", C.c_uint32), ("batch", C.c_uint32), ("tail", C.c_uint32)]


HOST_HOOK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_uint64, C.POINTER(C.POINTER(C.c_float)))
HOST_CHECK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_char), C.c_size_t)


class HostCallbackConfig(C.Structure):
    _fields_ = [("hook", HOST_HOOK_FN), ("check_termination", HOST_CHECK_FN), ("user", C.c_void_p),
                ("hook_mask", C.c_uint32)]


class CpmlConfig(C.Structure):
    _fields_ = [("layers", C.c_size_t * 6), ("grading_order", C.c_double), ("reflection_coeff", C.c_double),
                ("alpha_max", C.c_double), ("kappa_max", C.c_double), ("mode", C.c_int)]


# name -> (restype, argtypes); every function declared in include/oe/oe.h
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_sz = C.c_size_t
_d = C.c_double
_i = C.c_int
_u32 = C.c_uint32
_u64 = C.c_uint64
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_szp = C.POINTER(C.c_size_t)

SIGNATURES = {
    "oe_last_error": (C.c_char_p, []),
    "oe_version": (C.c_char_p, []),
    "oe_operator_create": (_i, [_sz, _sz, _sz, _dp, _pp]),
    "oe_operator_create_lines": (_i, [_dp, _sz, _dp, _sz, _dp, _sz, _pp]),
    "oe_operator_destroy": (None, [_vp]),
    "oe_operator_timestep": (_d, [_vp]),
    "oe_operator_dims": (None, [_vp, _szp]),
    "oe_operator_set_material": (_i, [_vp, _d, _d, _d, _d, _szp]),
    "oe_operator_set_pec": (_i, [_vp, _szp]),
    "oe_operator_add_lumped_element": (_i, [_vp, _i, _szp, _d]),
    "oe_engine_create": (_i, [_vp, _i, _pp]),
    "oe_engine_create_from_arrays": (_i, [_sz, _sz, _sz, _dp, _d, C.POINTER(_fp), _i, _pp]),
    "oe_engine_destroy": (None, [_vp]),
    "oe_engine_timestep": (_u64, [_vp]),
    "oe_engine_read_fields": (_i, [_vp, C.POINTER(_fp)]),
    "oe_engine_write_fields": (_i, [_vp, C.POINTER(_fp)]),
    "oe_engine_read_coefficients": (_i, [_vp, _i, _fp]),
    "oe_engine_material_table": (_i, [_vp, _i, C.POINTER(C.c_int)]),
    "oe_engine_fused_steps": (_i, [_vp, _i, C.POINTER(C.c_int)]),
    "oe_engine_graph_steps": (_i, [_vp, _i, C.POINTER(C.c_int)]),
    "oe_engine_reset": (_i, [_vp]),
    "oe_engine_energy": (_i, [_vp, _dp, _dp]),
    "oe_engine_run_batch": (_i, [_vp, C.POINTER(BatchDesc), C.POINTER(BatchResultC)]),
    "oe_engine_kernel_time": (_i, [_vp, C.c_char_p, C.POINTER(_u64), _dp]),
    "oe_engine_set_profiling": (_i, [_vp, _i]),
    "oe_engine_set_tiles": (_i, [_vp, _i, _i, _i, _i]),
    "oe_engine_get_tiles": (_i, [_vp, C.POINTER(_i)]),
    "oe_waveform_evaluate": (_i, [C.POINTER(Waveform), _d, _dp]),
    "oe_waveform_sample": (_i, [C.POINTER(Waveform), _u64, _u64, _d, _fp]),
    "oe_ext_cpml_create": (_i, [_vp, C.POINTER(CpmlConfig), _pp]),
    "oe_ext_mur_create": (_i, [_vp, C.POINTER(_i), _d, _pp]),
    "oe_ext_ade_create": (_i, [_vp, _i, _dp, C.POINTER(_u32), _sz, _i, _pp]),
    "oe_ext_field_probe_create": (_i, [_vp, _u32, _u32, _u32, _pp]),
    "oe_ext_probe_read": (_i, [_vp, _dp, _sz, _szp]),
    "oe_ext_dft_create": (_i, [_vp, C.POINTER(DftConfig), _pp]),
    "oe_ext_dft_read": (_i, [_vp, _dp, _dp, C.POINTER(_u64)]),
    "oe_ext_dft_read_series": (_i, [_vp, _dp, _sz, _szp]),
    "oe_ext_lumped_port_create": (_i, [_vp, _i, _szp, _i, _dp, _sz, _i, _u64, _pp]),
    "oe_ext_port_s11": (_i, [_vp, _d, _dp, _dp]),
    "oe_nf2ff_create": (_i, [_i, _d, _pp]),
    "oe_nf2ff_destroy": (None, [_vp]),
    "oe_nf2ff_add_surface": (_i, [_vp, _i, _i, _d, _dp, _sz, _dp, _sz, _dp, _dp]),
    "oe_nf2ff_calculate": (_i, [_vp, _dp, _sz, _dp, _sz, _d, _dp, C.POINTER(FarField)]),
    "oe_ext_nf2ff_create": (_i, [_vp, _szp, _i, _dp, _sz, _i, _u32, _pp]),
    "oe_ext_nf2ff_calculate": (_i, [_vp, _sz, _dp, _sz, _dp, _sz, _d, _dp, C.POINTER(FarField)]),
    "oe_sweep_create": (_i, [_i, _pp]),
    "oe_sweep_destroy": (None, [_vp]),
    "oe_sweep_add": (_i, [_vp, _vp, _vp, _vp, C.POINTER(Excitation), _sz, _szp]),
    "oe_sweep_run": (_i, [_vp, C.POINTER(SweepConfig), C.POINTER(SweepEntry)]),
    "oe_sweep_tuning": (_i, [_vp, C.POINTER(_i), _dp, C.POINTER(_i)]),
    "oe_sweep_set_ci": (_i, [_vp, _i]),
    "oe_sweep_energy_history": (_i, [_vp, _sz, C.POINTER(_u64), _dp, _sz, _szp]),
    "oe_extrap_create": (_i, [_i, C.POINTER(ExtrapDims), _fp, _u64, _pp]),
    "oe_extrap_destroy": (None, [_vp]),
    "oe_extrap_param_count": (_sz, [_vp]),
    "oe_extrap_params": (_i, [_vp, _fp, _sz]),
    "oe_extrap_set_params": (_i, [_vp, _fp, _sz]),
    "oe_extrap_forward": (_i, [_vp, _fp, _sz, _fp, _fp]),
    "oe_extrap_gradients": (_i, [_vp, _fp, _fp, _sz, _dp, _fp]),
    "oe_extrap_train_step": (_i, [_vp, _fp, _fp, _sz, C.c_float, _dp]),
    "oe_ext_upml_create": (_i, [_vp, _szp, _d, _d, _pp]),
    "oe_ext_upml_read": (_i, [_vp, _i, _i, _i, _fp, _sz, _szp, _szp, _szp]),
    "oe_ext_cpml_read_psi": (_i, [_vp, _i, _i, _i, _fp, _sz, _szp]),
    "oe_ext_ade_read_state": (_i, [_vp, _fp, _sz, _szp]),
    "oe_ext_host_callback_create": (_i, [_vp, C.POINTER(HostCallbackConfig), _pp]),
    "oe_ext_host_callback_reason": (_i, [_vp, C.c_char_p, _sz, _szp]),
    "oe_fields_energy_noalloc": (_i, [_vp, _sz, _vp, _i, _vp]),
    "oe_extension_reset": (_i, [_vp]),
    "oe_extension_destroy": (None, [_vp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load liboe.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m openems_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != OE_OK:
        raise OEError(status, lib().oe_last_error().decode(errors="replace"))
