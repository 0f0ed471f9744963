"""ctypes binding of libpawan_hip.so (the C ABI in include/pawan_hip.h).

The library is built in-tree (``pawan_amd/libpawan_hip.so``) by
``pawan_amd.build.build_library``.  Loading fails loudly when it is missing:
there is no other implementation of this path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpawan_hip.so")

PW_OK, PW_EINVAL, PW_ENOMEM, PW_EHIP, PW_ENONFINITE, PW_ENOTFOUND = 0, 1, 2, 3, 4, 5
KERNEL_HIGHORDER, KERNEL_GAUSSIAN = 0, 1
FP64, FP32C = 0, 1
SCHEME_DEFAULT, SCHEME_ONESIDED, SCHEME_SYMMETRIC = 0, 1, 2

_dp = C.POINTER(C.c_double)


class PwConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("kernel", C.c_int), ("precision", C.c_int),
                ("nu", C.c_double), ("uinf", C.c_double * 3), ("block", C.c_int),
                ("chunk", C.c_int), ("targets_per_thread", C.c_int), ("unroll", C.c_int),
                ("scheme", C.c_int), ("chunks", C.c_int)]


class PwVariant(C.Structure):
    _fields_ = [("block", C.c_int), ("targets_per_thread", C.c_int), ("unroll", C.c_int),
                ("scheme", C.c_int), ("chunks", C.c_int)]


class PwTuneEntry(C.Structure):
    _fields_ = [("variant", PwVariant), ("pair_ms", C.c_double), ("pairs_per_s", C.c_double)]


class PwJobSpec(C.Structure):
    _fields_ = [("n", C.c_size_t), ("pos", _dp), ("vort", _dp), ("radius", _dp),
                ("volume", _dp), ("nu", C.c_double), ("dt", C.c_double),
                ("uinf", C.c_double * 3), ("precision", C.c_int)]


class PwDiag(C.Structure):
    _fields_ = [("total_vorticity", C.c_double * 3), ("linear_impulse", C.c_double * 3),
                ("angular_impulse", C.c_double * 3), ("helicity", C.c_double),
                ("enstrophy", C.c_double), ("kinetic_energy", C.c_double)]


class PwServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("finished", C.c_int64),
                ("failed", C.c_int64), ("cancelled", C.c_int64), ("queued", C.c_int64),
                ("running", C.c_int64), ("steps", C.c_int64), ("pairs", C.c_double),
                ("busy_ms", C.c_double), ("step_p50_ms", C.c_double),
                ("step_p99_ms", C.c_double), ("job_p50_ms", C.c_double),
                ("job_p99_ms", C.c_double), ("hbm_in_use", C.c_size_t),
                ("hbm_budget", C.c_size_t), ("job_step_p50_ms", C.c_double),
                ("job_step_p99_ms", C.c_double)]


# name -> (restype, argtypes); every symbol declared in include/pawan_hip.h
SIGNATURES = {
    "pw_version": (C.c_char_p, []),
    "pw_config_default": (C.c_int, [C.POINTER(PwConfig)]),
    "pw_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(PwConfig)]),
    "pw_destroy": (C.c_int, [C.c_void_p]),
    "pw_last_error": (C.c_char_p, [C.c_void_p]),
    "pw_set_nu": (C.c_int, [C.c_void_p, C.c_double]),
    "pw_add_wake": (C.c_int, [C.c_void_p, C.c_size_t, _dp, _dp, _dp, _dp]),
    "pw_size": (C.c_size_t, [C.c_void_p]),
    "pw_num_wakes": (C.c_size_t, [C.c_void_p]),
    "pw_set_state": (C.c_int, [C.c_void_p, _dp]),
    "pw_get_state": (C.c_int, [C.c_void_p, _dp]),
    "pw_solve": (C.c_int, [C.c_void_p]),
    "pw_get_rates": (C.c_int, [C.c_void_p, _dp]),
    "pw_get_wake": (C.c_int, [C.c_void_p, C.c_size_t, _dp, _dp, _dp, _dp]),
    "pw_rk4_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_size_t]),
    "pw_euler_steps": (C.c_int, [C.c_void_p, C.c_double, C.c_size_t]),
    "pw_diagnostics": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pw_last_timing": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pw_get_variant": (C.c_int, [C.c_void_p, C.POINTER(PwVariant)]),
    "pw_rates_noalloc": (C.c_int, [C.c_int, C.c_double, _dp, C.c_size_t, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pw_noalloc_scratch_bytes": (C.c_size_t, []),
    "pw_diagnose": (C.c_int, [C.c_void_p, C.POINTER(PwDiag)]),
    "pw_resolve": (C.c_int, [C.c_void_p, _dp]),
    "pw_write": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pw_write_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int]),
    "pw_write_wait": (C.c_int, [C.c_void_p]),
    "pw_append_particles": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dp, _dp, _dp, _dp]),
    "pw_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(PwVariant)]),
    "pw_batch_destroy": (C.c_int, [C.c_void_p]),
    "pw_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "pw_batch_add": (C.c_int, [C.c_void_p, C.POINTER(PwJobSpec), C.POINTER(C.c_int64)]),
    "pw_batch_remove": (C.c_int, [C.c_void_p, C.c_int64]),
    "pw_batch_step": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int]),
    "pw_batch_solve": (C.c_int, [C.c_void_p]),
    "pw_batch_get": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp, _dp, _dp]),
    "pw_batch_status": (C.c_int, [C.c_void_p, C.c_int64]),
    "pw_batch_last_timing": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "pw_chunk_for": (C.c_size_t, [C.c_size_t, C.c_int]),
    "pw_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_size_t,
                                   C.POINTER(PwVariant)]),
    "pw_server_destroy": (C.c_int, [C.c_void_p]),
    "pw_server_last_error": (C.c_char_p, [C.c_void_p]),
    "pw_server_now_ms": (C.c_double, [C.c_void_p]),
    "pw_server_submit": (C.c_int, [C.c_void_p, C.POINTER(PwJobSpec), C.c_size_t, C.c_double,
                                   C.POINTER(C.c_int64)]),
    "pw_server_cancel": (C.c_int, [C.c_void_p, C.c_int64]),
    "pw_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "pw_server_job_state": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "pw_server_fetch": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp]),
    "pw_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(PwServerStats)]),
    "pw_job_footprint": (C.c_size_t, [C.c_size_t, C.c_int]),
    "pw_autotune": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_size_t, _dp, _dp, _dp, _dp, C.c_int,
                              C.POINTER(PwTuneEntry), C.c_size_t, C.POINTER(C.c_size_t),
                              C.POINTER(PwTuneEntry)]),
    "pw_tune_save": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_size_t, C.POINTER(PwTuneEntry)]),
    "pw_tune_lookup": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_size_t, C.POINTER(PwVariant)]),
}

_lib = None


class PawanError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """double* to the data of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)
