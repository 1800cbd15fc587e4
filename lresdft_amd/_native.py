"""ctypes binding of liblres_hip.so (the C-ABI declared in include/lres.h).

The library is built in-tree by ``python -m lresdft_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for the window step."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblres_hip.so")

LRES_FP64 = 0
LRES_MIXED = 1
LRES_MIX_PARTS = 1024   # workspace doubles of lres_diis_mix (include/lres.h)
LRES_MAX_HIST = 8       # history rows of lres_scf_residual

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); every entry point of include/lres.h
SIGNATURES = {
    "lres_last_error": (C.c_char_p, []),
    "lres_device_count": (C.c_int, []),
    "lres_build_windows": (C.c_int, [_f64p, C.c_int64, _f64p, C.c_int64, _i64p, C.c_int64, _f64p, _i64p,
                                     C.c_double, C.c_double, _i64p, _i64p, _i64p, _i64p]),
    "lres_build_windows_device": (C.c_int, [_f64p, C.c_int64, _f64p, C.c_int64, _i64p, C.c_int64, _f64p, _i64p,
                                            C.c_double, C.c_double, _i64p, _i64p, _i64p, _i64p]),
    "lres_plan_create": (C.c_int, [C.POINTER(_vp), C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int,
                                   _i64p, _i64p, C.c_int64, _i64p, _i64p, _i64p, _i64p, _f64p,
                                   _i64p, _i64p, _f64p, _vp]),
    "lres_plan_destroy": (C.c_int, [_vp]),
    "lres_plan_sizes": (C.c_int, [_vp, _i32p, _i32p]),
    "lres_plan_buckets": (C.c_int, [_vp, _i32p, C.c_int, C.POINTER(C.c_int)]),
    "lres_plan_set_precision": (C.c_int, [_vp, C.c_int]),
    "lres_set_hamiltonian_csr": (C.c_int, [_vp, _i64p, _i64p, _f64p, C.c_int64]),
    "lres_set_hamiltonian_csr_device": (C.c_int, [_vp, _vp, _vp, _vp]),
    "lres_plan_set_basis": (C.c_int, [_vp, _i64p, C.c_int64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p,
                                      _i64p, _i64p]),
    "lres_set_potential": (C.c_int, [_vp, _vp]),
    "lres_plan_set_stream": (C.c_int, [_vp, _vp]),
    "lres_diis_mix": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                _vp, _vp, _vp]),
    "lres_lda_potential": (C.c_int, [_vp, _vp, _vp, C.c_int64, _vp, _vp]),
    "lres_scf_residual": (C.c_int, [_vp, C.c_double, _vp, C.c_int64, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
    "lres_scf_residual_noalloc": (C.c_int, [_vp, C.c_double, _vp, C.c_int64, C.c_int, C.c_int, _vp, _vp, _vp,
                                            _vp]),
    "lres_harris_terms": (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp, _vp]),
    "lres_window_pass": (C.c_int, [_vp, C.c_double, C.c_double, _vp, _vp, _vp, _f64p]),
    "lres_density_last": (C.c_int, [_vp, _vp]),
    "lres_inertia": (C.c_int, [_vp, _f64p, C.c_int, _i64p, _vp]),
    "lres_lambda_min": (C.c_int, [_vp, C.c_double, _f64p]),
    "lres_mu_solve": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, _vp, _f64p, _f64p]),
    "lres_mu_solve_warm": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _f64p, _f64p]),
    "lres_step_enqueue": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_int, C.c_int, _vp]),
    "lres_step_finish": (C.c_int, [_vp, _vp, _f64p]),
    "lres_step_params": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int]),
    "lres_step_launch": (C.c_int, [_vp, C.c_int, _vp]),
    "lres_step_arm": (C.c_int, [_vp, C.c_int]),
    "lres_windows": (C.c_int, [_i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p,
                               _i64p, _i64p, C.c_int64, _i64p, _i64p, C.c_int64, C.c_int64, C.c_double,
                               C.c_double, C.c_double, C.c_int, _f64p, _f64p, _f64p]),
    "lres_last_timing": (C.c_int, [_vp, _f64p]),
    "lres_plan_stats": (C.c_int, [_vp, _f64p, C.c_int]),
    "lres_plan_edge_stats": (C.c_int, [_vp, _f64p]),
    "lres_plan_layout": (C.c_int, [_vp, _i64p, _i64p]),
    "lres_set_window_blocks": (C.c_int, [_vp, _vp, _vp]),
    "lres_get_window_blocks": (C.c_int, [_vp, _vp, _vp]),
    "lres_adjoint": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, _vp, _vp]),
    "lres_pass_blocks": (C.c_int, [_vp, _vp, _vp]),
    "lres_force_setup": (C.c_int, [_vp, _f64p, C.c_int64, _i64p, C.c_int64, C.c_double, _f64p, C.c_int, C.c_int,
                                   _f64p, _i64p, C.c_double]),
    "lres_band_forces": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp]),
}

_lib = None


def lib():
    """Load liblres_hip.so once; raise if it is absent (no fallback path exists).

    torch is imported first: its bundled HIP runtime carries the same soname
    (libamdhip64.so.7) as the one this library was linked against, so the
    dynamic linker then binds both to that single runtime.  Two HIP/HSA
    runtimes in one process cannot share the device."""
    global _lib
    if _lib is None:
        import torch  # noqa: F401  (see docstring)
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `python -m lresdft_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def loaded_runtime():
    """Path of the HIP runtime mapped into this process (for diagnostics)."""
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                return line.split()[-1]
    return None


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().lres_last_error().decode() or f"liblres_hip error {rc}")


def p64(a):
    return a.ctypes.data_as(_i64p)


def pf(a):
    return a.ctypes.data_as(_f64p)


def pi32(a):
    return a.ctypes.data_as(_i32p)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
