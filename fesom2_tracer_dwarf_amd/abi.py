"""ctypes binding of libfesomtra.so (the C ABI declared in include/fesomtra.h).

Argument types follow the reference's own ctypes harness (python/demo_tracer.py:50-99): explicit
``argtypes``/``restype``, Fortran-ordered float64 / int32 arrays, 1-based index values.  Loading
fails loudly if the library has not been built: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

# FESOMTRA_LIB selects another build of the library (e.g. one compiled with different options).
LIB_PATH = Path(os.environ.get("FESOMTRA_LIB", Path(__file__).resolve().parent / "libfesomtra.so"))

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_char_p = ctypes.c_char_p
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# name: (restype, argtypes)
_SIGNATURES = {
    "tracer_init_mpi": (_c_int, [_c_int]),
    "tracer_finalize_mpi": (_c_int, []),
    "tracer_init": (_c_int, [_c_char_p]),
    "tracer_advect_step": (_c_int, [_c_double]),
    "tracer_get_stats": (_c_int, [_pd, _pd, _pd]),
    "tracer_get_size": (_c_int, [_pi, _pi]),
    "tracer_get_values": (_c_int, [_pd, _c_int, _c_int]),
    "tracer_finalize": (_c_int, [_c_char_p]),
    "tracer_run_workflow": (_c_int, [_c_char_p, _c_int, _c_double]),
    "tracer_cleanup": (_c_int, []),
    "tracer_set_mesh_dims": (_c_int, [_c_int, _c_int, _c_int, _c_int]),
    "tracer_set_levels": (_c_int, [_pi, _pi, _c_int]),
    "tracer_set_thickness": (_c_int, [_pd, _c_int, _c_int]),
    "tracer_set_partition": (_c_int, [_c_int, _c_int]),
    "tracer_allocate_tracers": (_c_int, [_c_int, _c_int, _c_int, _c_int]),
    "tracer_set_values": (_c_int, [_c_int, _pd, _c_int, _c_int]),
    "tracer_set_velocity_zero": (_c_int, [_c_int, _c_int, _c_int]),
    "tracer_set_velocity": (_c_int, [_pd, _pd, _c_int, _c_int, _c_int]),
    "tracer_init_complete": (_c_int, []),
    "tracer_load_mesh_partition": (_c_int, [_c_char_p]),
    "tracer_load_mesh_from_files": (_c_int, [_c_char_p]),
    "tracer_init_ale_arrays": (_c_int, []),
    "tracer_init_for_advection": (_c_int, []),
    "tracer_set_coordinates": (_c_int, [_pd, _c_int]),
    "tracer_set_connectivity": (_c_int, [_pi, _pi, _pi, _pi, _pi, _c_int, _c_int, _c_int, _c_int]),
    "tracer_set_elem_levels": (_c_int, [_pi, _pi, _c_int]),
    "tracer_set_geometry": (_c_int, [_pd, _pd, _pd, _pd, _pd, _pd, _pd, _c_int, _c_int, _c_int, _c_int]),
    "tracer_set_vertical_grid": (_c_int, [_pd, _pd, _pd, _pd, _pd, _c_int, _c_int, _c_int]),
    "tracer_set_zbar": (_c_int, [_pd, _c_int]),
    "tracer_set_muscl_support": (_c_int, [_pi, _pi, _c_int, _c_int]),
    "tracer_set_velocity_elem": (_c_int, [_pd, _pd, _pd, _pd, _c_int, _c_int, _c_int]),
    "tracer_set_node_dt": (_c_int, [_pd, _c_int]),
    "tracer_set_ensemble": (_c_int, [_c_int]),
    "tracer_get_ensemble": (_c_int, [_pi]),
    "tracer_set_velocity_member": (_c_int, [_c_int, _pd, _pd, _pd, _pd, _c_int, _c_int, _c_int]),
    "tracer_set_tracer_member": (_c_int, [_c_int, _c_int]),
    "tracer_set_scheme": (_c_int, [_c_int, _c_char_p, _c_char_p, _c_char_p, _c_double, _c_double]),
    "tracer_advect_steps": (_c_int, [_c_int, _c_double]),
    "tracer_do_oce_adv_tra": (_c_int, [_c_double, _c_int]),
    "tracer_zero_tendencies": (_c_int, []),
    "tracer_get_tendencies": (_c_int, [_pd, _pd, _c_int, _c_int]),
    "tracer_set_values_ab": (_c_int, [_c_int, _pd, _c_int, _c_int]),
    "tracer_get_values_by_id": (_c_int, [_c_int, _pd, _c_int, _c_int]),
    "tracer_get_stats_by_id": (_c_int, [_c_int, _pd, _pd, _pd, _pd]),
    "tracer_get_work": (_c_int, [_c_char_p, _pd, ctypes.c_size_t]),
    "tracer_set_keep_work": (_c_int, [_c_int]),
    "tracer_set_fused_limiter": (_c_int, [_c_int]),
    "tracer_set_concurrent": (_c_int, [_c_int]),
    "tracer_get_fused_limiter": (_c_int, [_pi]),
    "tracer_set_values_device": (_c_int, [_c_int, _vp, _c_int, _c_int]),
    "tracer_get_values_device": (_c_int, [_c_int, _vp, _c_int, _c_int]),
    "tracer_last_elapsed_ms": (_c_int, [_pd]),
    "tracer_set_kernel_timing": (_c_int, [_c_int]),
    "tracer_get_kernel_times": (_c_int, [_pd, ctypes.POINTER(ctypes.c_long)]),
    "tracer_get_step_times": (_c_int, [_pd, _pd, ctypes.POINTER(ctypes.c_long)]),
    "tracer_set_values_ab_device": (_c_int, [_c_int, _vp, _c_int, _c_int]),
    "tracer_get_tendencies_device": (_c_int, [_vp, _vp, _c_int, _c_int]),
    "tracer_do_oce_adv_tra_noalloc": (_c_int, [_c_double, _c_int, _vp, _vp, _vp, _vp]),
    "tracer_set_precision": (_c_int, [_c_int]),
    "tracer_set_tracers_per_pass": (_c_int, [_c_int]),
    "tracer_get_tracers_per_pass": (_c_int, [_pi]),
    "tracer_set_graph": (_c_int, [_c_int]),
    "tracer_set_fused": (_c_int, [_c_int]),
    "tracer_set_wsplit": (_c_int, [_c_int]),
    "tracer_get_wsplit": (_c_int, [_pi]),
    "tracer_set_ab_mode": (_c_int, [_c_int, ctypes.c_double]),
    "tracer_get_ab_mode": (_c_int, [_pi, _pd]),
    "tracer_set_halo_tiles": (_c_int, [_c_int]),
    "tracer_get_halo_tiles": (_c_int, [_pi]),
    "tracer_set_rounding": (_c_int, [_c_int]),
    "tracer_set_compensated": (_c_int, [_c_int]),
    "tracer_checkpoint": (_c_int, [_c_int]),
    "tracer_set_launch_config": (_c_int, [_c_int, _c_int, _c_int, _c_int]),
    "tracer_get_launch_config": (_c_int, [_pi]),
    "tracer_last_error": (_c_char_p, []),
    "tracer_version": (_c_char_p, []),
}

_lib = None


class TracerError(RuntimeError):
    pass


def load(path: Path = LIB_PATH) -> ctypes.CDLL:
    """Load the library (RTLD_GLOBAL, as python/demo_tracer.py:50 does)."""
    global _lib
    if _lib is None:
        if not Path(path).exists():
            raise TracerError(f"{path} is missing: build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'`")
        lib = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.c_contiguous or a.size == 0)
    return a.ctypes.data_as(_pd)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and (a.flags.c_contiguous or a.size == 0)
    return a.ctypes.data_as(_pi)


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().tracer_last_error()
        raise TracerError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
