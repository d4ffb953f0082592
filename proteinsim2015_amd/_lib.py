"""ctypes binding of the C ABI declared in include/psim_gpu.h.

The shared library is built in-tree (``python -m proteinsim2015_amd.build``)
and loaded from this package directory.  There is no fallback: if the
library is missing or was built for another ABI version, loading fails.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpsim_gpu.so")
ABI_VERSION = 1

FORCE_CAL, FORCE_LJ, FORCE_AO = 1, 2, 4
ERR_INIT, ERR_BOUNDS, ERR_OVERLAP, ERR_NAN = 7701, 7703, 7704, 7705
E_ARG, E_HIP, E_NOMEM, E_UNSUPPORTED, E_IO = -1, -2, -3, -4, -5


class PsimParams(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int), ("cells_per_dim", C.c_int),
        ("box", C.c_double), ("cell_size", C.c_double), ("half_uv", C.c_double),
        ("radius", C.c_double), ("mass", C.c_double), ("kT", C.c_double),
        ("gamma", C.c_double), ("dt", C.c_double),
        ("vel_freq", C.c_int), ("force_mask", C.c_int), ("lj_num", C.c_int),
        ("kernel_path", C.c_int),
        ("cutoff", C.c_double), ("lj_well_depth", C.c_double),
        ("yukawa_strength", C.c_double), ("debye_length", C.c_double),
        ("ao_well_depth", C.c_double), ("seed", C.c_ulonglong),
    ]


class PsimReplica(C.Structure):
    _fields_ = [("n_particles", C.c_int), ("cells_per_dim", C.c_int),
                ("box", C.c_double), ("cell_size", C.c_double), ("cutoff", C.c_double),
                ("lj_well_depth", C.c_double), ("yukawa_strength", C.c_double),
                ("debye_length", C.c_double), ("ao_well_depth", C.c_double),
                ("lj_num", C.c_int), ("reserved", C.c_int)]


class Observables(C.Structure):
    _fields_ = [("temperature", C.c_double), ("mean_disp2", C.c_double),
                ("msd", C.c_double), ("mean_coordination", C.c_double),
                ("contacts", C.c_longlong), ("time", C.c_double), ("step", C.c_ulonglong)]


DP_AO_WELL, DP_AO_CUTOFF, DP_LJ_WELL, DP_YUKAWA, DP_DEBYE, DP_COUNT = range(6)
ENS_CLASSES = 4   # PSIM_ENS_CLASSES: occupancy classes of an ensemble's replicas


class KernelTimes(C.Structure):
    _fields_ = [("force_ms", C.c_double), ("force_launches", C.c_int),
                ("integrate_ms", C.c_double), ("integrate_launches", C.c_int),
                ("build_ms", C.c_double), ("build_launches", C.c_int),
                ("total_ms", C.c_double), ("steps", C.c_int)]


# every exported entry point of include/psim_gpu.h
EXPORTS = (
    "psim_gpu_abi_version", "psim_gpu_status_string", "psim_gpu_create",
    "psim_gpu_destroy", "psim_gpu_set_state", "psim_gpu_get_state", "psim_gpu_step",
    "psim_gpu_step_device_noise", "psim_gpu_forces", "psim_gpu_force_table",
    "psim_gpu_get_cells", "psim_gpu_last_error", "psim_gpu_clear_error",
    "psim_gpu_sync", "psim_gpu_set_profiling", "psim_gpu_kernel_times",
    "psim_gpu_stream", "psim_init_rsa",
    "psim_ens_create", "psim_ens_destroy", "psim_ens_total_particles", "psim_ens_set_state",
    "psim_ens_get_state", "psim_ens_set_graph", "psim_ens_step", "psim_ens_sync",
    "psim_ens_last_error", "psim_ens_autotune",
    "psim_gpu_set_force_params", "psim_gpu_track_start", "psim_gpu_observables",
    "psim_gpu_clusters", "psim_gpu_pair_histogram", "psim_gpu_pair_histogram_grad",
    "psim_adj_create", "psim_adj_destroy", "psim_adj_set_initial", "psim_adj_forward",
    "psim_adj_backward", "psim_init_reference",
    "psim_gpu_get_params", "psim_gpu_get_tracker", "psim_gpu_set_tracker",
    "psim_write_recovery", "psim_read_recovery", "psim_write_xyz", "psim_gpu_save",
    "psim_gpu_load", "psim_write_sys_config", "psim_gpu_kernel_info",
    "psim_gpu_pair_stats", "psim_ens_pair_stats", "psim_gpu_forces_noalloc",
    "psim_int_create", "psim_int_destroy", "psim_int_next", "psim_gpu_set_graph",
    "psim_adj_set_cutoff_window", "psim_adj_set_record_all", "psim_ens_class_info",
    "psim_ens_merged",
)

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ulonglong)
PAIRS_STEP, PAIRS_ADJOINT = 0, 1
NOISE_REFERENCE, NOISE_PHILOX = 0, 1

_SIGS = {
    "psim_gpu_abi_version": ([], C.c_int),
    "psim_gpu_status_string": ([C.c_int], C.c_char_p),
    "psim_gpu_create": ([C.POINTER(PsimParams), C.POINTER(_vp)], C.c_int),
    "psim_gpu_destroy": ([_vp], None),
    "psim_gpu_set_state": ([_vp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                            C.c_ulonglong], C.c_int),
    "psim_gpu_get_state": ([_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                            C.POINTER(C.c_ulonglong)], C.c_int),
    "psim_gpu_step": ([_vp, C.c_int, _dp], C.c_int),
    "psim_gpu_step_device_noise": ([_vp, C.c_int, _vp], C.c_int),
    "psim_gpu_forces": ([_vp, _dp], C.c_int),
    "psim_gpu_force_table": ([_vp, _dp, _ip, _ip, _dp], C.c_int),
    "psim_gpu_get_cells": ([_vp, _ip, _ip, _ip], C.c_int),
    "psim_gpu_last_error": ([_vp, _ip, _ip, _ip], C.c_int),
    "psim_gpu_clear_error": ([_vp], C.c_int),
    "psim_gpu_sync": ([_vp], C.c_int),
    "psim_gpu_set_profiling": ([_vp, C.c_int], C.c_int),
    "psim_gpu_kernel_times": ([_vp, C.POINTER(KernelTimes), C.c_int], C.c_int),
    "psim_gpu_stream": ([_vp], _vp),
    "psim_init_rsa": ([C.c_int, C.c_double, C.c_double, C.c_ulonglong, _dp], C.c_int),
    "psim_ens_create": ([C.POINTER(PsimParams), C.POINTER(PsimReplica), C.c_int,
                         C.POINTER(_vp)], C.c_int),
    "psim_ens_destroy": ([_vp], None),
    "psim_ens_total_particles": ([_vp], C.c_int),
    "psim_ens_set_state": ([_vp, _dp, _dp, C.c_double], C.c_int),
    "psim_ens_get_state": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
    "psim_ens_set_graph": ([_vp, C.c_int], C.c_int),
    "psim_ens_step": ([_vp, C.c_int], C.c_int),
    "psim_ens_sync": ([_vp], C.c_int),
    "psim_ens_last_error": ([_vp, _ip, _ip, _ip], C.c_int),
    "psim_ens_autotune": ([_vp, C.c_int, _ip, _ip, C.POINTER(C.c_float)], C.c_int),
    "psim_gpu_set_force_params": ([_vp, C.POINTER(PsimParams)], C.c_int),
    "psim_gpu_track_start": ([_vp], C.c_int),
    "psim_gpu_observables": ([_vp, C.c_double, C.POINTER(Observables), _ip], C.c_int),
    "psim_gpu_clusters": ([_vp, C.c_double, C.c_int, _ip, _ip, _ip], C.c_int),
    "psim_gpu_pair_histogram": ([_vp, C.c_double, C.c_double, C.c_int, C.c_double, _dp],
                                C.c_int),
    "psim_gpu_pair_histogram_grad": ([_vp, C.c_double, C.c_double, C.c_int, C.c_double,
                                      _dp, _dp], C.c_int),
    "psim_adj_create": ([_vp, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
    "psim_adj_destroy": ([_vp], None),
    "psim_adj_set_initial": ([_vp], C.c_int),
    "psim_adj_set_cutoff_window": ([_vp, C.c_double], C.c_int),
    "psim_adj_set_record_all": ([_vp, C.c_int], C.c_int),
    "psim_ens_class_info": ([_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int)], C.c_int),
    "psim_ens_merged": ([_vp], C.c_int),
    "psim_adj_forward": ([_vp], C.c_int),
    "psim_adj_backward": ([_vp, _dp, _dp, _dp], C.c_int),
    "psim_init_reference": ([C.c_int, C.c_int, C.c_double, C.c_uint, C.c_int, C.c_double, _dp,
                             _dp, _dp], C.c_int),
    "psim_gpu_get_params": ([_vp, C.POINTER(PsimParams)], C.c_int),
    "psim_gpu_kernel_info": ([_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "psim_gpu_pair_stats": ([_vp, C.c_int, _ip, _up], C.c_int),
    "psim_ens_pair_stats": ([_vp, _ip, _up], C.c_int),
    "psim_gpu_forces_noalloc": ([C.POINTER(PsimParams), _vp, _vp, _vp], C.c_int),
    "psim_int_create": ([C.POINTER(PsimParams), C.c_int, C.POINTER(_vp)], C.c_int),
    "psim_gpu_set_graph": ([_vp, C.c_int], C.c_int),
    "psim_int_destroy": ([_vp], None),
    "psim_int_next": ([_vp, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip], C.c_int),
    "psim_gpu_get_tracker": ([_vp, _dp, _ip], C.c_int),
    "psim_gpu_set_tracker": ([_vp, _dp, _ip], C.c_int),
    "psim_write_recovery": ([C.c_char_p, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double],
                            C.c_int),
    "psim_read_recovery": ([C.c_char_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp],
                           C.c_int),
    "psim_write_xyz": ([C.c_char_p, C.c_int, _dp, C.c_double], C.c_int),
    "psim_gpu_save": ([_vp, C.c_char_p], C.c_int),
    "psim_gpu_load": ([_vp, C.c_char_p], C.c_int),
    "psim_write_sys_config": ([C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_int, C.c_int,
                              C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int],
                             C.c_int),
}

_LIB = None


def load() -> C.CDLL:
    """Load libpsim_gpu.so; raises if it is missing (no CPU fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m proteinsim2015_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    v = lib.psim_gpu_abi_version()
    if v != ABI_VERSION:
        raise RuntimeError(f"libpsim_gpu.so ABI {v} != {ABI_VERSION}")
    _LIB = lib
    return lib


class PsimError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = load().psim_gpu_status_string(status).decode()
        super().__init__(f"{what}: {msg}" if what else msg)


def check(status: int, what: str = "") -> int:
    if status != 0:
        raise PsimError(status, what)
    return status
