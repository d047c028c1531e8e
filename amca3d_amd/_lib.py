"""ctypes binding of libamca3d.so (include/amca3d.h).

The library is built in-tree (amca3d_amd/_build/) by ``amca3d_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(PKG_DIR, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libamca3d.so")
CLI_PATH = os.path.join(BUILD_DIR, "amca3d")


class DomainDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("h", C.c_double), ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("T0", C.c_double), ("Tm", C.c_double), ("Tdot", C.c_double),
        ("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double),
        ("trunc", C.c_double), ("time_step_factor", C.c_double),
        ("max_dt", C.c_double), ("final_time", C.c_double),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("time", C.c_double), ("dt", C.c_double),
        ("n_liquid", C.c_int64), ("n_grains", C.c_int32), ("n_orient", C.c_int32),
        ("n_active", C.c_int32), ("n_sites", C.c_int32), ("done", C.c_int32),
        ("max_wave", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepRecord(C.Structure):
    _fields_ = [
        ("time_before", C.c_double), ("dt", C.c_double),
        ("n_nucleated", C.c_int32), ("n_active_before", C.c_int32),
        ("n_captured", C.c_int32), ("n_active_after", C.c_int32),
        ("n_liquid", C.c_int64), ("n_wave1", C.c_int32), ("max_wave", C.c_int32),
        ("n_sites", C.c_int32),
    ]


class SimInfo(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("h", C.c_double),
        ("termination_time", C.c_double), ("time_step", C.c_double),
        ("n_sites", C.c_int32), ("n_surface_sites", C.c_int32), ("n_bulk_sites", C.c_int32),
        ("n_site_draws", C.c_int64), ("coupled", C.c_int32), ("n_frames", C.c_int32),
        ("fe_nodes", C.c_int32), ("n_init_grains", C.c_int32),
    ]


I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
F64P = C.POINTER(C.c_double)
VPP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every entry must be exported by the library
SIGNATURES = {
    "amca_domain_create": (C.c_int, [C.POINTER(DomainDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    "amca_domain_destroy": (C.c_int, [C.c_void_p]),
    "amca_domain_set_sites": (C.c_int, [C.c_void_p, C.c_int32, I32P, F64P]),
    "amca_domain_push_orientations": (C.c_int, [C.c_void_p, C.c_int32, F64P, F64P]),
    "amca_domain_step": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "amca_domain_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Counters)]),
    "amca_domain_run_fixed": (C.c_int, [C.c_void_p, C.c_int64]),
    "amca_domain_counters": (C.c_int, [C.c_void_p, C.POINTER(Counters)]),
    "amca_domain_enable_log": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_domain_read_log": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(StepRecord), I32P]),
    "amca_domain_download": (C.c_int, [C.c_void_p, I32P, I32P, F64P, C.c_int32, I32P, F64P, I32P]),
    "amca_domain_download_grains": (C.c_int, [C.c_void_p, C.c_int32, F64P, I32P, I32P, I32P]),
    "amca_domain_device_fields": (C.c_int, [C.c_void_p, VPP, VPP, VPP, VPP]),
    "amca_domain_enable_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_kernel_count": (C.c_int32, []),
    "amca_kernel_name": (C.c_char_p, [C.c_int32]),
    "amca_domain_kernel_times": (C.c_int, [C.c_void_p, F64P, I64P]),
    "amca_envelope_selftest": (C.c_int, [C.c_int32, F64P, F64P, F64P, C.c_double, C.c_double,
                                         I32P, F64P]),
    "amca_last_error": (C.c_char_p, []),
    "amca_version": (C.c_char_p, []),
    "amca_copy_probe": (C.c_int, [C.c_int64, C.c_int32, F64P, F64P]),
    "amca_sim_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "amca_sim_create_from_file": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "amca_sim_prepare": (C.c_int, [C.c_void_p]),
    "amca_sim_initialize": (C.c_int, [C.c_void_p]),
    "amca_sim_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Counters)]),
    "amca_sim_domain": (C.c_void_p, [C.c_void_p]),
    "amca_sim_remelt": (C.c_void_p, [C.c_void_p]),
    "amca_sim_finalize": (C.c_int, [C.c_void_p]),
    "amca_sim_enable_vtr": (C.c_int, [C.c_void_p, C.c_int32]),
    "amca_write_final_grains": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32,
                                          C.c_int32, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_int32]),
    "amca_sim_get_info": (C.c_int, [C.c_void_p, C.POINTER(SimInfo)]),
    "amca_sim_initial_sites": (C.c_int, [C.c_void_p, C.c_int32, I32P, F64P, F64P, I32P]),
    "amca_sim_destroy": (C.c_int, [C.c_void_p]),
    "amca_sim_last_error": (C.c_char_p, []),
}

_lib = None


class AmcaError(RuntimeError):
    pass


def load():
    """Load libamca3d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AmcaError(f"{LIB_PATH} is missing: build it with amca3d_amd.build.build()")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, where="", sim=False):
    if status != 0:
        lib = load()
        msg = (lib.amca_sim_last_error() if sim else lib.amca_last_error()).decode()
        raise AmcaError(f"{where}: status {status}: {msg}")
