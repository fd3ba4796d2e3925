"""ctypes binding of libmusica_gpu.so (the C-ABI declared in include/musica_gpu.h).

The library is the product: every call here goes to the HIP kernels.  There is
no CPU fallback -- if the shared library is missing or fails to load, the
import of :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmusica_gpu.so")

MGPU_NGH = 5
FACE_PERIODIC = 0
FACE_INTERFACE = 1
FACE_WALL = 2
FACE_CHARAC = 3
FACE_INLET = 4
FACE_SLIP = 5
FILTER_NONE, FILTER_SF, FILTER_SHOCK, FILTER_ARTVISC, FILTER_ARTVISC_SHOCK = 0, 1, 2, 3, 4
PREC_FP64, PREC_MIXED = 0, 1
GEOM_CART, GEOM_C3, GEOM_C2 = 0, 1, 2
LAUNCH_EAGER = 1
LAUNCH_STAGE_CLASSIC, LAUNCH_STAGE_PIPELINED = 2, 4
OK, EINVAL, EDEVICE, ENOMEM, EUNSUP, ESTATE = 0, -1, -2, -3, -4, -5
K_STAGE, K_VISC, K_FILTER, K_HALO = 0, 1, 2, 3
K_NAMES = {K_STAGE: "stage", K_VISC: "visc", K_FILTER: "filter", K_HALO: "halo"}

# every symbol include/musica_gpu.h declares
EXPORTS = (
    "mgpu_create",
    "mgpu_set_metrics_cart",
    "mgpu_set_metrics_c3",
    "mgpu_set_metrics_c2",
    "mgpu_set_wall",
    "mgpu_set_wall_c3",
    "mgpu_set_wall_c3_slip",
    "mgpu_put_state",
    "mgpu_get_state",
    "mgpu_step",
    "mgpu_autotune",
    "mgpu_stage_config",
    "mgpu_stage_pipelined",
    "mgpu_sync",
    "mgpu_profile",
    "mgpu_kernel_time",
    "mgpu_last_error",
    "mgpu_diagnostics",
    "mgpu_diagnostics_noalloc",
    "mgpu_max_cfl",
    "mgpu_destroy",
    "mgpu_exchange_plan",
    "mgpu_server_create",
    "mgpu_server_submit",
    "mgpu_server_tick",
    "mgpu_server_status",
    "mgpu_server_take",
    "mgpu_server_memory",
    "mgpu_server_batching",
    "mgpu_server_last_error",
    "mgpu_server_destroy",
)


class MgpuBlock(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("face_kind", C.c_int32 * 6),
        ("nbr_block", C.c_int32 * 6),
        ("nbr_face", C.c_int32 * 6),
        ("orient", C.c_int32 * 6),
    ]


class MgpuCase(C.Structure):
    _fields_ = [
        ("nblocks", C.c_int32),
        ("curvilinear", C.c_int32),
        ("iorder_visc", C.c_int32),
        ("filter", C.c_int32),
        ("is_ducros", C.c_int32),
        ("nrk", C.c_int32),
        ("precision", C.c_int32),
        ("launch_flags", C.c_int32),
        ("blocks", C.POINTER(MgpuBlock)),
        ("crk", C.c_double * 8),
        ("deltat", C.c_double),
        ("gam", C.c_double), ("rg", C.c_double),
        ("c0", C.c_double), ("S0", C.c_double),
        ("diffscale", C.c_double),
        ("cok_ratio", C.c_double),
        ("dissip_coeff", C.c_double),
        ("dissip_shock", C.c_double),
        ("csens", C.c_double),
        ("is_adiab", C.c_int32),
        ("is_sw_edoh", C.c_int32),
        ("t_wall", C.c_double),
        ("p_ref", C.c_double),
        ("is_wall2", C.c_int32),
        ("reserved0", C.c_int32),
    ]


class MgpuDiag(C.Structure):
    _fields_ = [("mass", C.c_double), ("kinetic", C.c_double), ("enstrophy", C.c_double), ("umax", C.c_double),
                ("cells", C.c_int64)]


class MgpuError(RuntimeError):
    pass


_lib = None

_dp = C.POINTER(C.c_double)


def lib() -> C.CDLL:
    """Load libmusica_gpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MgpuError(
            f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    L = C.CDLL(LIB_PATH)
    L.mgpu_create.argtypes = [C.POINTER(MgpuCase), C.POINTER(C.c_void_p)]
    L.mgpu_set_metrics_cart.argtypes = [C.c_void_p, C.c_int] + [_dp] * 6
    L.mgpu_set_metrics_c3.argtypes = [C.c_void_p, C.c_int, C.POINTER(_dp), C.POINTER(_dp)]
    L.mgpu_set_metrics_c2.argtypes = [C.c_void_p, C.c_int, C.POINTER(_dp), C.POINTER(_dp), _dp, _dp]
    L.mgpu_set_wall.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.mgpu_set_wall_c3.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(_dp)]
    L.mgpu_set_wall_c3_slip.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(_dp)]
    L.mgpu_put_state.argtypes = [C.c_void_p, C.c_int] + [_dp] * 5
    L.mgpu_get_state.argtypes = [C.c_void_p, C.c_int] + [_dp] * 5
    L.mgpu_step.argtypes = [C.c_void_p, C.c_int]
    L.mgpu_autotune.argtypes = [C.c_void_p, C.c_int]
    L.mgpu_stage_config.argtypes = [C.c_void_p, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 3
    L.mgpu_stage_pipelined.argtypes = [C.c_void_p, C.c_int]
    L.mgpu_stage_pipelined.restype = C.c_int
    L.mgpu_sync.argtypes = [C.c_void_p]
    L.mgpu_profile.argtypes = [C.c_void_p, C.c_int]
    L.mgpu_kernel_time.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.mgpu_last_error.argtypes = [C.c_void_p]
    L.mgpu_last_error.restype = C.c_char_p
    L.mgpu_destroy.argtypes = [C.c_void_p]
    L.mgpu_destroy.restype = None
    L.mgpu_diagnostics.argtypes = [C.c_void_p, C.POINTER(MgpuDiag)]
    L.mgpu_diagnostics.restype = C.c_int
    L.mgpu_diagnostics_noalloc.argtypes = [C.c_void_p, C.POINTER(MgpuDiag)]
    L.mgpu_diagnostics_noalloc.restype = C.c_int
    L.mgpu_max_cfl.argtypes = [C.c_void_p, _dp]
    L.mgpu_max_cfl.restype = C.c_int
    L.mgpu_exchange_plan.argtypes = [C.POINTER(MgpuCase), C.c_int, C.POINTER(C.c_int64), C.c_int64,
                                     C.POINTER(C.c_int64)]
    L.mgpu_exchange_plan.restype = C.c_int
    L.mgpu_server_create.argtypes = [C.c_int64, C.POINTER(C.c_void_p)]
    L.mgpu_server_submit.argtypes = [C.c_void_p, C.POINTER(MgpuCase), C.POINTER(_dp), C.POINTER(_dp), C.c_int,
                                     C.POINTER(C.c_int64)]
    L.mgpu_server_tick.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mgpu_server_status.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mgpu_server_take.argtypes = [C.c_void_p, C.c_int64] + [_dp] * 5
    L.mgpu_server_memory.argtypes = [C.c_void_p]
    L.mgpu_server_memory.restype = C.c_int64
    L.mgpu_server_batching.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.mgpu_server_batching.restype = C.c_int
    L.mgpu_server_last_error.argtypes = [C.c_void_p]
    L.mgpu_server_last_error.restype = C.c_char_p
    L.mgpu_server_destroy.argtypes = [C.c_void_p]
    L.mgpu_server_destroy.restype = None
    for name in ("mgpu_server_create", "mgpu_server_submit", "mgpu_server_tick", "mgpu_server_status",
                 "mgpu_server_take"):
        getattr(L, name).restype = C.c_int
    for name in ("mgpu_create", "mgpu_set_metrics_cart", "mgpu_set_metrics_c3", "mgpu_set_metrics_c2", "mgpu_set_wall",
                 "mgpu_set_wall_c3", "mgpu_set_wall_c3_slip",
                 "mgpu_put_state",
                 "mgpu_get_state", "mgpu_step", "mgpu_autotune", "mgpu_stage_config", "mgpu_sync", "mgpu_profile", "mgpu_kernel_time"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int, handle=None, what: str = "") -> None:
    if rc != 0:
        msg = lib().mgpu_last_error(handle).decode(errors="replace")
        raise MgpuError(f"{what} failed (rc={rc}): {msg}")


def dptr(a):
    """double* of a C-contiguous float64 numpy array."""
    assert a.dtype.name == "float64" and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)
