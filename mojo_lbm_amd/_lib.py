"""ctypes binding of liblbm_hip.so (the C ABI declared in include/lbm_hip.h).

The library is the only compute path: every call here goes to the HIP kernels,
and a missing or unloadable library raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblbm_hip.so"

# enums (include/lbm_hip.h)
DOUBLE_BUFFER, ESOTERIC_PULL = 0, 1
COLL_SRT, COLL_TRT, COLL_KBC, COLL_RLBM = 0, 1, 2, 3
STORE_F32, STORE_F16C, STORE_F64 = 0, 1, 2
LES_NONE, LES_SMAGORINSKY, LES_MLP = 0, 1, 2
U_XYZD, U_DXYZ = 0, 1
OPT_XCD_REMAP, OPT_STATIC_FLAGS, OPT_ADDR64, OPT_STEP_LANES, OPT_STAGED_PULL = 1, 2, 3, 4, 5
BATCH_OPT_GRID_BLOCKS, BATCH_OPT_GRAPH_STEPS = 2, 3
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_NO_DEVICE = 0, -1, -2, -3, -4


class LbmDesc(C.Structure):
    """lbm_desc of include/lbm_hip.h."""

    _fields_ = [
        ("D", C.c_int32), ("Q", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("tx", C.c_int32), ("ty", C.c_int32), ("tz", C.c_int32),
        ("method", C.c_int32), ("collision", C.c_int32), ("storage", C.c_int32),
        ("ddf_shift", C.c_int32), ("les", C.c_int32), ("include_moving_boundary", C.c_int32),
        ("included_bcs", C.c_uint32),
        ("smagorinsky_cs", C.c_double),
    ]


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"liblbm_hip error {code}: {msg}")
        self.code = code


# Exported symbols and their signatures, in header order.
_vp, _i, _d, _sz, _cp = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_char_p
SIGNATURES = {
    "lbm_version": ([], _i),
    "lbm_last_error": ([], _cp),
    "lbm_device_count": ([C.POINTER(_i)], _i),
    "lbm_create": ([C.POINTER(LbmDesc), C.POINTER(_vp)], _i),
    "lbm_destroy": ([_vp], _i),
    "lbm_f_bytes": ([C.POINTER(LbmDesc)], _sz),
    "lbm_bc_bytes": ([C.POINTER(LbmDesc)], _sz),
    "lbm_flags_bytes": ([C.POINTER(LbmDesc)], _sz),
    "lbm_step_db": ([_vp, _vp, _vp, _vp, _vp, _d, _vp], _i),
    "lbm_step_ep": ([_vp, _vp, _vp, _vp, _d, _i, _vp], _i),
    "lbm_macroscopic": ([_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp], _i),
    "lbm_q_criterion": ([_vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _vp], _i),
    "lbm_reduce_moments_f64": ([_vp, _vp, _vp, _vp, _vp, _i, _vp], _i),
    "lbm_init_equilibrium": ([_vp, _vp, _d, C.POINTER(_d), _vp], _i),
    "lbm_set_exterior_wall": ([_vp, _vp, _vp, _cp, _i, C.POINTER(_d), _d, _vp], _i),
    "lbm_set_links": ([_vp, _vp, _vp, C.c_uint32], _i),
    "lbm_set_links_clamped": ([_vp, _vp, _vp, C.c_uint32, C.c_float], _i),
    "lbm_set_link_node_force": ([_vp, _vp], _i),
    "lbm_set_mlp": ([_vp, _vp, C.c_int32], _i),
    "lbm_mlp_forward": ([_vp, _vp, C.c_int64, _vp, _vp], _i),
    "lbm_mlp_closure": ([_vp, _vp, C.c_int64, _vp, _vp], _i),
    "lbm_mlp_workspace_bytes": ([C.c_int64], _sz),
    "lbm_mlp_loss_grad": ([_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp], _i),
    "lbm_mlp_loss_grad_noalloc": ([_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp], _i),
    "lbm_mlp_adam": ([_vp, _vp, _vp, _vp, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, _vp], _i),
    "lbm_box_filter": ([_vp, _vp, _i, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                        C.c_int32, C.c_int32, _vp, _vp], _i),
    "lbm_les_samples": ([_vp, C.c_int32, C.c_int32, C.c_int32, _d, _d, _vp, C.c_int64, _vp, _vp, _vp], _i),
    "lbm_init_fields": ([_vp, _vp, _vp, _vp, _vp], _i),
    "lbm_init_f_from_fields": ([_vp, _vp, _d, _vp, _vp, _d, C.c_int32, C.c_int32, _vp], _i),
    "lbm_link_force": ([_vp, _vp, _vp], _i),
    "lbm_f16c_convert": ([_vp, _vp, _sz, _i, _vp], _i),
    "lbm_set_option": ([_vp, _i, _i], _i),
    "lbm_alloc": ([_sz, C.POINTER(_vp)], _i),
    "lbm_free": ([_vp], _i),
    "lbm_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i),
    "lbm_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i),
    "lbm_memset": ([_vp, _i, _sz, _vp], _i),
    "lbm_synchronize": ([_vp], _i),
    "lbm_batch_create": ([C.POINTER(LbmDesc), C.c_int32, _vp, _vp, C.POINTER(_vp)], _i),
    "lbm_batch_destroy": ([_vp], _i),
    "lbm_batch_set_case": ([_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_uint32], _i),
    "lbm_batch_set_tile": ([_vp, C.c_int32, C.c_int32, C.c_int32], _i),
    "lbm_batch_set_option": ([_vp, _i, _i], _i),
    "lbm_batch_init_equilibrium": ([_vp, _vp], _i),
    "lbm_batch_run": ([_vp, C.c_int64, _vp], _i),
    "lbm_batch_force": ([_vp, _vp, _vp], _i),
    "lbm_batch_case_desc": ([_vp, C.c_int32, C.POINTER(LbmDesc)], _i),
    "lbm_batch_export_case": ([_vp, C.c_int32, _vp, _vp, _vp, _vp, C.POINTER(C.c_uint32)], _i),
    "lbm_batch_get_f": ([_vp, C.c_int32, _vp, _vp], _i),
    "lbm_batch_set_f": ([_vp, C.c_int32, _vp, _vp], _i),
    "lbm_batch_autotune": ([_vp, C.c_int32, _vp, _vp], _i),
    "lbm_batch_info": ([_vp, _vp], _i),
}

_lib = None


def lib() -> C.CDLL:
    """Loads liblbm_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m mojo_lbm_amd.build` "
                "(there is no fallback path)"
            )
        lib_ = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib_, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib_
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise LbmError(rc, lib().lbm_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().lbm_device_count(C.byref(n)))
    return n.value


def dvec(values) -> C.Array | None:
    """Host double array for u arguments (None stays NULL)."""
    if values is None:
        return None
    arr = (C.c_double * len(values))(*[float(v) for v in values])
    return arr


class Handle:
    """Owns one lbm_handle (lbm_create / lbm_destroy)."""

    def __init__(self, desc: LbmDesc):
        self.desc = desc
        h = C.c_void_p()
        check(lib().lbm_create(C.byref(desc), C.byref(h)))
        self.ptr = h

    def close(self) -> None:
        if getattr(self, "ptr", None):
            lib().lbm_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
