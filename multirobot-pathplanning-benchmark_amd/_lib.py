"""ctypes binding of libmrcc.so (the C ABI in include/mrcc.h).

The library is built in-tree by ``build.py`` (``__graft_entry__.build()``).  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmrcc.so")

MRCC_OK = 0
MRCC_ERR_INVALID = -1
MRCC_ERR_HIP = -2
MRCC_ERR_NOMEM = -3
MRCC_ERR_UNSUPPORTED = -4
MRCC_ERR_INJECTED = -5

KIND_DISKS, KIND_PLANAR, KIND_ARMS = 1, 2, 3

_c = ctypes
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int32)
_u8p = _c.POINTER(_c.c_uint8)
_vp = _c.c_void_p
_i64 = _c.c_int64
_i32 = _c.c_int32
_u32 = _c.c_uint32


class MrccError(RuntimeError):
    """A libmrcc call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libmrcc error {code}: {msg}")
        self.code = code


class EdgeBatch(_c.Structure):
    _fields_ = [("q1", _vp), ("q2", _vp), ("row_stride", _i64), ("dim_stride", _i64),
                ("idx1", _vp), ("idx2", _vp), ("n_edges", _i64), ("resolution", _c.c_double),
                ("N", _vp), ("N_start", _vp), ("N_max", _vp), ("include_endpoints", _i32),
                ("mode_ids", _vp), ("robot_mask", _u32), ("out", _vp)]


class Tune(_c.Structure):
    _fields_ = [("slots_per_pass", _i32), ("slots_per_ticket", _i32), ("waves_per_cu", _i32),
                ("sorted", _i32), ("persistent", _i32), ("use_graph", _i32),
                ("collect_stats", _i32), ("time_kernels", _i32)]


# name -> (restype, argtypes); every symbol of include/mrcc.h
SIGNATURES = {
    "mrcc_abi_version": (_c.c_int, []),
    "mrcc_last_error": (_c.c_char_p, []),
    "mrcc_init": (_c.c_int, [_c.c_int]),
    "mrcc_debug_inject_failure": (_c.c_int, [_c.c_int, _c.c_int]),
    "mrcc_scene_create_disks": (_c.c_int, [_c.c_int, _c.c_int, _dp, _c.c_int, _dp, _dp, _c.c_int, _dp, _dp,
                                           _c.POINTER(_vp)]),
    "mrcc_scene_create_planar_boxes": (_c.c_int, [_c.c_int, _dp, _c.c_int, _dp, _dp, _c.POINTER(_vp)]),
    "mrcc_scene_create_arms": (_c.c_int, [_c.c_int, _c.c_int, _dp, _dp, _ip, _dp, _c.c_int, _dp,
                                          _c.POINTER(_vp)]),
    "mrcc_scene_create_arm_chains": (_c.c_int, [_c.c_int, _c.c_int, _dp, _dp, _ip, _dp, _c.c_int, _dp,
                                                _c.POINTER(_vp)]),
    "mrcc_scene_set_modes": (_c.c_int, [_vp, _c.c_int, _ip, _c.c_int, _u8p, _dp, _dp]),
    "mrcc_scene_set_broadphase": (_c.c_int, [_vp, _c.c_int]),
    "mrcc_scene_set_joint_types": (_c.c_int, [_vp, _u8p, _dp]),
    "mrcc_scene_destroy": (_c.c_int, [_vp]),
    "mrcc_scene_info": (_c.c_int, [_vp, _ip]),
    "mrcc_check_configs": (_c.c_int, [_vp, _vp, _i64, _i64, _i64, _vp, _u32, _vp, _vp]),
    "mrcc_edge_plan_create": (_c.c_int, [_vp, _c.POINTER(EdgeBatch), _c.POINTER(Tune), _vp,
                                         _c.POINTER(_vp)]),
    "mrcc_edge_plan_run": (_c.c_int, [_vp, _vp]),
    "mrcc_edge_plan_stats": (_c.c_int, [_vp, _c.POINTER(_i64)]),
    "mrcc_edge_plan_timing": (_c.c_int, [_vp, _dp]),
    "mrcc_edge_plan_destroy": (_c.c_int, [_vp]),
    "mrcc_check_edges": (_c.c_int, [_vp, _c.POINTER(EdgeBatch), _c.POINTER(Tune), _vp]),
    "mrcc_check_edge": (_c.c_int, [_vp, _vp, _vp, _c.c_double, _i32, _i32, _i32, _i32, _i32, _u32, _vp, _vp]),
    # the rows of the per-call entry points as addresses (a numpy array's .ctypes.data is the
    # cheapest pointer to take per call)
    "mrcc_check_edge_host": (_c.c_int, [_vp, _vp, _vp, _c.c_double, _i32, _i32, _i32, _i32, _i32, _u32,
                                        _c.POINTER(_i32)]),
    "mrcc_check_config_host": (_c.c_int, [_vp, _vp, _i32, _u32, _c.POINTER(_i32)]),
    "mrcc_check_edges_host": (_c.c_int, [_vp, _vp, _vp, _i64, _c.c_double, _vp, _vp, _vp, _i32, _vp, _u32,
                                         _vp]),
    "mrcc_check_edges_per_call": (_c.c_int, [_vp, _c.POINTER(EdgeBatch), _vp]),
    "mrcc_check_edges_noalloc": (_c.c_int, [_vp, _c.POINTER(EdgeBatch), _c.POINTER(Tune), _vp, _vp]),
    "mrcc_bfs_indices": (_c.c_int, [_i32, _ip]),
    "mrcc_edge_num_states": (_i64, [_i32, _i32, _i32, _i32]),
    "mrcc_arm_fk": (_c.c_int, [_vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "mrcc_sincos": (_c.c_int, [_vp, _i64, _vp, _vp, _vp]),
    "mrcc_clearance_forward": (_c.c_int, [_vp, _vp, _i64, _i32, _c.c_float, _vp, _vp]),
    "mrcc_clearance_backward": (_c.c_int, [_vp, _vp, _i64, _i32, _c.c_float, _vp, _vp, _vp]),
    "mrcc_config_cost": (_c.c_int, [_vp, _i64, _vp, _i64, _i64, _i32, _vp, _i32, _i32, _i32, _c.c_double,
                                    _vp, _vp]),
}

# metric / reduction codes of mrcc_config_cost
COST_MAX, COST_EUCLIDEAN, COST_ABSMAX = 0, 1, 2
REDUCE_MAX_SUM, REDUCE_SUM, REDUCE_MAX = 0, 1, 2

_lib = None


def lib():
    """Loads libmrcc.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              f"(hipcc for gfx950); there is no CPU fallback")
        # PyTorch ships its own HIP runtime.  Loading it first lets the library bind to that same
        # runtime, so device buffers and streams from torch are valid here; loaded the other way
        # round, torch could no longer initialise its device.
        try:
            import torch

            torch.cuda.is_available()
        except ImportError:
            pass
        L = _c.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != MRCC_OK:
        msg = lib().mrcc_last_error()
        raise MrccError(rc, msg.decode() if msg else "")


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array."""
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)
