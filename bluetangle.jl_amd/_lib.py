"""ctypes binding of libbt.so (the C ABI in include/bt.h).

There is no CPU fallback: if the HIP library is missing or no device is
visible, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbt.so")

BT_SV, BT_DM = 0, 1
BT_C64, BT_C128 = 0, 1
BT_BASIS_Z, BT_BASIS_X, BT_BASIS_Y, BT_BASIS_R = 0, 1, 2, 3
BT_OP_GATE, BT_OP_CHANNEL, BT_OP_MEASURE, BT_OP_RESET, BT_OP_IF = 0, 1, 2, 3, 4

ERRORS = {-1: "BT_ERR_ARG", -2: "BT_ERR_UNSUPPORTED", -3: "BT_ERR_HIP", -4: "BT_ERR_NOMEM", -5: "BT_ERR_DRAWS",
          -6: "BT_ERR_STATE", -7: "BT_ERR_BUDGET", -8: "BT_ERR_NODEVICE"}

# exported entry points of include/bt.h
SYMBOLS = [
    "bt_last_error", "bt_version", "bt_init", "bt_uniform",
    "bt_state_create", "bt_state_destroy", "bt_state_info", "bt_state_set_basis", "bt_state_copy",
    "bt_state_upload", "bt_state_download", "bt_state_device_ptr", "bt_state_sync", "bt_state_bytes",
    "bt_apply_full", "bt_apply_fused5", "bt_run_fused", "bt_autotune_fused", "bt_inner", "bt_expect", "bt_sample", "bt_sample_noalloc", "bt_run_batch_sample", "bt_graph_create_fused", "bt_graph_create_dm", "bt_graph_launch",
    "bt_graph_passes", "bt_graph_destroy",
    "bt_apply_matrix", "bt_reduced_rho", "bt_apply_kraus_traj", "bt_apply_kraus_dm", "bt_measure", "bt_reset",
    "bt_probabilities", "bt_norm2", "bt_trace_purity", "bt_gather",
    "bt_circuit_create", "bt_circuit_destroy", "bt_circuit_max_draws", "bt_run_circuit", "bt_run_circuit_dm",
    "bt_run_batch", "bt_run_batch_persistent", "bt_release_cache", "bt_set_tile_bits", "bt_set_tuning", "bt_get_tuning",
    "bt_autotune_batch",
    "bt_server_create", "bt_server_submit", "bt_server_poll", "bt_server_fetch", "bt_server_stats",
    "bt_server_reserve",
    "bt_server_destroy",
]


class BTError(RuntimeError):
    """A nonzero status from libbt (the reference throws strings; hilbert.jl:22-26)."""

    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


class IrOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("k", ctypes.c_int32), ("bits", ctypes.c_int32 * 5),
                ("prob_bits", ctypes.c_int32 * 3), ("ctrl_mask", ctypes.c_uint64), ("mat", ctypes.c_int64),
                ("nkraus", ctypes.c_int32), ("basis", ctypes.c_int32), ("record", ctypes.c_int32),
                ("br0_begin", ctypes.c_int32), ("br0_end", ctypes.c_int32), ("br1_begin", ctypes.c_int32),
                ("br1_end", ctypes.c_int32)]


class RunStats(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_int64), ("reduce_passes", ctypes.c_int64), ("bytes", ctypes.c_double),
                ("boundary_draws", ctypes.c_int64), ("draws", ctypes.c_int64), ("kernel_ms", ctypes.c_double),
                ("launches", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("tile_bits", ctypes.c_int32), ("coalesce_bits", ctypes.c_int32), ("tiles_per_block", ctypes.c_int32),
                ("ms_per_traj", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int32), ("nout", ctypes.c_int32),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double)]


_lib = None
_lock = threading.Lock()


def lib():
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbt.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        i64p, u64p = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_uint64)
        sig = {
            "bt_last_error": (ctypes.c_char_p, []),
            "bt_version": (ctypes.c_int, []),
            "bt_init": (ctypes.c_int, [ctypes.c_int]),
            "bt_uniform": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "bt_state_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
            "bt_state_destroy": (ctypes.c_int, [vp]),
            "bt_state_info": (ctypes.c_int, [vp, ip, ip, ip]),
            "bt_state_set_basis": (ctypes.c_int, [vp, ctypes.c_uint64]),
            "bt_state_copy": (ctypes.c_int, [vp, vp]),
            "bt_state_upload": (ctypes.c_int, [vp, dp, ctypes.c_uint64]),
            "bt_state_download": (ctypes.c_int, [vp, dp, ctypes.c_uint64]),
            "bt_state_device_ptr": (ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
            "bt_state_sync": (ctypes.c_int, [vp]),
            "bt_apply_full": (ctypes.c_int, [vp, dp]),
            "bt_apply_fused5": (ctypes.c_int, [vp, dp, ip]),
            "bt_inner": (ctypes.c_int, [vp, vp, dp, dp]),
            "bt_expect": (ctypes.c_int, [vp, ctypes.c_int, dp, ip, ctypes.c_uint64, ctypes.c_uint64, dp, dp]),
            "bt_sample": (ctypes.c_int, [vp, ctypes.c_int64, dp, u64p, i64p]),
            "bt_sample_noalloc": (ctypes.c_int, [vp, ctypes.c_int64, dp, u64p, i64p]),
            "bt_run_batch_sample": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                                   ip, ctypes.c_int64, i64p, u64p, i64p, ctypes.POINTER(RunStats)]),
            "bt_graph_create_fused": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.POINTER(vp)]),
            "bt_graph_launch": (ctypes.c_int, [vp]),
            "bt_graph_passes": (ctypes.c_int64, [vp]),
            "bt_graph_destroy": (ctypes.c_int, [vp]),
            "bt_run_fused": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, i64p,
                                            ctypes.POINTER(RunStats)]),
            "bt_state_bytes": (ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "bt_apply_matrix": (ctypes.c_int, [vp, ctypes.c_int, dp, ip, ctypes.c_uint64]),
            "bt_reduced_rho": (ctypes.c_int, [vp, ctypes.c_int, ip, dp, dp]),
            "bt_apply_kraus_traj": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, ip, ip, ctypes.c_double, ip,
                                                   dp]),
            "bt_apply_kraus_dm": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, ip]),
            "bt_measure": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, ip, dp]),
            "bt_reset": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ip]),
            "bt_probabilities": (ctypes.c_int, [vp, dp]),
            "bt_norm2": (ctypes.c_int, [vp, dp]),
            "bt_trace_purity": (ctypes.c_int, [vp, dp, dp]),
            "bt_gather": (ctypes.c_int, [vp, u64p, ctypes.c_uint64, dp]),
            "bt_circuit_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(IrOp), ctypes.c_int, ctypes.c_int, dp,
                                                 ctypes.c_int64, ctypes.POINTER(vp)]),
            "bt_circuit_destroy": (ctypes.c_int, [vp]),
            "bt_circuit_max_draws": (ctypes.c_int64, [vp]),
            "bt_run_circuit": (ctypes.c_int, [vp, vp, ctypes.c_uint64, ctypes.c_uint64, dp, ctypes.c_int64, ip,
                                              ctypes.c_int64, i64p, ctypes.POINTER(RunStats)]),
            "bt_run_circuit_dm": (ctypes.c_int, [vp, vp, ctypes.POINTER(RunStats)]),
            "bt_run_batch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ip,
                                            ctypes.c_int64, i64p, dp, ctypes.POINTER(RunStats)]),
            "bt_run_batch_persistent": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int64,
                                                       ctypes.c_int64, ip, ctypes.c_int64, i64p, dp,
                                                       ctypes.POINTER(RunStats)]),
            "bt_set_tile_bits": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
            "bt_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "bt_get_tuning": (ctypes.c_int, [ctypes.c_int, ip, ip, ip]),
            "bt_autotune_batch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                                 ctypes.POINTER(Tuning)]),
            "bt_server_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(vp)]),
            "bt_server_submit": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                ctypes.c_uint64, i64p]),
            "bt_server_poll": (ctypes.c_int, [vp, ctypes.POINTER(Result), ctypes.c_int, ctypes.c_int, ip]),
            "bt_server_fetch": (ctypes.c_int, [vp, ctypes.c_int64, ip, ctypes.c_int64, i64p, dp, ctypes.c_uint64]),
            "bt_server_stats": (ctypes.c_int, [vp, ip, ip, u64p]),
            "bt_autotune_fused": (ctypes.c_int, [vp, vp, ip, dp]),
            "bt_graph_create_dm": (ctypes.c_int, [vp, vp, ctypes.POINTER(vp)]),
            "bt_server_reserve": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]),
            "bt_server_destroy": (ctypes.c_int, [vp]),
            "bt_release_cache": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def check(rc):
    if rc != 0:
        raise BTError(rc, lib().bt_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
