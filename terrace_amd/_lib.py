"""ctypes binding of libterrace_hip.so (include/terrace_hip.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libterrace_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "terrace_hip.h")

TG_OK, TG_EINVAL, TG_ENOMEM, TG_EDEVICE, TG_EINTERNAL = 0, -1, -2, -3, -4
OP_BFS, OP_CC, OP_BC_FWD, OP_BC_BACK, OP_PR, OP_TC, OP_BF = 1, 2, 3, 4, 5, 6, 7


class TerraceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class OpArgs(ctypes.Structure):
    _fields_ = [
        ("parent", ctypes.c_void_p),
        ("level", ctypes.c_void_p),
        ("depth", ctypes.c_uint32),
        ("ids", ctypes.c_void_p),
        ("prev", ctypes.c_void_p),
        ("val", ctypes.c_void_p),
        ("src_val", ctypes.c_void_p),
        ("visited", ctypes.c_void_p),
        ("dist", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int),
        ("arrival_s", ctypes.c_double),
        ("src", ctypes.c_uint32),
        ("iters", ctypes.c_uint32),
        ("ins_src", ctypes.c_void_p),
        ("ins_dst", ctypes.c_void_p),
        ("n_ins", ctypes.c_uint64),
        ("del_src", ctypes.c_void_p),
        ("del_dst", ctypes.c_void_p),
        ("n_del", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("arrival_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("version", ctypes.c_uint64),
        ("digest", ctypes.c_uint64),
        ("value", ctypes.c_uint64),
        ("fvalue", ctypes.c_double),
        ("group", ctypes.c_uint64),
    ]


_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "tg_last_error": (ctypes.c_char_p, []),
    "tg_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(_P)]),
    "tg_destroy": (ctypes.c_int, [_P]),
    "tg_create_weighted": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(_P)]),
    "tg_is_weighted": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int)]),
    "tg_insert_batch_weighted": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint64]),
    "tg_insert_batch_weighted_dev": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint64]),
    "tg_edge_weights": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P]),
    "tg_insert_batch": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "tg_delete_batch": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "tg_insert_batch_dev": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "tg_delete_batch_dev": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64]),
    "tg_update_batch": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P, _P, ctypes.c_uint64]),
    "tg_update_batch_dev": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P, _P, ctypes.c_uint64]),
    "tg_add_edge": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
    "tg_remove_edge": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
    "tg_is_edge": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_is_edge_batch": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P]),
    "tg_degree": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_degrees_all": (ctypes.c_int, [_P, _P]),
    "tg_num_edges": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_num_vertices": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_memory_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_export_adjacency": (ctypes.c_int, [_P, _P, _P]),
    "tg_export_adjacency_dev": (ctypes.c_int, [_P, _P, _P]),
    "tg_neighbors": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_tiers": (ctypes.c_int, [_P, _P]),
    "tg_bfs": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P]),
    "tg_pagerank": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "tg_cc": (ctypes.c_int, [_P, _P]),
    "tg_bc": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "tg_tc": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_bfs_dev": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_pagerank_dev": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "tg_cc_dev": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_bc_dev": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "tg_sssp_bf": (ctypes.c_int, [_P, ctypes.c_uint32, _P]),
    "tg_sssp_bf_dev": (ctypes.c_int, [_P, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "tg_frontier_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "tg_frontier_destroy": (ctypes.c_int, [_P]),
    "tg_frontier_set": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "tg_frontier_set_all": (ctypes.c_int, [_P]),
    "tg_frontier_size": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_frontier_get": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_edge_map": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.POINTER(OpArgs), _P, ctypes.c_uint32]),
    "tg_vertex_map": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "tg_frontier_copy": (ctypes.c_int, [_P, _P]),
    "tg_bfs_multi": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "tg_bc_multi": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "tg_bfs_multi_dev": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "tg_bc_multi_dev": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "tg_apply_ops": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_uint64]),
    "tg_autotune": (ctypes.c_int, [_P, _P]),
    "tg_get_tuning": (ctypes.c_int, [_P, _P]),
    "tg_set_tuning": (ctypes.c_int, [_P, _P]),
    "tg_ltc": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P]),
    "tg_server_create": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_P)]),
    "tg_server_destroy": (ctypes.c_int, [_P]),
    "tg_server_submit": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_server_run": (ctypes.c_int, [_P]),
    "tg_server_start": (ctypes.c_int, [_P]),
    "tg_server_finish": (ctypes.c_int, [_P]),
    "tg_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "tg_server_result": (ctypes.c_int, [_P, ctypes.c_uint64, _P]),
    "tg_server_result_data": (ctypes.c_int, [_P, ctypes.c_uint64, _P, ctypes.c_uint64]),
    "tg_server_summary": (ctypes.c_int, [_P, _P]),
    "tg_rmat_dev": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int, _P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "tg_dev_alloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    "tg_dev_free": (ctypes.c_int, [_P]),
    "tg_memcpy_h2d": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "tg_memcpy_d2h": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "tg_memcpy_d2d": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
    "tg_read_adjacency_graph": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(ctypes.c_uint64), _P, _P, ctypes.c_uint64]),
    "tg_load_adjacency_graph": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "tg_device_sync": (ctypes.c_int, []),
    "tg_profile_enable": (ctypes.c_int, [_P, ctypes.c_int]),
    "tg_profile_reset": (ctypes.c_int, [_P]),
    "tg_profile_get": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_uint64)]),
    "tg_stats": (ctypes.c_int, [_P, _P]),
    "tg_pma_leaf_counts": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P]),
    "tg_tier_counts": (ctypes.c_int, [_P, _P]),
    "tg_debug_fail_next": (ctypes.c_int, [_P, ctypes.c_int]),
    "tg_count_common": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P]),
    "tg_count_common_dev": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P]),
    "tg_count_common_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint64, _P]),
    "tg_frontier_create_n": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(_P)]),
    "tg_frontier_set_dense": (ctypes.c_int, [_P, _P, ctypes.c_uint64]),
}


def declared_symbols(header: str = HEADER) -> list[str]:
    """Function names declared in include/terrace_hip.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libterrace_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m terrace_amd.build` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != TG_OK:
        msg = lib().tg_last_error()
        raise TerraceError(rc, msg.decode() if msg else "")


def ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceBuffer:
    """Raw HBM allocation owned from Python (freed on close / GC)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p()
        check(lib().tg_dev_alloc(max(int(nbytes), 1), ctypes.byref(p)))
        self.ptr = p
        self.nbytes = int(nbytes)

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def offset(self, nbytes: int) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.addr + int(nbytes))

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().tg_memcpy_h2d(self.ptr, ptr(a), a.nbytes))

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(int(count), dtype)
        assert out.nbytes <= self.nbytes
        check(lib().tg_memcpy_d2h(ptr(out), self.ptr, out.nbytes))
        return out

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().tg_dev_free(self.ptr)
        self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
