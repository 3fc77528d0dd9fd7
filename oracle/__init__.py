"""CPU oracle for the Terrace hot path — TEST INFRASTRUCTURE ONLY.

numpy/ctypes front end to ``oracle/oracle.c`` (the plain-C restatement of
the reference).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this package, and only as the
checker; the product (``terrace_amd``) never does.

Parity pins for every function are in ``tests/test_oracle.py``:
golden vectors written by the reference build (``oracle/_ref``), the
survey's known-answer values for the generator, closed forms and brute
force on tiny inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile liboracle.so (and the reference build when the reference is
    present on this machine)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/include"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    c = ctypes
    L.orc_hash64.restype = c.c_uint64
    L.orc_hash64.argtypes = [c.c_uint64]
    L.orc_hash32.restype = c.c_uint32
    L.orc_hash32.argtypes = [c.c_uint32]
    L.orc_aspen_ith_rand.restype = c.c_uint64
    L.orc_aspen_ith_rand.argtypes = [c.c_uint64, c.c_uint64]
    L.orc_rmat_edges.restype = c.c_int
    L.orc_rmat_edges.argtypes = [c.c_uint32, c.c_uint32, c.c_double, c.c_double,
                                 c.c_double, c.c_int, c.c_uint64, c.c_uint64,
                                 _u32p, _u32p]
    L.orc_sort_u64.restype = None
    L.orc_sort_u64.argtypes = [_u64p, c.c_uint64]
    L.orc_sort_unique_u64.restype = c.c_uint64
    L.orc_sort_unique_u64.argtypes = [_u64p, c.c_uint64]
    L.orc_build_keys.restype = c.c_uint64
    L.orc_build_keys.argtypes = [_u32p, _u32p, c.c_uint64, c.c_int, _u64p]
    L.orc_set_union.restype = c.c_uint64
    L.orc_set_union.argtypes = [_u64p, c.c_uint64, _u64p, c.c_uint64, _u64p]
    L.orc_set_difference.restype = c.c_uint64
    L.orc_set_difference.argtypes = [_u64p, c.c_uint64, _u64p, c.c_uint64, _u64p]
    L.orc_keys_to_csr.restype = None
    L.orc_keys_to_csr.argtypes = [_u64p, c.c_uint64, c.c_uint32, _u64p, _u32p]
    L.orc_edges_to_csr.restype = c.c_uint64
    L.orc_edges_to_csr.argtypes = [c.c_uint32, _u32p, _u32p, c.c_uint64, c.c_int, c.c_int, _u64p, _u32p]
    L.orc_csr_update.restype = c.c_uint64
    L.orc_csr_update.argtypes = [c.c_uint32, _u64p, _u32p, _u64p, c.c_uint64, _u64p, c.c_uint64, _u64p, _u32p]
    L.orc_tier_insert.restype = None
    L.orc_tier_insert.argtypes = [c.c_void_p, c.c_uint32, _u64p, _u32p, _u64p, c.c_uint64]
    L.orc_bfs_levels.restype = None
    L.orc_bfs_levels.argtypes = [c.c_uint32, _u64p, _u32p, c.c_uint32, _u32p]
    L.orc_bfs_check_parents.restype = c.c_uint64
    L.orc_bfs_check_parents.argtypes = [c.c_uint32, _u64p, _u32p, c.c_uint32, _i32p, _u32p]
    L.orc_pagerank.restype = None
    L.orc_pagerank.argtypes = [c.c_uint32, _u64p, _u32p, c.c_uint32, _f64p]
    L.orc_cc.restype = None
    L.orc_cc.argtypes = [c.c_uint32, _u64p, _u32p, _u32p]
    L.orc_bc.restype = None
    L.orc_bc.argtypes = [c.c_uint32, _u64p, _u32p, c.c_uint32, _f64p]
    L.orc_tc.restype = c.c_uint64
    L.orc_tc.argtypes = [c.c_uint32, _u64p, _u32p]
    L.orc_ltc.restype = c.c_uint64
    L.orc_ltc.argtypes = [_u64p, _u32p, c.c_uint32]
    L.orc_sssp_bf.restype = None
    L.orc_sssp_bf.argtypes = [c.c_uint32, _u64p, _u32p, c.c_uint32, _i32p]
    L.orc_sssp_bf_weighted.restype = None
    L.orc_sssp_bf_weighted.argtypes = [c.c_uint32, _u64p, _u32p, _u32p, c.c_uint32, _i32p]
    L.orc_count_common.restype = c.c_uint32
    L.orc_count_common.argtypes = [_u64p, _u32p, c.c_uint32, c.c_uint32]
    _lib = L
    return L


# ---------------------------------------------------------------- generator
def hash32(x: int) -> int:
    return lib().orc_hash32(x)


def hash64(x: int) -> int:
    return lib().orc_hash64(x)


def aspen_ith_rand(i: int, state: int = 0) -> int:
    return lib().orc_aspen_ith_rand(state, i)


def rmat_edges(n: int, seed: int, a: float, b: float, c: float, count: int,
               start: int = 0, nominal: bool = True):
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    rc = lib().orc_rmat_edges(n, seed & 0xFFFFFFFF, a, b, c, int(nominal), start,
                              count, src, dst)
    if rc != 0:
        raise ValueError("rMat: n must be a power of two and a+b+c <= 1")
    return src, dst


# ---------------------------------------------------------------- edge sets
def build_keys(src, dst, symmetric=True) -> np.ndarray:
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    out = np.empty(max(2 * len(src), 1), np.uint64)
    m = lib().orc_build_keys(src, dst, len(src), int(symmetric), out)
    return out[:m].copy()


def sort_unique(keys) -> np.ndarray:
    k = np.array(keys, dtype=np.uint64, copy=True)
    m = lib().orc_sort_unique_u64(k, len(k)) if len(k) else 0
    return k[:m]


def set_union(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.empty(max(len(a) + len(b), 1), np.uint64)
    m = lib().orc_set_union(a, len(a), b, len(b), out)
    return out[:m].copy()


def set_difference(a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    out = np.empty(max(len(a), 1), np.uint64)
    m = lib().orc_set_difference(a, len(a), b, len(b), out)
    return out[:m].copy()


class CSR:
    """Sorted adjacency of a directed key set."""

    def __init__(self, keys: np.ndarray, n: int):
        keys = np.ascontiguousarray(keys, np.uint64)
        self.n = n
        self.m = len(keys)
        self.off = np.empty(n + 1, np.uint64)
        self.nbr = np.empty(max(self.m, 1), np.uint32)
        lib().orc_keys_to_csr(keys, self.m, n, self.off, self.nbr)

    @classmethod
    def from_edges(cls, n: int, src, dst, symmetric=True, drop_loops=True) -> "CSR":
        """The CSR of an edge list without sorting 64-bit keys (equal to
        CSR(build_keys(src, dst, symmetric), n) when drop_loops)."""
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        g = cls.__new__(cls)
        g.n = n
        g.off = np.empty(n + 1, np.uint64)
        nbr = np.empty(max(len(src) * (2 if symmetric else 1), 1), np.uint32)
        g.m = int(lib().orc_edges_to_csr(n, src, dst, len(src), int(symmetric), int(drop_loops), g.off, nbr))
        g.nbr = nbr[:max(g.m, 1)]
        return g

    def updated(self, ins, dels) -> "CSR":
        """(self ∪ ins) \\ dels as a new CSR (sorted unique keys)."""
        ins = np.ascontiguousarray(ins, np.uint64)
        dels = np.ascontiguousarray(dels, np.uint64)
        g = CSR.__new__(CSR)
        g.n = self.n
        g.off = np.empty(self.n + 1, np.uint64)
        nbr = np.empty(max(self.m + len(ins), 1), np.uint32)
        g.m = int(lib().orc_csr_update(self.n, self.off, self.nbr, ins, len(ins), dels, len(dels), g.off, nbr))
        g.nbr = nbr[:max(g.m, 1)]
        return g

    def degrees(self) -> np.ndarray:
        return (self.off[1:] - self.off[:-1]).astype(np.uint32)

    def neighbors(self, v: int) -> np.ndarray:
        return self.nbr[int(self.off[v]):int(self.off[v + 1])]


# ---------------------------------------------------------------- tier model
TIER_DTYPE = np.dtype([("degree", np.uint32), ("btree", np.uint8)], align=False)


def tier_model(n: int) -> np.ndarray:
    # orc_tier_t is {uint32 degree; uint8 btree;} padded to 8 bytes
    return np.zeros(n, dtype=np.dtype([("degree", np.uint32), ("btree", np.uint8),
                                       ("_pad", np.uint8, 3)]))


def tier_insert(tiers: np.ndarray, before: CSR, keys: np.ndarray) -> None:
    keys = np.ascontiguousarray(keys, np.uint64)
    lib().orc_tier_insert(tiers.ctypes.data, before.n, before.off, before.nbr, keys, len(keys))


# ---------------------------------------------------------------- algorithms
def bfs_levels(g: CSR, src: int) -> np.ndarray:
    out = np.empty(g.n, np.uint32)
    lib().orc_bfs_levels(g.n, g.off, g.nbr, src, out)
    return out


def bfs_check_parents(g: CSR, src: int, parents, levels) -> int:
    parents = np.ascontiguousarray(parents, np.int32)
    levels = np.ascontiguousarray(levels, np.uint32)
    return int(lib().orc_bfs_check_parents(g.n, g.off, g.nbr, src, parents, levels))


def pagerank(g: CSR, iters: int) -> np.ndarray:
    out = np.empty(g.n, np.float64)
    lib().orc_pagerank(g.n, g.off, g.nbr, iters, out)
    return out


def cc(g: CSR) -> np.ndarray:
    out = np.empty(g.n, np.uint32)
    lib().orc_cc(g.n, g.off, g.nbr, out)
    return out


def bc(g: CSR, src: int) -> np.ndarray:
    out = np.empty(g.n, np.float64)
    lib().orc_bc(g.n, g.off, g.nbr, src, out)
    return out


def tc(g: CSR) -> int:
    return int(lib().orc_tc(g.n, g.off, g.nbr))


def count_common(g: CSR, s: int, d: int) -> int:
    return int(lib().orc_count_common(g.off, g.nbr, s, d))


def ltc(g: CSR, v: int) -> int:
    return int(lib().orc_ltc(g.off, g.nbr, v))


def sssp_bf(g: CSR, src: int) -> np.ndarray:
    out = np.empty(g.n, np.int32)
    lib().orc_sssp_bf(g.n, g.off, g.nbr, src, out)
    return out


def sssp_bf_weighted(g: CSR, w: np.ndarray, src: int) -> np.ndarray:
    """w: one length per CSR entry (g.nbr order)."""
    out = np.empty(g.n, np.int32)
    lib().orc_sssp_bf_weighted(g.n, g.off, g.nbr, np.ascontiguousarray(w[:g.m], np.uint32), src, out)
    return out


# ---------------------------------------------------------------- weighted edge sets
def weighted_insert(keys, wts, ins_keys, ins_w):
    """The weighted build's add_edge_batch as a set model: within the batch
    the first weight of a key counts; keys already present keep theirs."""
    ins_keys = np.asarray(ins_keys, np.uint64)
    ins_w = np.asarray(ins_w, np.uint32)
    uk, first = np.unique(ins_keys, return_index=True)
    uw = ins_w[first]
    new = ~np.isin(uk, keys)
    allk = np.concatenate([keys, uk[new]])
    allw = np.concatenate([wts, uw[new]])
    order = np.argsort(allk, kind="stable")
    return allk[order], allw[order]


def weighted_delete(keys, wts, del_keys):
    keep = ~np.isin(keys, np.asarray(del_keys, np.uint64))
    return keys[keep], wts[keep]


def rel_l1(a: np.ndarray, b: np.ndarray) -> float:
    """Relative L1 distance over entries finite in both (NaN/inf masked)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mask = np.isfinite(a) & np.isfinite(b)
    den = np.abs(b[mask]).sum()
    num = np.abs(a[mask] - b[mask]).sum()
    return float(num / den) if den > 0 else float(num)
