"""Host-side mirror of the reference's graphstore::Graph and Ligra surface.

Same names and argument meaning as include/graph.h (class Graph),
include/VertexSubset.cpp (VertexSubset), include/Map.cpp (edgeMap) and the
algorithm headers (BFS_with_edge_map, PR_S, CC, BC, TC), backed by
libterrace_hip.so on the GPU.  Every call goes through the C ABI; there is
no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (OP_BC_BACK, OP_BC_FWD, OP_BF, OP_BFS, OP_CC, OP_PR, OP_TC, DeviceBuffer, OpArgs, Request,
                   Result, TerraceError, check, lib, ptr)

__all__ = ["Graph", "VertexSubset", "edgeMap", "BFS_with_edge_map", "PR_S", "CC", "BC", "TC",
           "bfs_multi", "local_triangles", "Server", "rmat_device", "TerraceError", "DeviceBuffer"]


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


class Graph:
    """graphstore::Graph (include/graph.h:73-230) resident in HBM."""

    def __init__(self, size: int, weighted: bool = False):
        h = ctypes.c_void_p()
        create = lib().tg_create_weighted if weighted else lib().tg_create
        check(create(int(size), ctypes.byref(h)))
        self._h = h
        self._n = int(size)

    @property
    def weighted(self) -> bool:
        out = ctypes.c_int()
        check(lib().tg_is_weighted(self._h, ctypes.byref(out)))
        return bool(out.value)

    def add_edge_batch_weighted(self, srcs, dests, weights) -> None:
        """Weighted add_edge_batch (include/graph.h:430-436): the first
        weight given for an edge is kept; existing edges keep theirs."""
        s, d, w = _u32(srcs), _u32(dests), _u32(weights)
        if not (len(s) == len(d) == len(w)):
            raise ValueError("srcs, dests and weights differ in length")
        check(lib().tg_insert_batch_weighted(self._h, ptr(s), ptr(d), ptr(w), len(s)))

    def edge_weights(self, srcs, dests) -> np.ndarray:
        """Weight of each edge, 0xFFFFFFFF where absent."""
        s, d = _u32(srcs), _u32(dests)
        out = np.empty(len(s), np.uint32)
        check(lib().tg_edge_weights(self._h, ptr(s), ptr(d), len(s), ptr(out)))
        return out

    @classmethod
    def from_adjacency_graph(cls, path: str) -> "Graph":
        """A graph holding the edges of a Ligra AdjacencyGraph file (the load
        in src/graph_bm.cc:236-266)."""
        h = ctypes.c_void_p()
        check(lib().tg_load_adjacency_graph(str(path).encode(), ctypes.byref(h)))
        G = cls.__new__(cls)
        G._h = h
        n = ctypes.c_uint32()
        check(lib().tg_num_vertices(h, ctypes.byref(n)))
        G._n = n.value
        return G

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().tg_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    # -- updates (graph.h:434-502, 507-641, 675-791)
    def add_edge_batch(self, srcs, dests, edge_count=None, perm=None) -> None:
        """Insert edges; any order, duplicates ignored.  perm is accepted for
        signature compatibility and ignored (the device sorts)."""
        s, d = _u32(srcs), _u32(dests)
        if len(s) != len(d):
            raise ValueError("srcs and dests differ in length")
        n = len(s) if edge_count is None else int(edge_count)
        check(lib().tg_insert_batch(self._h, ptr(s), ptr(d), n))

    def remove_edge_batch(self, srcs, dests, edge_count=None) -> None:
        s, d = _u32(srcs), _u32(dests)
        if len(s) != len(d):
            raise ValueError("srcs and dests differ in length")
        n = len(s) if edge_count is None else int(edge_count)
        check(lib().tg_delete_batch(self._h, ptr(s), ptr(d), n))

    def update_batch(self, ins_src, ins_dst, del_src, del_dst) -> None:
        """A mixed batch: add_edge_batch(ins) then remove_edge of every
        (del) edge (src/graph_bm.cc:354-378), prepared together in one call."""
        a, b, c, d = _u32(ins_src), _u32(ins_dst), _u32(del_src), _u32(del_dst)
        if len(a) != len(b) or len(c) != len(d):
            raise ValueError("source and destination arrays differ in length")
        check(lib().tg_update_batch(self._h, ptr(a), ptr(b), len(a), ptr(c), ptr(d), len(c)))

    def update_batch_dev(self, ins_src: DeviceBuffer, ins_dst: DeviceBuffer, n_ins: int, del_src: DeviceBuffer,
                         del_dst: DeviceBuffer, n_del: int) -> None:
        check(lib().tg_update_batch_dev(self._h, ins_src.ptr, ins_dst.ptr, int(n_ins), del_src.ptr, del_dst.ptr,
                                        int(n_del)))

    def add_edge_batch_dev(self, src: DeviceBuffer, dst: DeviceBuffer, count: int) -> None:
        check(lib().tg_insert_batch_dev(self._h, src.ptr, dst.ptr, int(count)))

    def remove_edge_batch_dev(self, src: DeviceBuffer, dst: DeviceBuffer, count: int) -> None:
        check(lib().tg_delete_batch_dev(self._h, src.ptr, dst.ptr, int(count)))

    def add_edge(self, s: int, d: int) -> int:
        r = ctypes.c_int()
        check(lib().tg_add_edge(self._h, s, d, ctypes.byref(r)))
        return r.value

    def remove_edge(self, s: int, d: int) -> int:
        r = ctypes.c_int()
        check(lib().tg_remove_edge(self._h, s, d, ctypes.byref(r)))
        return r.value

    # -- queries
    def is_edge(self, s: int, d: int) -> int:
        r = ctypes.c_uint32()
        check(lib().tg_is_edge(self._h, s, d, ctypes.byref(r)))
        return r.value

    def is_edge_batch(self, srcs, dests) -> np.ndarray:
        s, d = _u32(srcs), _u32(dests)
        out = np.zeros(len(s), np.uint8)
        check(lib().tg_is_edge_batch(self._h, ptr(s), ptr(d), len(s), ptr(out)))
        return out

    def count_common(self, s: int, d: int) -> int:
        """Graph::count_common (include/graph.h:1085-1255): common
        neighbours of s and d below min(s, d)."""
        return int(self.count_common_batch([s], [d])[0])

    def count_common_batch(self, srcs, dests) -> np.ndarray:
        s, d = _u32(srcs), _u32(dests)
        if len(s) != len(d):
            raise ValueError("srcs and dests differ in length")
        out = np.zeros(len(s), np.uint32)
        check(lib().tg_count_common(self._h, ptr(s), ptr(d), len(s), ptr(out)))
        return out

    def degree(self, v: int) -> int:
        r = ctypes.c_uint32()
        check(lib().tg_degree(self._h, v, ctypes.byref(r)))
        return r.value

    def degrees(self) -> np.ndarray:
        out = np.empty(self._n, np.uint32)
        check(lib().tg_degrees_all(self._h, ptr(out)))
        return out

    def neighbors(self, v: int) -> np.ndarray:
        deg = ctypes.c_uint32()
        check(lib().tg_neighbors(self._h, v, None, 0, ctypes.byref(deg)))
        out = np.empty(deg.value, np.uint32)
        if deg.value:
            check(lib().tg_neighbors(self._h, v, ptr(out), deg.value, ctypes.byref(deg)))
        return out

    def get_num_edges(self) -> int:
        r = ctypes.c_uint64()
        check(lib().tg_num_edges(self._h, ctypes.byref(r)))
        return r.value

    def get_num_vertices(self) -> int:
        return self._n

    def get_size(self) -> int:
        r = ctypes.c_uint64()
        check(lib().tg_memory_bytes(self._h, ctypes.byref(r)))
        return r.value

    def export_adjacency(self):
        """(offsets[n+1] uint64, neighbours[m] uint32), sorted per vertex."""
        m = self.get_num_edges()
        off = np.empty(self._n + 1, np.uint64)
        nbr = np.empty(max(m, 1), np.uint32)
        check(lib().tg_export_adjacency(self._h, ptr(off), ptr(nbr)))
        return off, nbr[:m]

    def tiers(self) -> np.ndarray:
        out = np.empty(self._n, np.uint8)
        check(lib().tg_tiers(self._h, ptr(out)))
        return out

    def stats(self) -> dict:
        out = np.zeros(18, np.uint64)
        check(lib().tg_stats(self._h, ptr(out)))
        keys = ["pma_rebuilds", "pma_windows", "pma_window_leaves", "tree_blocks_written",
                "promotions", "trees", "pma_slots", "tree_small_updates", "pma_leaf_merges",
                "pma_elements", "last_ins_keys", "last_ins_sources", "last_del_keys", "last_del_sources",
                "pma_low_windows", "pma_shrinks", "sum_ins_sources", "sum_del_sources"]
        return {k: int(v) for k, v in zip(keys, out)}

    def pma_leaf_counts(self) -> np.ndarray:
        """Element count of every PMA leaf (64 slots each)."""
        nl = ctypes.c_uint64()
        check(lib().tg_pma_leaf_counts(self._h, None, 0, ctypes.byref(nl)))
        out = np.zeros(nl.value, np.uint32)
        check(lib().tg_pma_leaf_counts(self._h, ptr(out), nl.value, ctypes.byref(nl)))
        return out

    def apply_ops(self, src, dst, ops) -> None:
        """Coalesced updates: ops[i] 0 = insert, 1 = delete (src[i], dst[i]),
        applied in order in one device pass (last op on an edge wins)."""
        s, d, o = _u32(src), _u32(dst), _u32(ops)
        assert len(s) == len(d) == len(o)
        check(lib().tg_apply_ops(self._h, ptr(s), ptr(d), ptr(o), len(s)))

    def autotune(self) -> dict:
        """Sweep push/pull thresholds and the pull block size on this graph;
        later algorithm calls use the fastest (results are unchanged)."""
        r = np.zeros(12, np.float64)
        check(lib().tg_autotune(self._h, ptr(r)))
        keys = ["thr_bfs", "thr_ms", "thr_cc", "pull_block", "bfs_ref_ms", "bfs_tuned_ms", "ms_ref_ms",
                "ms_tuned_ms", "cc_ref_ms", "cc_tuned_ms", "pr_ref_ms", "pr_tuned_ms"]
        out = {k: float(v) for k, v in zip(keys, r)}
        for k in keys[:4]:
            out[k] = int(out[k])
        return out

    TUNING_KEYS = ("thr_bfs", "thr_ms", "thr_cc", "pull_block", "pr_graph", "pr_sliced", "push_wave", "level_graph")

    def get_tuning(self) -> dict:
        t = np.zeros(len(self.TUNING_KEYS), np.uint32)
        check(lib().tg_get_tuning(self._h, ptr(t)))
        return {k: int(v) for k, v in zip(self.TUNING_KEYS, t)}

    def set_tuning(self, **kw) -> None:
        cur = self.get_tuning()
        cur.update(kw)
        t = np.array([cur[k] for k in self.TUNING_KEYS], np.uint32)
        check(lib().tg_set_tuning(self._h, ptr(t)))

    def tier_counts(self) -> dict:
        out = np.zeros(4, np.uint64)
        check(lib().tg_tier_counts(self._h, ptr(out)))
        return {"pma_vertices": int(out[0]), "pma_edges": int(out[1]),
                "tree_vertices": int(out[2]), "tree_edges": int(out[3])}

    # -- profiling
    def profile(self, on: bool = True) -> None:
        check(lib().tg_profile_enable(self._h, int(on)))

    def profile_reset(self) -> None:
        check(lib().tg_profile_reset(self._h))

    def profile_get(self, name: str):
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        check(lib().tg_profile_get(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class VertexSubset:
    """VertexSubset (include/VertexSubset.cpp): VertexSubset(G, e) holds one
    vertex, VertexSubset(G, 0, all=True) every vertex.  With G = None and
    max_el = n it is the reference's graph-less VertexSubset(e, n, all)
    (VertexSubset.cpp:109), bound to a graph by its first edgeMap;
    VertexSubset.dense(els) is VertexSubset(bool *els, len) (:119)."""

    def __init__(self, G: Graph = None, e=None, all: bool = False, max_el: int = None):
        h = ctypes.c_void_p()
        if G is None:
            if not max_el:
                raise ValueError("a VertexSubset without a graph needs max_el")
            check(lib().tg_frontier_create_n(int(max_el), ctypes.byref(h)))
        else:
            check(lib().tg_frontier_create(G.handle, ctypes.byref(h)))
        self._h, self._g = h, G
        if all:
            check(lib().tg_frontier_set_all(h))
        elif e is not None:
            v = _u32(np.atleast_1d(e))
            check(lib().tg_frontier_set(h, ptr(v), len(v)))

    @classmethod
    def dense(cls, els) -> "VertexSubset":
        els = np.ascontiguousarray(els, dtype=np.uint8)
        vs = cls(None, max_el=len(els))
        check(lib().tg_frontier_set_dense(vs._h, ptr(els), len(els)))
        return vs

    @property
    def handle(self):
        return self._h

    def get_n(self) -> int:
        r = ctypes.c_uint64()
        check(lib().tg_frontier_size(self._h, ctypes.byref(r)))
        return r.value

    def not_empty(self) -> bool:
        return self.get_n() > 0

    def to_array(self) -> np.ndarray:
        c = ctypes.c_uint64()
        n = self.get_n()
        out = np.empty(max(n, 1), np.uint32)
        check(lib().tg_frontier_get(self._h, ptr(out), n, ctypes.byref(c)))
        return out[:c.value]

    def del_(self) -> None:
        if self._h is not None and self._h.value:
            lib().tg_frontier_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.del_()
        except Exception:
            pass


def edgeMap(G: Graph, vs: VertexSubset, op: int, args: OpArgs, output: bool = True,
            threshold: int = 20):
    """edgeMap(G, vs, F, output, threshold) (include/Map.cpp:66-74) with F
    one of the device ops (OP_BFS, OP_CC, OP_BC_FWD, OP_BC_BACK, OP_PR,
    OP_BF, OP_TC; OP_TC takes no output frontier)."""
    out = VertexSubset(G) if output else None
    check(lib().tg_edge_map(G.handle, vs.handle, int(op), ctypes.byref(args),
                            out.handle if out else None, int(threshold)))
    return out


# -- algorithm entry points (src/graph_bm.cc test_* drivers)
def BFS_with_edge_map(G: Graph, src: int, with_levels: bool = False):
    """include/BFS.h:41-84: parents (int32, -1 unreached)."""
    parents = np.empty(G.get_num_vertices(), np.int32)
    levels = np.empty(G.get_num_vertices(), np.uint32)
    check(lib().tg_bfs(G.handle, int(src), ptr(parents), ptr(levels)))
    return (parents, levels) if with_levels else parents


def PR_S(G: Graph, maxIters: int) -> np.ndarray:
    """include/Pagerank.h:106-184 (undamped, FP64)."""
    out = np.empty(G.get_num_vertices(), np.float64)
    check(lib().tg_pagerank(G.handle, int(maxIters), ptr(out)))
    return out


def CC(G: Graph) -> np.ndarray:
    """include/Components.h:92-120."""
    out = np.empty(G.get_num_vertices(), np.uint32)
    check(lib().tg_cc(G.handle, ptr(out)))
    return out


def BC(G: Graph, start: int) -> np.ndarray:
    """include/BC.h:99-158 (unreachable -> NaN)."""
    out = np.empty(G.get_num_vertices(), np.float64)
    check(lib().tg_bc(G.handle, int(start), ptr(out)))
    return out


def SSSP_BF(G: Graph, start: int) -> np.ndarray:
    """include/BellmanFordUnweighted.h:60-130: int32 hop distances, INT_MAX/2
    where unreachable."""
    out = np.empty(G.get_num_vertices(), np.int32)
    check(lib().tg_sssp_bf(G.handle, int(start), ptr(out)))
    return out


def TC(G: Graph) -> int:
    """include/TC.h:73-87."""
    r = ctypes.c_uint64()
    check(lib().tg_tc(G.handle, ctypes.byref(r)))
    return r.value


def bfs_multi(G: Graph, sources) -> np.ndarray:
    """BFS levels from up to 64 sources in one traversal: array [k, n]."""
    srcs = _u32(sources)
    out = np.empty((len(srcs), G.get_num_vertices()), np.uint32)
    check(lib().tg_bfs_multi(G.handle, ptr(srcs), len(srcs), ptr(out)))
    return out


def read_adjacency_graph(path: str):
    """Ligra AdjacencyGraph file (include/io_util.h:112-145): (n, src, dst)
    with the edges in file order."""
    n, m = ctypes.c_uint32(), ctypes.c_uint64()
    check(lib().tg_read_adjacency_graph(str(path).encode(), ctypes.byref(n), ctypes.byref(m), None, None, 0))
    src = np.empty(m.value, np.uint32)
    dst = np.empty(m.value, np.uint32)
    check(lib().tg_read_adjacency_graph(str(path).encode(), ctypes.byref(n), ctypes.byref(m), ptr(src), ptr(dst),
                                        m.value))
    return n.value, src, dst


def write_adjacency_graph(path: str, n: int, src, dst) -> None:
    """Writes edges (sorted by source) as a Ligra AdjacencyGraph file."""
    src, dst = _u32(src), _u32(dst)
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.searchsorted(src, np.arange(n, dtype=np.uint32), side="left").astype(np.int64)
    with open(path, "w") as f:
        f.write(f"AdjacencyGraph\n{n}\n{len(dst)}\n")
        f.write("\n".join(map(str, off.tolist())))
        f.write("\n")
        f.write("\n".join(map(str, dst.tolist())))
        f.write("\n")


def bc_multi(G: Graph, sources) -> np.ndarray:
    """Betweenness from up to 64 sources in one traversal: array [k, n]
    (NaN where a vertex is unreachable from that source)."""
    srcs = _u32(sources)
    out = np.empty((len(srcs), G.get_num_vertices()), np.float64)
    check(lib().tg_bc_multi(G.handle, ptr(srcs), len(srcs), ptr(out)))
    return out


def local_triangles(G: Graph, verts) -> np.ndarray:
    """Triangles through each vertex of verts."""
    v = _u32(verts)
    out = np.zeros(len(v), np.uint64)
    check(lib().tg_ltc(G.handle, ptr(v), len(v), ptr(out)))
    return out


REQ_UPDATE, REQ_BFS, REQ_CC, REQ_PR, REQ_LTC = 1, 2, 3, 4, 5


SERVE_MODES = {"plain": 0, "strict": 1, "admission": 2}


class Server:
    """Request loop over one graph (tg_server_*).  mode "plain" replays the
    stream one request at a time; "strict" batches arrival-ordered runs (each
    query sees exactly the updates that arrived before it); "admission"
    pools queries and admits them together on the version current at
    admission.  batched=True is "strict"."""

    def __init__(self, G: Graph, batched: bool = False, mem_budget: int = 1 << 40, keep_results: bool = False,
                 mode: str | None = None):
        m = SERVE_MODES[mode] if mode is not None else int(bool(batched))
        h = ctypes.c_void_p()
        check(lib().tg_server_create(G.handle, m, int(mem_budget), int(keep_results),
                                     ctypes.byref(h)))
        self._h, self._g = h, G

    def submit(self, type: int, arrival: float, src: int = 0, iters: int = 10,
               ins=None, dels=None) -> int:
        r = Request(type=type, arrival_s=float(arrival), src=int(src), iters=int(iters))
        keep = []  # host arrays referenced until submit returns (it copies them)
        if ins is not None:
            s, d = _u32(ins[0]), _u32(ins[1])
            r.ins_src, r.ins_dst, r.n_ins = ptr(s), ptr(d), len(s)
            keep += [s, d]
        if dels is not None:
            s, d = _u32(dels[0]), _u32(dels[1])
            r.del_src, r.del_dst, r.n_del = ptr(s), ptr(d), len(s)
            keep += [s, d]
        rid = ctypes.c_uint64()
        check(lib().tg_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        del keep
        return rid.value

    def run(self) -> None:
        check(lib().tg_server_run(self._h))

    def start(self) -> None:
        """Run the loop on a thread of its own; submit() stays open and a
        request submitted now arrives at its submission time."""
        check(lib().tg_server_start(self._h))

    def finish(self) -> None:
        """Close the input, wait for every request and join the loop."""
        check(lib().tg_server_finish(self._h))

    def wait(self, rid: int, timeout: float = 60.0) -> int:
        """Status of request rid (1 pending, 0 done, < 0 error) after waiting
        up to timeout seconds for it to complete."""
        st = ctypes.c_int()
        check(lib().tg_server_wait(self._h, int(rid), float(timeout), ctypes.byref(st)))
        return st.value

    def result(self, rid: int) -> dict:
        r = Result()
        check(lib().tg_server_result(self._h, int(rid), ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in Result._fields_}

    def result_data(self, rid: int, dtype) -> np.ndarray:
        out = np.empty(self._g.get_num_vertices(), dtype)
        check(lib().tg_server_result_data(self._h, int(rid), ptr(out), out.nbytes))
        return out

    def summary(self) -> dict:
        out = np.zeros(10, np.float64)
        check(lib().tg_server_summary(self._h, ptr(out)))
        keys = ["makespan_s", "requests_per_s", "p50_s", "p99_s", "p50_update_s", "p99_update_s",
                "p50_query_s", "p99_query_s", "groups", "rejected"]
        return dict(zip(keys, out.tolist()))

    def close(self) -> None:
        if self._h is not None and self._h.value:
            lib().tg_server_destroy(self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rmat_device(n: int, seed: int, a: float, b: float, c: float, count: int, start: int = 0,
                nominal: bool = True, symmetric: bool = False, drop_loops: bool = False):
    """rMat<uint32_t> edges generated in HBM (include/rmat_util.h:81-115).
    Returns (src DeviceBuffer, dst DeviceBuffer, entries written)."""
    cap = count * (2 if symmetric else 1)
    s, d = DeviceBuffer(cap * 4), DeviceBuffer(cap * 4)
    w = ctypes.c_uint64()
    flags = (1 if symmetric else 0) | (2 if drop_loops else 0)
    check(lib().tg_rmat_dev(int(n), int(seed) & 0xFFFFFFFF, a, b, c, int(nominal), int(start), int(count),
                            flags, s.ptr, d.ptr, ctypes.byref(w)))
    return s, d, w.value


OPS = {"BFS": OP_BFS, "CC": OP_CC, "BC_FWD": OP_BC_FWD, "BC_BACK": OP_BC_BACK, "PR": OP_PR, "BF": OP_BF,
       "TC": OP_TC}
