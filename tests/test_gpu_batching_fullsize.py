"""GPU tests of the batching workload at full size: rMAT scale 24
(16,777,216 vertex ids), 136M generated edges (a=.5 b=.1 c=.1), symmetrized.

The oracle's own generator gives the key set, and the device-built graph's
adjacency must equal it.  Then:
* 64 BFS sources in one multi-source traversal: every lane equals the
  single-source GPU BFS, and 8 lanes equal the oracle's BFS levels;
* 64 BC sources in one multi-source traversal: 4 lanes within 1e-9 of the
  oracle's BC (relative L1, NaN where unreachable);
* 512 small update batches (1..1000 edges, log-uniform sizes, 90% inserts of
  fresh rMAT edges / 10% deletes of existing edges) coalesced 64 per launch
  through apply_ops: the adjacency equals the set model with the last op per
  edge winning."""
import ctypes

import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import DeviceBuffer, check, lib, ptr

pytestmark = pytest.mark.gpu
FP_TOL = 1e-9
UNREACHED = np.iinfo(np.uint32).max
N = 1 << 24
M0 = 136_000_000
PROBS = (0.5, 0.1, 0.1)


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


@pytest.fixture(scope="module")
def batching():
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    say("batching workload: oracle keys")
    keys = oracle.build_keys(*oracle.rmat_edges(N, seed, *PROBS, M0))
    say("batching workload: device build", len(keys))
    s, d, w = tg.rmat_device(N, seed, *PROBS, M0, nominal=True, symmetric=True, drop_loops=True)
    G = tg.Graph(N)
    G.add_edge_batch_dev(s, d, w)
    del s, d
    st = {"G": G, "keys": keys, "g": oracle.CSR(keys, N)}
    yield st
    G.close()


def test_batching_graph_equals_oracle(batching):
    G, g = batching["G"], batching["g"]
    off, nbr = G.export_adjacency()
    assert np.array_equal(off, g.off), "offsets"
    assert np.array_equal(nbr, g.nbr[:g.m]), "neighbours"
    assert g.m > 260_000_000


def test_bfs_multi_full_size(batching):
    G, g = batching["G"], batching["g"]
    say("batching workload: 64-lane BFS")
    rng = np.random.default_rng(4242)
    active = np.nonzero(g.degrees())[0]
    srcs = np.ascontiguousarray(active[rng.integers(0, len(active), 64)], np.uint32)
    srcs[1] = 9
    mlev = DeviceBuffer(64 * 4 * N)
    check(lib().tg_bfs_multi_dev(G.handle, ptr(srcs), 64, mlev.ptr))
    par, lev = DeviceBuffer(4 * N), DeviceBuffer(4 * N)
    depth = ctypes.c_uint32()
    lane = np.empty(N, np.uint32)
    for j in range(64):
        check(lib().tg_bfs_dev(G.handle, int(srcs[j]), par.ptr, lev.ptr, ctypes.byref(depth)))
        single = lev.download(np.uint32, N)
        check(lib().tg_memcpy_d2h(ptr(lane), mlev.offset(4 * N * j), 4 * N))
        assert np.array_equal(lane, single), f"lane {j} differs from single-source BFS"
        if j % 8 == 0:
            assert np.array_equal(lane, oracle.bfs_levels(g, int(srcs[j]))), f"lane {j} differs from the oracle"
    mlev.close(), par.close(), lev.close()


def test_bc_multi_full_size(batching):
    G, g = batching["G"], batching["g"]
    say("batching workload: 64-lane BC")
    rng = np.random.default_rng(4343)
    active = np.nonzero(g.degrees())[0]
    srcs = np.ascontiguousarray(active[rng.integers(0, len(active), 64)], np.uint32)
    mbc = DeviceBuffer(64 * 8 * N)
    check(lib().tg_bc_multi_dev(G.handle, ptr(srcs), 64, mbc.ptr))
    lane = np.empty(N, np.float64)
    for j in (0, 21, 42, 63):
        check(lib().tg_memcpy_d2h(ptr(lane), mbc.offset(8 * N * j), 8 * N))
        ref = oracle.bc(g, int(srcs[j]))
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(lane), fin), f"lane {j}: unreachable set"
        assert oracle.rel_l1(lane[fin], ref[fin]) <= FP_TOL, f"lane {j}"
    mbc.close()


def test_coalesced_small_batches_full_size(batching):
    G = batching["G"]
    say("batching workload: 512 coalesced small batches")
    rng = np.random.default_rng(4444)
    keys = batching["keys"]
    sizes = np.round(10 ** rng.uniform(0, 3, 512)).astype(np.int64)
    n_ins = np.maximum(1, (sizes * 9) // 10)
    n_del = sizes - n_ins
    ins_s, ins_d = oracle.rmat_edges(N, oracle.aspen_ith_rand(9000) & 0xFFFFFFFF, *PROBS, int(n_ins.sum()))
    pick = keys[rng.integers(0, len(keys), int(n_del.sum()))]
    del_s, del_d = split(pick)
    src, dst, ops = [], [], []
    io = do = 0
    for b in range(512):
        src += [ins_s[io:io + n_ins[b]], del_s[do:do + n_del[b]]]
        dst += [ins_d[io:io + n_ins[b]], del_d[do:do + n_del[b]]]
        ops += [np.zeros(n_ins[b], np.uint32), np.ones(n_del[b], np.uint32)]
        io += n_ins[b]
        do += n_del[b]
    # 64 batches per launch, in arrival order
    per = []
    for c in range(0, 512, 64):
        lo, hi = 2 * c, 2 * (c + 64)
        per.append((np.concatenate(src[lo:hi]), np.concatenate(dst[lo:hi]), np.concatenate(ops[lo:hi])))
    for s_, d_, o_ in per:
        G.apply_ops(s_, d_, o_)
    # the set model: the last op of each edge in arrival order decides
    allk = np.concatenate([(s_.astype(np.uint64) << np.uint64(32)) | d_ for s_, d_, _ in per])
    allo = np.concatenate([o_ for _, _, o_ in per])
    order = np.argsort(allk, kind="stable")
    k_sorted, o_sorted = allk[order], allo[order]
    last = np.ones(len(k_sorted), bool)
    last[:-1] = k_sorted[1:] != k_sorted[:-1]
    ins_final = k_sorted[last & (o_sorted == 0)]
    del_final = k_sorted[last & (o_sorted == 1)]
    want = oracle.set_difference(oracle.set_union(keys, ins_final), del_final)
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(want, N)
    assert np.array_equal(off, ref.off)
    assert np.array_equal(nbr, ref.nbr[:ref.m])
