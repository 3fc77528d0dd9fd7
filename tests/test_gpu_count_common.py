"""GPU tests of Graph::count_common (include/graph.h:901-1255) on the device,
of TC.h's edgeMap + countF form (TC.h:51-87) and of the graph-less
VertexSubset constructors (include/VertexSubset.cpp:109-128), through the C
ABI, against the oracle (count_common pinned by brute-force intersection in
tests/test_oracle.py).  Integer results bit-exact."""
import ctypes

import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import TerraceError

pytestmark = pytest.mark.gpu
TG_EINVAL = -1


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def all_tier_graph(n=1 << 14, m0=300_000, seed=17, holes=True):
    """Skewed rMAT with all three tiers; then a delete batch of random
    existing edges (both directions) leaves empty PMA slots and shrunken
    tree blocks for the walks to skip."""
    keys = oracle.build_keys(*oracle.rmat_edges(n, seed, 0.57, 0.19, 0.19, m0))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    if holes:
        rng = np.random.default_rng(seed)
        pick = keys[rng.choice(len(keys), len(keys) // 10, replace=False)]
        s, d = split(pick)
        dels = oracle.sort_unique(np.concatenate([pick, (d.astype(np.uint64) << np.uint64(32)) | s]))
        G.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(keys, dels)
    tiers = G.tiers()
    assert (tiers == 1).any() and (tiers == 2).any()
    return G, keys, tiers


def test_count_common_all_tier_pairs():
    n = 1 << 14
    say("count_common tier pairs")
    G, keys, tiers = all_tier_graph(n)
    g = oracle.CSR(keys, n)
    deg = g.degrees()
    rng = np.random.default_rng(3)
    by_tier = [np.nonzero((tiers == 0) & (deg > 0))[0], np.nonzero(tiers == 1)[0], np.nonzero(tiers == 2)[0]]
    src, dst = [], []
    for a in range(3):
        for b in range(3):
            src.append(by_tier[a][rng.integers(0, len(by_tier[a]), 300)])
            dst.append(by_tier[b][rng.integers(0, len(by_tier[b]), 300)])
    # adjacent pairs (the pairs TC asks about), equal endpoints, isolated vertices
    e = rng.integers(0, len(keys), 3000)
    es, ed = split(keys[e])
    src += [es, np.arange(0, n, 97)]
    dst += [ed, np.arange(0, n, 97)]
    iso = np.nonzero(deg == 0)[0]
    if len(iso):
        src.append(iso[:50])
        dst.append(by_tier[2][rng.integers(0, len(by_tier[2]), min(50, len(iso)))])
    s = np.concatenate(src).astype(np.uint32)
    d = np.concatenate(dst).astype(np.uint32)
    got = G.count_common_batch(s, d)
    want = np.array([oracle.count_common(g, int(a), int(b)) for a, b in zip(s, d)], np.uint32)
    assert np.array_equal(got, want), np.nonzero(got != want)[0][:10]
    assert want.max() > 0
    # the single-pair form and symmetry
    i = int(np.argmax(want))
    assert G.count_common(int(s[i]), int(d[i])) == int(want[i]) == G.count_common(int(d[i]), int(s[i]))
    with pytest.raises(TerraceError) as ex:
        G.count_common_batch([0, n], [1, 2])
    assert ex.value.code == TG_EINVAL
    G.close()


def edge_map_tc(G, vs):
    cnt = tg.DeviceBuffer(8)
    cnt.upload(np.zeros(1, np.uint64))
    args = tg.OpArgs(count=cnt.ptr)
    assert tg.edgeMap(G, vs, tg.OP_TC, args, output=False) is None
    v = int(cnt.download(np.uint64, 1)[0])
    cnt.close()
    return v


def test_tc_as_edge_map_with_count_f():
    """TC.h: edgeMap(G, VertexSubset(0, n, true), countF, false) equals the
    oracle's TC; over a partial frontier it is the sum of count_common(u, v)
    over the frontier's edges with v < u."""
    n = 1 << 14
    say("TC edgeMap countF")
    G, keys, tiers = all_tier_graph(n)
    g = oracle.CSR(keys, n)
    assert edge_map_tc(G, tg.VertexSubset(None, 0, all=True, max_el=n)) == oracle.tc(g)
    assert tg.TC(G) == oracle.tc(g)
    rng = np.random.default_rng(5)
    members = np.unique(np.concatenate([rng.integers(0, n, 200), np.nonzero(tiers == 2)[0][:20]]))
    want = 0
    for u in members:
        for v in g.nbr[g.off[u]:g.off[u + 1]]:
            if v < u:
                want += oracle.count_common(g, int(u), int(v))
    assert edge_map_tc(G, tg.VertexSubset(G, members)) == want
    G.close()


def test_graphless_vertex_subsets():
    """VertexSubset(e, n) and VertexSubset(bool *els, n) carry no graph; the
    first edgeMap binds them.  BFS levels from one source and from a dense
    set of sources match the oracle; a bound subset is refused by another
    graph."""
    n = 1 << 12
    say("graph-less vertex subsets")
    keys = oracle.build_keys(*oracle.rmat_edges(n, 9, 0.5, 0.1, 0.1, 30000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    g = oracle.CSR(keys, n)
    for sources in ([int(np.argmax(g.degrees()))], [3, 40, 1000, 2047]):
        lv = np.minimum.reduce([oracle.bfs_levels(g, v) for v in sources])  # distance to the set
        parent = tg.DeviceBuffer(n * 4)
        init = np.full(n, -1, np.int32)
        init[sources] = sources
        parent.upload(init)
        if len(sources) == 1:
            vs = tg.VertexSubset(None, sources[0], max_el=n)
        else:
            els = np.zeros(n, np.uint8)
            els[sources] = 1
            vs = tg.VertexSubset.dense(els)
        assert vs.to_array().tolist() == sorted(sources)
        depth = 0
        while vs.not_empty():
            depth += 1
            nxt = tg.edgeMap(G, vs, tg.OPS["BFS"], tg.OpArgs(parent=parent.ptr, depth=depth), True)
            assert nxt.to_array().tolist() == np.nonzero(lv == depth)[0].tolist(), (sources, depth)
            vs = nxt
        parent.close()
    H = tg.Graph(n)
    vs = tg.VertexSubset(None, 1, max_el=n)
    parent = tg.DeviceBuffer(n * 4)
    parent.upload(np.full(n, -1, np.int32))
    tg.edgeMap(G, vs, tg.OPS["BFS"], tg.OpArgs(parent=parent.ptr, depth=1), True)
    with pytest.raises(TerraceError) as ex:
        tg.edgeMap(H, vs, tg.OPS["BFS"], tg.OpArgs(parent=parent.ptr, depth=1), True)
    assert ex.value.code == TG_EINVAL
    with pytest.raises(TerraceError):  # vertex count differs
        tg.edgeMap(tg.Graph(n // 2), tg.VertexSubset(None, 1, max_el=n), tg.OPS["BFS"],
                   tg.OpArgs(parent=parent.ptr, depth=1), True)
    parent.close()
    G.close()
    H.close()
