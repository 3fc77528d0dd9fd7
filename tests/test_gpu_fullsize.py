"""GPU tests at the benchmark workloads' full sizes (rMAT scale 22).

The graphs come from the device generator (bit-exact with the oracle's rMAT
at smaller sizes, tests/test_gpu_parity.py) and are pinned at full size by
the counts measured from the reference's generator: directed edges, maximum
degree, one vertex's degree and the vertices per tier.  Updates are checked
against the set model, and BFS, PageRank and CC against the oracle."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
FP_TOL = 1e-9
UNREACHED = np.iinfo(np.uint32).max


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def sym(keys):
    keys = np.asarray(keys, np.uint64)
    rev = ((keys & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (keys >> np.uint64(32))
    return oracle.sort_unique(np.concatenate([keys, rev]))


def device_graph(n, probs, m0):
    """Static seed ith_rand(100), nominal generator, symmetrized, no
    self-loops: the device buffers and the sorted unique keys."""
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    s, d, w = tg.rmat_device(n, seed, *probs, m0, nominal=True, symmetric=True, drop_loops=True)
    src, dst = s.download(np.uint32, w), d.download(np.uint32, w)
    keys = oracle.sort_unique((src.astype(np.uint64) << np.uint64(32)) | dst)
    return s, d, w, keys


def assert_adjacency(G, keys, n, what):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off[1:] - off[:-1], ref.degrees().astype(off.dtype)), what
    assert np.array_equal(nbr, ref.nbr[:ref.m]), what
    assert G.stats()["pma_elements"] == G.tier_counts()["pma_edges"], what
    return ref


def check_queries(G, g, src, what):
    parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
    lv = oracle.bfs_levels(g, src)
    assert np.array_equal(levels, lv), f"{what}: BFS levels"
    assert oracle.bfs_check_parents(g, src, parents, lv) == 0, f"{what}: BFS parents"
    assert oracle.rel_l1(tg.PR_S(G, 10), oracle.pagerank(g, 10)) <= FP_TOL, f"{what}: PR"
    assert np.array_equal(tg.CC(G), oracle.cc(g)), f"{what}: CC"


def test_plain_workload_full_size():
    """rMAT scale 22, 69M generated edges (a=.5 b=.1 c=.1): 135,910,260
    directed edges, maximum degree 1,441, deg(9) = 707, 17 B-tree vertices.
    Two mixed batches (about 9e4 inserts generated with ith_rand(1000 + t),
    1e4 deletes of existing edges in both directions) match the set model;
    BFS from 9, PageRank 10 iterations and CC match the oracle."""
    n = 1 << 22
    say("full size plain workload: build")
    s, d, w, keys = device_graph(n, (0.5, 0.1, 0.1), 69_000_000)
    assert len(keys) == 135_910_260
    G = tg.Graph(n)
    G.add_edge_batch_dev(s, d, w)
    del s, d
    deg = G.degrees()
    assert int(deg.max()) == 1441 and int(deg[9]) == 707
    assert int((G.tiers() == 2).sum()) == 17
    assert_adjacency(G, keys, n, "static")
    rng = np.random.default_rng(5)
    for t in range(2):
        say("full size plain workload: batch", t)
        seed = oracle.aspen_ith_rand(1000 + t) & 0xFFFFFFFF
        ins = oracle.build_keys(*oracle.rmat_edges(n, seed, 0.5, 0.1, 0.1, 45000))
        dels = sym(keys[rng.choice(len(keys), 5000, replace=False)])
        G.add_edge_batch(*split(ins))
        keys = oracle.set_union(keys, ins)
        G.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(keys, dels)
        ref = assert_adjacency(G, keys, n, f"batch {t}")
    say("full size plain workload: queries")
    check_queries(G, ref, 9, "plain full size")
    G.close()


def test_skew_workload_full_size():
    """Skewed rMAT scale 22, 120M generated edges (a=.57 b=.19 c=.19):
    224,914,926 directed edges, maximum degree 243,995; vertices with edges
    by tier: 1,685,167 inline only, 989,920 PMA, 35,443 B-tree.  A 1e6-edge
    batch whose sources are the top-1% degree vertices, inserted and then
    deleted, leaves the adjacency as it was (round trip); BFS from the
    largest hub, PageRank 10 iterations and CC match the oracle."""
    n = 1 << 22
    say("full size skew workload: build")
    s, d, w, keys = device_graph(n, (0.57, 0.19, 0.19), 120_000_000)
    assert len(keys) == 224_914_926
    G = tg.Graph(n)
    G.add_edge_batch_dev(s, d, w)
    del s, d
    deg = G.degrees()
    assert int(deg.max()) == 243_995
    tiers = G.tiers()
    assert int(((tiers == 0) & (deg > 0)).sum()) == 1_685_167
    assert int((tiers == 1).sum()) == 989_920
    assert int((tiers == 2).sum()) == 35_443
    ref = assert_adjacency(G, keys, n, "static")
    say("full size skew workload: hub round trip")
    rng = np.random.default_rng(9)
    top = np.argsort(deg)[-(n // 100):]
    active = np.nonzero(deg)[0]
    bs = top[rng.integers(0, len(top), 1_000_000)].astype(np.uint64)
    bd = active[rng.integers(0, len(active), 1_000_000)].astype(np.uint64)
    batch = oracle.sort_unique((bs << np.uint64(32)) | bd)
    batch = batch[(batch >> np.uint64(32)) != (batch & np.uint64(0xFFFFFFFF))]
    fresh = oracle.set_difference(batch, keys)  # deleting only these restores the graph
    G.add_edge_batch(*split(batch))
    assert_adjacency(G, oracle.set_union(keys, batch), n, "hub insert")
    G.remove_edge_batch(*split(fresh))
    assert_adjacency(G, keys, n, "hub round trip")
    say("full size skew workload: queries")
    check_queries(G, ref, int(np.argmax(deg)), "skew full size")
    G.close()
