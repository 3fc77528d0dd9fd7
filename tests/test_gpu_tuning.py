"""GPU tests of the tuning and batching entry points: the autotuner and the
HIP-graph PageRank replay leave every result unchanged, and coalesced update
ops equal applying the same ops one at a time."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_autotune_keeps_results():
    n = 1 << 13
    print("[gpu-test] autotune", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 61, 0.57, 0.19, 0.19, 120000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    g = oracle.CSR(keys, n)
    src = int(np.argmax(g.degrees()))
    _, lv0 = tg.BFS_with_edge_map(G, src, with_levels=True)
    cc0, pr0, bc0 = tg.CC(G), tg.PR_S(G, 10), tg.BC(G, src)
    ms0 = tg.bfs_multi(G, np.arange(0, n, n // 64, dtype=np.uint32)[:64])
    rep = G.autotune()
    tun = G.get_tuning()
    assert tun["thr_bfs"] == rep["thr_bfs"] and tun["pull_block"] == rep["pull_block"]
    assert rep["pull_block"] in (128, 256, 512)
    for k in ("thr_bfs", "thr_ms", "thr_cc"):
        assert rep[k] in (5, 10, 20, 40, 80, 160, 320)
    _, lv1 = tg.BFS_with_edge_map(G, src, with_levels=True)
    assert np.array_equal(lv0, lv1) and np.array_equal(lv1, oracle.bfs_levels(g, src))
    assert np.array_equal(tg.CC(G), cc0)
    assert np.array_equal(tg.PR_S(G, 10), pr0)  # same sum order in every block size
    np.testing.assert_allclose(tg.BC(G, src), bc0, rtol=1e-9)
    assert np.array_equal(tg.bfs_multi(G, np.arange(0, n, n // 64, dtype=np.uint32)[:64]), ms0)
    assert tun["push_wave"] in (4096, 16384, 65536)
    # forced extremes of the threshold and of the PMA push mapping (always
    # 16 lanes, always a wave) give the same levels and labels
    for thr in (1, 100000):
        for pw in (0, 1 << 30):
            G.set_tuning(thr_bfs=thr, thr_ms=thr, thr_cc=thr, push_wave=pw)
            _, lv = tg.BFS_with_edge_map(G, src, with_levels=True)
            assert np.array_equal(lv, lv0)
            assert np.array_equal(tg.CC(G), cc0)
    with pytest.raises(Exception):
        G.set_tuning(pull_block=96)
    G.close()


def test_pagerank_graph_replay_matches_eager():
    """PageRank iterations replayed from a captured HIP graph equal the eager
    iterations, on several calls with eager work and updates in between."""
    n = 1 << 12
    print("[gpu-test] pagerank graph replay", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 67, 0.57, 0.19, 0.19, 60000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    rng = np.random.default_rng(3)
    for rnd in range(3):
        for iters in (1, 2, 5, 10):
            G.set_tuning(pr_graph=0)
            eager = tg.PR_S(G, iters)
            G.set_tuning(pr_graph=1)
            replay = tg.PR_S(G, iters)
            assert np.array_equal(eager, replay), (rnd, iters)
            g = oracle.CSR(keys, n)
            assert oracle.rel_l1(replay, oracle.pagerank(g, iters)) <= 1e-9
        tg.CC(G)  # eager work between the replays
        fresh = oracle.build_keys(*oracle.rmat_edges(n, 700 + rnd, 0.57, 0.19, 0.19, 2000))
        G.add_edge_batch(*split(fresh))
        keys = oracle.set_union(keys, fresh)
    G.close()


def test_pagerank_sliced_matches_tiered_pull():
    """The XCD-sliced PageRank (slice partials added in slice order) and the
    tiered pull agree within the FP64 bar, with the oracle too, through
    updates (the slices are rebuilt when the graph changes)."""
    n = 1 << 13
    print("[gpu-test] pagerank sliced", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 89, 0.57, 0.19, 0.19, 100000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    for rnd in range(2):
        g = oracle.CSR(keys, n)
        want = oracle.pagerank(g, 10)
        G.set_tuning(pr_sliced=1)
        a = tg.PR_S(G, 10)
        assert np.array_equal(a, tg.PR_S(G, 10))  # deterministic, reused slices
        G.set_tuning(pr_sliced=0)
        b = tg.PR_S(G, 10)
        assert oracle.rel_l1(a, want) <= 1e-9 and oracle.rel_l1(b, want) <= 1e-9
        assert oracle.rel_l1(a, b) <= 1e-12
        G.set_tuning(pr_sliced=1, pr_graph=0)
        assert np.array_equal(tg.PR_S(G, 10), a)  # eager iterations equal the replayed ones
        G.set_tuning(pr_graph=1)
        assert np.array_equal(tg.PR_S(G, 0), np.full(n, 1.0 / n))
        fresh = oracle.build_keys(*oracle.rmat_edges(n, 300 + rnd, 0.57, 0.19, 0.19, 3000))
        G.add_edge_batch(*split(fresh))
        keys = oracle.set_union(keys, fresh)
    G.close()


def test_pagerank_sliced_hub_rows_span_tiles():
    """A star with 60000 leaves (the hub's row in every slice spans several
    2048-step merge-path tiles, so its sum is carried across tiles), a
    path and isolated vertices (empty rows in every slice): the sliced and
    tiered PageRank agree with the oracle."""
    n = 1 << 17
    print("[gpu-test] pagerank sliced hub", flush=True)
    hub = 5
    leaves = np.arange(1000, 61000, dtype=np.uint64)
    path = np.arange(70000, 70500, dtype=np.uint64)
    src = np.concatenate([np.full(leaves.size, hub, np.uint64), leaves, path[:-1], path[1:]])
    dst = np.concatenate([leaves, np.full(leaves.size, hub, np.uint64), path[1:], path[:-1]])
    keys = np.unique((src << np.uint64(32)) | dst)
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    want = oracle.pagerank(oracle.CSR(keys, n), 7)
    G.set_tuning(pr_sliced=1)
    a = tg.PR_S(G, 7)
    G.set_tuning(pr_sliced=0)
    b = tg.PR_S(G, 7)
    assert oracle.rel_l1(a, want) <= 1e-12 and oracle.rel_l1(b, want) <= 1e-12
    assert np.all(a[71000:] == 0.0) and np.all(b[71000:] == 0.0)  # isolated after one iteration
    G.close()


def test_apply_ops_equals_sequential():
    n = 1 << 12
    print("[gpu-test] apply_ops", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 71, 0.5, 0.1, 0.1, 30000))
    G1, G2 = tg.Graph(n), tg.Graph(n)
    G1.add_edge_batch(*split(keys))
    G2.add_edge_batch(*split(keys))
    rng = np.random.default_rng(9)
    exist = keys[rng.choice(len(keys), 400, replace=False)]
    fresh = oracle.build_keys(*oracle.rmat_edges(n, 73, 0.5, 0.1, 0.1, 400), symmetric=False)
    pool = np.concatenate([exist, fresh])
    batches = []
    for b in range(20):
        k = pool[rng.integers(0, len(pool), int(rng.integers(1, 200)))]
        ops = rng.integers(0, 2, len(k)).astype(np.uint32)
        batches.append((k, ops))
    # sequential: each op on its own, in order (oracle and G2)
    state = keys
    for k, ops in batches:
        for key, op in zip(k, ops):
            s, d = split(np.array([key], np.uint64))
            if op == 0:
                G2.add_edge_batch(s, d)
                state = oracle.set_union(state, np.array([key], np.uint64))
            else:
                G2.remove_edge_batch(s, d)
                state = oracle.set_difference(state, np.array([key], np.uint64))
    # coalesced: all 20 batches in one pass
    allk = np.concatenate([k for k, _ in batches])
    allo = np.concatenate([o for _, o in batches])
    G1.apply_ops(*split(allk), allo)
    ref = oracle.CSR(state, n)
    for G in (G1, G2):
        off, nbr = G.export_adjacency()
        assert np.array_equal(off, ref.off) and np.array_equal(nbr, ref.nbr[:ref.m])
    with pytest.raises(Exception):
        G1.apply_ops(np.array([1], np.uint32), np.array([2], np.uint32), np.array([2], np.uint32))
    G1.close(), G2.close()
