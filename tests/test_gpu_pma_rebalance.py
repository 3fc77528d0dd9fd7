"""GPU tests of the PMA's density maintenance under deletes
(src/PMA.cc:1520-1746 and half_list, PMA.cc:722-813): a leaf left at or
below its lower density bound (0.125 of its 64 slots) climbs to the lowest
window above that level's lower bound, which is respread evenly; an array
whose root falls to its own lower bound (0.25) shrinks.

* Deleting 90% of an all-tier scale-16 graph in ten batches (plain delete
  batches and mixed update batches) shrinks the PMA, leaves no leaf at or
  below 8 elements after any batch, keeps the adjacency equal to the set
  model, and BFS / PageRank / CC on what remains equal the oracle.
* Deleting the edges of a source range empties a block of leaves while the
  array as a whole stays dense: the underfull windows are respread, the
  array keeps its size.
* Deleting every edge shrinks the PMA to its smallest size (two leaves),
  and the emptied graph takes the edges back."""
import time

import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
FP_TOL = 1e-9
LEAF_LOW = 8  # floor(0.125 * 64): a leaf at or below this count is underfull


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def sym(keys):
    s, d = split(keys)
    both = np.concatenate([keys, (d.astype(np.uint64) << np.uint64(32)) | s])
    return oracle.sort_unique(both)


def assert_set_model(G, keys, n, what):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off, ref.off), what
    assert np.array_equal(nbr, ref.nbr[:ref.m]), what
    assert G.stats()["pma_elements"] == G.tier_counts()["pma_edges"], what


def assert_leaves_above_bound(G, what):
    cnt = G.pma_leaf_counts()
    if len(cnt) > 2:  # the smallest array may hold fewer elements
        assert (cnt >= LEAF_LOW).all(), f"{what}: leaf at {cnt.min()}"
    assert int(cnt.sum()) == G.stats()["pma_elements"], what


def build(n, m, seed):
    keys = sym(oracle.build_keys(*oracle.rmat_edges(n, seed, 0.57, 0.19, 0.19, m)))
    keys = keys[(keys >> np.uint64(32)) != (keys & np.uint64(0xFFFFFFFF))]
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    return G, keys


def time_pr(G, iters=10, reps=3):
    tg.PR_S(G, iters)
    t0 = time.perf_counter()
    for _ in range(reps):
        tg.PR_S(G, iters)
    return (time.perf_counter() - t0) / reps


def test_delete_ninety_percent_shrinks_and_respreads():
    n = 1 << 16
    say("PMA rebalance: scale-16 all-tier graph, delete 90% in ten batches")
    G, keys = build(n, 1_200_000, 16)
    tiers = G.tiers()
    assert (tiers == 0).any() and (tiers == 1).any() and (tiers == 2).any()
    assert_leaves_above_bound(G, "after the build")
    slots0 = G.stats()["pma_slots"]
    pr_before = time_pr(G)
    rng = np.random.default_rng(90)
    # undirected edges (s < d) to delete, both directions each
    und = keys[(keys >> np.uint64(32)) < (keys & np.uint64(0xFFFFFFFF))]
    doomed = und[rng.permutation(len(und))[: int(0.9 * len(und))]]
    chunks = np.array_split(doomed, 10)
    for b, chunk in enumerate(chunks):
        dels = sym(chunk)
        if b % 2:
            G.remove_edge_batch(*split(dels))
        else:
            # a mixed batch: the deletes and a few inserts of edges already present
            ins = keys[rng.integers(0, len(keys), 200)]
            ins = sym(ins)
            ins = oracle.set_difference(ins, dels)
            G.update_batch(*split(ins), *split(dels))
        keys = oracle.set_difference(keys, dels)
        assert_leaves_above_bound(G, f"batch {b}")
        if b in (0, 4, 9):
            assert_set_model(G, keys, n, f"batch {b}")
    st = G.stats()
    say("PMA slots", slots0, "->", st["pma_slots"], "shrinks", st["pma_shrinks"], "low windows",
        st["pma_low_windows"])
    assert st["pma_shrinks"] >= 1
    assert st["pma_slots"] * 4 <= slots0
    g = oracle.CSR(keys, n)
    active = np.nonzero(g.degrees())[0]
    for src in (int(active[0]), int(active[len(active) // 2])):
        parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
        lv = oracle.bfs_levels(g, src)
        assert np.array_equal(levels, lv), "BFS levels"
        assert oracle.bfs_check_parents(g, src, parents, lv) == 0, "BFS parents"
    assert oracle.rel_l1(tg.PR_S(G, 10), oracle.pagerank(g, 10)) <= FP_TOL, "PR"
    assert np.array_equal(tg.CC(G), oracle.cc(g)), "CC"
    pr_after = time_pr(G)
    say(f"PageRank 10 iterations: {pr_before * 1e3:.2f} ms before, {pr_after * 1e3:.2f} ms after the deletes")
    G.close()


def test_source_range_delete_respreads_windows():
    n = 1 << 15
    say("PMA rebalance: delete every edge of a run of sources")
    G, keys = build(n, 600_000, 15)
    st0 = G.stats()
    src = (keys >> np.uint64(32)).astype(np.int64)
    deg = np.bincount(src, minlength=n)
    tiers = G.tiers()
    # PMA-tier sources from n/4 on (13 edges inline, the rest in the PMA):
    # a prefix whose PMA elements empty whole leaves but keep the array
    # above its root lower bound (a quarter of its slots)
    cand = np.nonzero((tiers == 1) & (np.arange(n) >= n // 4))[0]
    in_pma = np.cumsum(deg[cand] - 13)
    room = (st0["pma_elements"] - st0["pma_slots"] // 4) // 2
    take = int(np.searchsorted(in_pma, room, side="right"))
    assert take > 0 and in_pma[take - 1] >= 64 * 16, (room, in_pma[:4])
    dels = keys[np.isin(src, cand[:take])]
    G.remove_edge_batch(*split(dels))
    keys = oracle.set_difference(keys, dels)
    st = G.stats()
    assert st["pma_elements"] == st0["pma_elements"] - int(in_pma[take - 1])
    assert st["pma_low_windows"] > st0["pma_low_windows"]
    assert st["pma_shrinks"] == st0["pma_shrinks"] and st["pma_slots"] == st0["pma_slots"]
    assert_leaves_above_bound(G, "after the range delete")
    assert_set_model(G, keys, n, "after the range delete")
    G.close()


def test_delete_everything_then_reinsert():
    n = 1 << 14
    say("PMA rebalance: delete every edge, insert them again")
    G, keys = build(n, 150_000, 14)
    assert G.stats()["pma_slots"] > 128
    G.remove_edge_batch(*split(keys))
    st = G.stats()
    assert st["pma_elements"] == 0 and st["pma_slots"] == 128 and G.get_num_edges() == 0
    assert_set_model(G, np.zeros(0, np.uint64), n, "empty")
    G.add_edge_batch(*split(keys))
    assert_set_model(G, keys, n, "inserted again")
    assert_leaves_above_bound(G, "inserted again")
    G.close()
