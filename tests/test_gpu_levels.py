"""GPU tests of BFS, CC and BC forward levels run on device-resident
counts (each level's direction chosen on the device, include/Map.cpp:66-74)
and of the captured HIP graphs that replay them and PageRank's iterations.

* Captured and one-by-one launches give the same BFS levels and CC labels,
  equal to the oracle, at thresholds that force all-push, all-pull and
  mixed directions (a push after a pull reads the frontier's bitmap); BC
  within 1e-9 of the oracle both ways (its pushes add FP64 values with
  atomics, so runs agree to rounding, not bit for bit).
* A captured graph replays several times, with other work and host
  synchronisation between the replays, and every replay equals the eager
  result; after an update batch the next call captures again and follows
  the new graph (the oracle on the new key set).
* PageRank's cached graphs (tiered pull and sliced) give the same ranks on
  every call and follow an update."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
FP_TOL = 1e-9


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def sym(keys):
    s, d = split(keys)
    both = np.concatenate([keys, (d.astype(np.uint64) << np.uint64(32)) | s])
    both = oracle.sort_unique(both)
    return both[(both >> np.uint64(32)) != (both & np.uint64(0xFFFFFFFF))]


@pytest.fixture(scope="module")
def graph():
    n = 1 << 15
    keys = sym(oracle.build_keys(*oracle.rmat_edges(n, 151, 0.57, 0.19, 0.19, 500_000)))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    tiers = G.tiers()
    assert (tiers == 0).any() and (tiers == 1).any() and (tiers == 2).any()
    st = {"G": G, "keys": keys, "n": n}
    yield st
    G.close()


def check_bc(got, ref, what):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin), f"{what}: unreachable set"
    assert oracle.rel_l1(got[fin], ref[fin]) <= FP_TOL, what


def check_bfs(G, g, src, what):
    parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
    lv = oracle.bfs_levels(g, src)
    assert np.array_equal(levels, lv), f"{what}: BFS levels"
    assert oracle.bfs_check_parents(g, src, parents, lv) == 0, f"{what}: BFS parents"
    return levels


@pytest.mark.parametrize("thr", [1, 20, 1 << 20])
def test_levels_captured_equal_eager(graph, thr):
    G, n = graph["G"], graph["n"]
    g = oracle.CSR(graph["keys"], n)
    active = np.nonzero(g.degrees())[0]
    srcs = [int(active[0]), int(active[len(active) // 3]), int(np.argmax(g.degrees()))]
    say("levels: threshold", thr)
    G.set_tuning(thr_bfs=thr, thr_cc=thr, level_graph=0)
    eager = [check_bfs(G, g, s, f"eager thr {thr}") for s in srcs]
    cc_eager = tg.CC(G)
    assert np.array_equal(cc_eager, oracle.cc(g)), "eager CC"
    bc_ref = {s: oracle.bc(g, s) for s in srcs[1:]}
    for s in srcs[1:]:
        check_bc(tg.BC(G, s), bc_ref[s], f"eager BC thr {thr}")
    G.set_tuning(level_graph=1)
    for rep in range(3):
        for s, lv in zip(srcs, eager):
            assert np.array_equal(check_bfs(G, g, s, f"captured thr {thr} rep {rep}"), lv)
        assert np.array_equal(tg.CC(G), cc_eager), f"captured CC rep {rep}"
        for s in srcs[1:]:
            check_bc(tg.BC(G, s), bc_ref[s], f"captured BC thr {thr} rep {rep}")
        G.export_adjacency()  # other work and a host synchronisation between replays
    G.set_tuning(thr_bfs=40, thr_cc=5, level_graph=1)


def test_levels_follow_updates(graph):
    G, n = graph["G"], graph["n"]
    rng = np.random.default_rng(77)
    keys = graph["keys"]
    src = int(np.nonzero(oracle.CSR(keys, n).degrees())[0][5])
    for rnd in range(3):
        say("levels after update", rnd)
        und = keys[(keys >> np.uint64(32)) < (keys & np.uint64(0xFFFFFFFF))]
        dels = sym(und[rng.integers(0, len(und), 20_000)])
        ins = sym(oracle.build_keys(*oracle.rmat_edges(n, 900 + rnd, 0.57, 0.19, 0.19, 20_000)))
        ins = oracle.set_difference(ins, dels)
        G.update_batch(*split(ins), *split(dels))
        keys = oracle.set_difference(oracle.set_union(keys, ins), dels)
        g = oracle.CSR(keys, n)
        check_bfs(G, g, src, f"round {rnd}")
        assert np.array_equal(tg.CC(G), oracle.cc(g)), f"round {rnd}: CC"
        check_bc(tg.BC(G, src), oracle.bc(g, src), f"round {rnd}: BC")
        pr = tg.PR_S(G, 10)
        assert oracle.rel_l1(pr, oracle.pagerank(g, 10)) <= FP_TOL, f"round {rnd}: PR"
        assert np.array_equal(tg.PR_S(G, 10), pr), f"round {rnd}: PR replay"
    graph["keys"] = keys


@pytest.mark.parametrize("sliced", [0, 1])
def test_pagerank_cached_graph(graph, sliced):
    G, n = graph["G"], graph["n"]
    g = oracle.CSR(graph["keys"], n)
    say("PageRank cached graph, sliced", sliced)
    G.set_tuning(pr_sliced=sliced, pr_graph=1)
    ref = oracle.pagerank(g, 10)
    first = tg.PR_S(G, 10)
    assert oracle.rel_l1(first, ref) <= FP_TOL
    for rep in range(3):
        G.export_adjacency()
        assert np.array_equal(tg.PR_S(G, 10), first), f"replay {rep}"
        assert oracle.rel_l1(tg.PR_S(G, 7), oracle.pagerank(g, 7)) <= FP_TOL, f"7 iterations, replay {rep}"
    G.set_tuning(pr_graph=0)
    assert oracle.rel_l1(tg.PR_S(G, 10), ref) <= FP_TOL
    G.set_tuning(pr_sliced=1, pr_graph=1)
