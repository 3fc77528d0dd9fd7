"""Pins for the CPU oracle (oracle/oracle.c).

Each oracle function is checked against something that is not itself:
  * golden vectors written by the reference build (tests/golden/*.npz),
  * the generator's known-answer values recorded from the reference,
  * closed forms and brute force on tiny graphs,
  * a planted mistake in a copy of the oracle's output, which must fail.
"""
import itertools
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GRAPHS = ["graph_s10", "graph_skew12", "graph_promote11"]


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def keys_of(pairs):
    return [(s << 32) | d for s, d in pairs]


# --------------------------------------------------------------- generator
def test_aspen_known_answer():
    # random_aspen().ith_rand(100) (rmat_util.h:57-70), recorded from the
    # reference build
    assert oracle.aspen_ith_rand(100) == 98155223774525950


def test_rmat_known_answer():
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    s, d = oracle.rmat_edges(1 << 20, seed, 0.5, 0.1, 0.1, 2, nominal=False)
    assert list(zip(s.tolist(), d.tolist())) == [(208863, 212575), (907695, 907630)]


def test_rmat_matches_reference_generator():
    z = load("rmat_kat")
    k = 0
    while f"pairs{k}" in z:
        n, seed, count = (int(x) for x in z[f"params{k}"])
        a, b, c = (float(x) for x in z[f"probs{k}"])
        s, d = oracle.rmat_edges(n, seed, a, b, c, count, nominal=False)
        ref = z[f"pairs{k}"]
        assert np.array_equal(s, ref[:, 0]) and np.array_equal(d, ref[:, 1]), k
        k += 1
    assert k >= 4


def test_rmat_nominal_quadrants():
    # with r = hash32/2^32 the top-level quadrant frequencies follow (a,b,c,d)
    n = 1 << 20
    s, d = oracle.rmat_edges(n, 5, 0.5, 0.1, 0.1, 200000, nominal=True)
    half = n // 2
    q = (s >= half).astype(int) * 2 + (d >= half).astype(int)
    freq = np.bincount(q, minlength=4) / len(q)
    assert np.allclose(freq, [0.5, 0.1, 0.1, 0.3], atol=0.01)


def test_rmat_rejects_bad_params():
    with pytest.raises(ValueError):
        oracle.rmat_edges(1000, 1, 0.5, 0.1, 0.1, 4)
    with pytest.raises(ValueError):
        oracle.rmat_edges(1024, 1, 0.6, 0.3, 0.3, 4)


# --------------------------------------------------------------- edge sets
def test_sort_unique_and_sets_vs_numpy():
    rng = np.random.default_rng(1)
    a = rng.integers(0, 1 << 40, 5000, dtype=np.uint64)
    a = np.concatenate([a, a[:100]])
    b = rng.integers(0, 1 << 40, 3000, dtype=np.uint64)
    ua, ub = np.unique(a), np.unique(b)
    assert np.array_equal(oracle.sort_unique(a), ua)
    assert np.array_equal(oracle.set_union(ua, ub), np.union1d(ua, ub))
    assert np.array_equal(oracle.set_difference(ua, ub), np.setdiff1d(ua, ub))


def test_build_keys_drops_loops_and_symmetrizes():
    k = oracle.build_keys(np.array([1, 2, 3, 3], np.uint32), np.array([2, 2, 1, 1], np.uint32))
    assert k.tolist() == sorted(keys_of([(1, 2), (2, 1), (3, 1), (1, 3)]))


# --------------------------------------------------------------- golden graphs
@pytest.mark.parametrize("name", GRAPHS)
def test_adjacency_after_every_batch_matches_reference(name):
    z = load(name)
    n = int(z["n"][0])
    keys = z["base"]
    for k in range(int(z["nbatches"][0]) + 1):
        if k:
            keys = oracle.set_union(keys, z[f"batch{k - 1}"])
        g = oracle.CSR(keys, n)
        assert np.array_equal(g.degrees(), z[f"deg{k}"]), k
        assert np.array_equal(g.nbr[:g.m], z[f"nbrs{k}"]), k


def final_graph(z):
    keys = z["base"]
    for k in range(int(z["nbatches"][0])):
        keys = oracle.set_union(keys, z[f"batch{k}"])
    return oracle.CSR(keys, int(z["n"][0]))


def depths_from_parents(parents):
    n = len(parents)
    depth = np.full(n, np.iinfo(np.uint32).max, np.uint32)
    for v in range(n):
        if parents[v] < 0:
            continue
        d, u = 0, v
        while parents[u] != u:
            u = parents[u]
            d += 1
        depth[v] = d
    return depth


@pytest.mark.parametrize("name", GRAPHS)
def test_bfs_matches_reference(name):
    z = load(name)
    g = final_graph(z)
    src = int(z["src"][0])
    lv = oracle.bfs_levels(g, src)
    assert np.array_equal(lv, depths_from_parents(z["bfs_parents"]))
    assert oracle.bfs_check_parents(g, src, z["bfs_parents"], lv) == 0


@pytest.mark.parametrize("name", GRAPHS)
def test_pagerank_matches_reference(name):
    z = load(name)
    g = final_graph(z)
    assert oracle.rel_l1(oracle.pagerank(g, int(z["iters"][0])), z["pr"]) <= 1e-12


@pytest.mark.parametrize("name", GRAPHS)
def test_cc_matches_reference(name):
    z = load(name)
    assert np.array_equal(oracle.cc(final_graph(z)), z["cc"])


@pytest.mark.parametrize("name", GRAPHS)
def test_bc_matches_reference(name):
    z = load(name)
    g = final_graph(z)
    bc = oracle.bc(g, int(z["src"][0]))
    ref = z["bc"]
    lv = oracle.bfs_levels(g, int(z["src"][0]))
    reach = lv != np.iinfo(np.uint32).max
    assert np.all(np.isnan(bc[~reach]))
    assert np.all(np.isfinite(bc[reach]))
    assert oracle.rel_l1(bc[reach], ref[reach]) <= 1e-9


@pytest.mark.parametrize("name", GRAPHS)
def test_tc_matches_reference(name):
    z = load(name)
    assert oracle.tc(final_graph(z)) == int(z["tc"][0])


@pytest.mark.parametrize("name", GRAPHS)
def test_sssp_bf_matches_reference(name):
    z = load(name)
    assert np.array_equal(oracle.sssp_bf(final_graph(z), int(z["src"][0])), z["sssp"])


def test_sssp_bf_is_hop_distance():
    # unit edge lengths: the Bellman-Ford fixed point is the BFS level, and
    # unreachable vertices keep INT_MAX / 2
    n = 1 << 11
    keys = oracle.build_keys(*oracle.rmat_edges(n, 83, 0.57, 0.19, 0.19, 9000))
    g = oracle.CSR(keys, n)
    for src in (0, 5, int(np.argmax(g.degrees()))):
        lv = oracle.bfs_levels(g, src).astype(np.int64)
        want = np.where(lv == np.iinfo(np.uint32).max, (2**31 - 1) // 2, lv)
        got = oracle.sssp_bf(g, src)
        assert np.array_equal(got, want)


# --------------------------------------------------------------- closed forms
def undirected(n, edges):
    ks = []
    for a, b in edges:
        ks += [(a << 32) | b, (b << 32) | a]
    return oracle.CSR(oracle.sort_unique(np.array(ks, np.uint64)), n)


def test_tc_brute_force():
    rng = np.random.default_rng(3)
    n = 40
    edges = {(int(a), int(b)) for a, b in rng.integers(0, n, (300, 2)) if a != b}
    g = undirected(n, edges)
    adj = np.zeros((n, n), bool)
    for a, b in edges:
        adj[a, b] = adj[b, a] = True
    brute = sum(1 for a, b, c in itertools.combinations(range(n), 3)
                if adj[a, b] and adj[b, c] and adj[a, c])
    assert oracle.tc(g) == brute
    # count_common(s,d) only counts common neighbours below min(s, d)
    for s, d in list(edges)[:50]:
        want = sum(1 for w in range(min(s, d)) if adj[s, w] and adj[d, w])
        assert oracle.count_common(g, s, d) == want


def test_cc_brute_force():
    g = undirected(10, [(1, 2), (2, 3), (5, 9), (7, 8), (8, 6)])
    assert oracle.cc(g).tolist() == [0, 1, 1, 1, 4, 5, 6, 6, 6, 5]


def test_pagerank_closed_forms():
    # undamped: on a cycle every vertex keeps 1/n
    n = 12
    g = undirected(n, [(i, (i + 1) % n) for i in range(n)])
    assert np.allclose(oracle.pagerank(g, 10), 1.0 / n, rtol=0, atol=1e-15)
    # star with k leaves: mass alternates between centre and leaves
    k = 5
    g = undirected(k + 1, [(0, i) for i in range(1, k + 1)])
    p = np.full(k + 1, 1.0 / (k + 1))
    for _ in range(3):
        q = p / np.maximum((np.array([k] + [1] * k)), 1)
        p = np.array([q[1:].sum()] + [q[0]] * k)
    assert np.allclose(oracle.pagerank(g, 3), p, rtol=1e-15)
    # isolated vertices end at 0
    g = undirected(4, [(0, 1)])
    pr = oracle.pagerank(g, 4)
    assert pr[2] == 0.0 and pr[3] == 0.0


def test_bc_closed_forms():
    # path 0-1-2-3-4 from source 0: dependency of vertex k = n-1-k
    g = undirected(5, [(i, i + 1) for i in range(4)])
    assert np.allclose(oracle.bc(g, 0), [4, 3, 2, 1, 0])
    # diamond 0-{1,2}-3: each middle vertex carries half of vertex 3
    g = undirected(5, [(0, 1), (0, 2), (1, 3), (2, 3)])
    bc = oracle.bc(g, 0)
    assert np.allclose(bc[:4], [3, 0.5, 0.5, 0]) and np.isnan(bc[4])


def test_bfs_levels_closed_form():
    g = undirected(6, [(0, 1), (1, 2), (2, 3), (0, 4)])
    assert oracle.bfs_levels(g, 0).tolist() == [0, 1, 2, 3, 1, 0xFFFFFFFF]
    bad = np.array([0, 0, 1, 2, 0, -1], np.int32)
    lv = oracle.bfs_levels(g, 0)
    assert oracle.bfs_check_parents(g, 0, bad, lv) == 0
    bad[3] = 1  # not adjacent to 3
    assert oracle.bfs_check_parents(g, 0, bad, lv) == 1


# --------------------------------------------------------------- tier model
def run_tiers(n, batches):
    tiers = oracle.tier_model(n)
    keys = np.zeros(0, np.uint64)
    for b in batches:
        b = oracle.sort_unique(np.array(b, np.uint64))
        before = oracle.CSR(keys, n)
        oracle.tier_insert(tiers, before, b)
        keys = oracle.set_union(keys, b)
        assert np.array_equal(tiers["degree"], oracle.CSR(keys, n).degrees())
    return tiers


def test_tier_routing_rule():
    e = lambda s, ds: [(s << 32) | d for d in ds]  # noqa: E731
    # one batch of 14 -> inline + PMA; one batch of 1025 -> B-tree (graph.h:276)
    t = run_tiers(2000, [e(0, range(1, 15)) + e(1, range(2, 1027))])
    assert t["btree"][0] == 0 and t["btree"][1] == 1
    # D=1000 then a run of 30 -> D+size > 1024 -> B-tree (graph.h:344)
    t = run_tiers(2000, [e(0, range(1, 1001)), e(0, range(1001, 1031))])
    assert t["btree"][0] == 1
    # D=1000 then 10 -> stays PMA; duplicates count toward size (graph.h:262)
    t = run_tiers(2000, [e(0, range(1, 1001)), e(0, range(1001, 1011))])
    assert t["btree"][0] == 0
    t = run_tiers(2000, [e(0, range(1, 1001)), e(0, range(900, 1011))])
    assert t["btree"][0] == 1 and t["degree"][0] == 1010
    # run entirely inside the inline slots: no overflow, no routing
    t = run_tiers(2000, [e(0, range(1, 1001)), e(0, range(1, 5))])
    assert t["btree"][0] == 0


@pytest.mark.parametrize("name", ["graph_s10", "graph_skew12", "graph_promote11"])
def test_tier_model_matches_reference_build(name):
    """The tier model's B-tree membership after the static build and after
    every batch equals the reference build's own (vertices whose
    aux_neighbors is set), dumped by oracle/_ref/ref_harness."""
    z = load(name)
    n = int(z["n"][0])
    tiers = oracle.tier_model(n)
    keys = np.zeros(0, np.uint64)
    for k in range(int(z["nbatches"][0]) + 1):
        b = z["base"] if k == 0 else z[f"batch{k - 1}"]
        oracle.tier_insert(tiers, oracle.CSR(keys, n), b)
        keys = oracle.set_union(keys, b)
        assert np.array_equal(tiers["btree"], z[f"btree{k}"]), (name, k)
        assert np.array_equal(tiers["degree"], z[f"deg{k}"]), (name, k)
    if name != "graph_s10":
        assert z[f"btree{int(z['nbatches'][0])}"].any()


# --------------------------------------------------------------- CSR builders
@pytest.mark.parametrize("symmetric", [True, False])
def test_edges_to_csr_equals_key_sort(symmetric):
    """The row-wise CSR builder equals the CSR of the sorted unique key
    set (duplicates, self-loops, an empty row, a hub row)."""
    rng = np.random.default_rng(21)
    n = 3000
    s = np.concatenate([rng.integers(0, n, 50_000), np.full(5000, 7), [5, 5, 9]]).astype(np.uint32)
    d = np.concatenate([rng.integers(0, n, 50_000), rng.integers(0, n, 5000), [5, 6, 6]]).astype(np.uint32)
    s[s == 11] = 12  # vertex 11 has no out-edges of its own
    got = oracle.CSR.from_edges(n, s, d, symmetric=symmetric)
    want = oracle.CSR(oracle.build_keys(s, d, symmetric=symmetric), n)
    assert got.m == want.m
    assert np.array_equal(got.off, want.off)
    assert np.array_equal(got.nbr[:got.m], want.nbr[:want.m])


def test_csr_update_equals_set_model():
    rng = np.random.default_rng(22)
    n = 2048
    keys = oracle.build_keys(*oracle.rmat_edges(n, 5, 0.57, 0.19, 0.19, 20000))
    g = oracle.CSR(keys, n)
    ins = oracle.sort_unique((rng.integers(0, n, 4000).astype(np.uint64) << np.uint64(32))
                             | rng.integers(0, n, 4000).astype(np.uint64))
    dels = oracle.sort_unique(np.concatenate([keys[rng.integers(0, len(keys), 3000)], ins[:100],
                                              (np.uint64(n - 1) << np.uint64(32)) + np.arange(5, dtype=np.uint64)]))
    got = g.updated(ins, dels)
    want = oracle.CSR(oracle.set_difference(oracle.set_union(keys, ins), dels), n)
    assert got.m == want.m
    assert np.array_equal(got.off, want.off)
    assert np.array_equal(got.nbr[:got.m], want.nbr[:want.m])


# --------------------------------------------------------------- local triangle count
def test_ltc_brute_force_and_sum():
    """Triangles through v: brute force over pairs of neighbours on random
    graphs, the sum over all v equals 3 x TC, and K_n gives C(n-1, 2)."""
    rng = np.random.default_rng(8)
    for n, m in ((30, 150), (50, 400)):
        edges = {(int(a), int(b)) for a, b in rng.integers(0, n, (m, 2)) if a != b}
        g = undirected(n, edges)
        adj = np.zeros((n, n), bool)
        for a, b in edges:
            adj[a, b] = adj[b, a] = True
        for v in range(n):
            nb = np.nonzero(adj[v])[0]
            brute = sum(1 for a, b in itertools.combinations(nb, 2) if adj[a, b])
            assert oracle.ltc(g, v) == brute, (n, v)
        assert sum(oracle.ltc(g, v) for v in range(n)) == 3 * oracle.tc(g)
    k = 9
    g = undirected(k, [(a, b) for a, b in itertools.combinations(range(k), 2)])
    assert all(oracle.ltc(g, v) == (k - 1) * (k - 2) // 2 for v in range(k))


# --------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught():
    z = load("graph_skew12")
    g = final_graph(z)
    # damped PR (the d=0.85 variant the reference comments out) must fail
    pr = oracle.pagerank(g, 10)
    damped = 0.85 * pr + 0.15 / g.n
    assert oracle.rel_l1(damped, z["pr"]) > 1e-3
    # an off-by-one BFS level must fail the parent check
    lv = oracle.bfs_levels(g, int(z["src"][0]))
    lv2 = lv.copy()
    lv2[lv2 == 2] = 3
    assert oracle.bfs_check_parents(g, int(z["src"][0]), z["bfs_parents"], lv2) > 0
    # BC with the final division dropped (D - 1/σ instead of (D - 1/σ)·σ)
    bc = oracle.bc(g, int(z["src"][0]))
    reach = np.isfinite(bc)
    assert oracle.rel_l1(bc[reach] / 2.0, z["bc"][reach]) > 1e-3


# --------------------------------------------------------------- weighted SSSP
def dijkstra(g, w, src):
    import heapq
    inf = (2**31 - 1) // 2
    dist = np.full(g.n, inf, np.int64)
    dist[src] = 0
    pq = [(0, src)]
    while pq:
        d, u = heapq.heappop(pq)
        if d > dist[u]:
            continue
        for e in range(int(g.off[u]), int(g.off[u + 1])):
            v = int(g.nbr[e])
            nd = d + int(w[e])
            if nd < dist[v]:
                dist[v] = nd
                heapq.heappush(pq, (nd, v))
    return np.minimum(dist, inf).astype(np.int32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sssp_bf_weighted_equals_dijkstra(seed):
    # positive lengths: the Bellman-Ford fixed point is the shortest distance
    n = 1 << 9
    keys = oracle.build_keys(*oracle.rmat_edges(n, 90 + seed, 0.57, 0.19, 0.19, 3000))
    g = oracle.CSR(keys, n)
    rng = np.random.default_rng(seed)
    w = rng.integers(1, 1000, g.m).astype(np.uint32)
    for src in (0, int(np.argmax(g.degrees())), 7):
        assert np.array_equal(oracle.sssp_bf_weighted(g, w, src), dijkstra(g, w, src))
    # unit lengths reduce to the unweighted form
    assert np.array_equal(oracle.sssp_bf_weighted(g, np.ones(g.m, np.uint32), 0), oracle.sssp_bf(g, 0))


def test_weighted_set_model():
    k0 = np.array([5, 9], np.uint64)
    w0 = np.array([50, 90], np.uint32)
    # duplicate key 7 in one batch: the first weight counts; 9 exists: kept
    k, w = oracle.weighted_insert(k0, w0, [7, 9, 7, 1], [70, 99, 71, 10])
    assert k.tolist() == [1, 5, 7, 9] and w.tolist() == [10, 50, 70, 90]
    k, w = oracle.weighted_delete(k, w, [5, 8])
    assert k.tolist() == [1, 7, 9] and w.tolist() == [10, 70, 90]
