"""GPU parity: the HIP path through the C ABI against the oracle and the
reference-build golden fixtures.

Bars: adjacency, degrees, BFS levels, CC labels, TC and tier assignment
bit-exact; BFS parents validated; PR and BC within 1e-9 relative L1 (FP64,
different summation order).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GRAPHS = ["graph_s10", "graph_skew12", "graph_promote11"]
UNREACHED = np.iinfo(np.uint32).max
FP_TOL = 1e-9


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def assert_adjacency(G, keys, n, what=""):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    deg = (off[1:] - off[:-1]).astype(np.uint32)
    bad = np.nonzero(deg != ref.degrees())[0]
    assert len(bad) == 0, f"{what}: degree mismatch at {bad[:10]}"
    assert np.array_equal(nbr, ref.nbr[:ref.m]), f"{what}: neighbour lists differ"
    assert np.array_equal(G.degrees(), ref.degrees()), what
    assert G.get_num_edges() == ref.m
    assert G.stats()["pma_elements"] == G.tier_counts()["pma_edges"], what  # host-kept PMA count


def check_algorithms(G, g, src, iters=10, what=""):
    parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
    lv = oracle.bfs_levels(g, src)
    assert np.array_equal(levels, lv), f"{what}: BFS levels differ"
    assert oracle.bfs_check_parents(g, src, parents, lv) == 0, f"{what}: BFS parents invalid"
    pr = tg.PR_S(G, iters)
    assert oracle.rel_l1(pr, oracle.pagerank(g, iters)) <= FP_TOL, f"{what}: PR"
    assert np.array_equal(tg.CC(G), oracle.cc(g)), f"{what}: CC"
    bc = tg.BC(G, src)
    ref = oracle.bc(g, src)
    reach = lv != UNREACHED
    assert np.all(np.isnan(bc[~reach])), f"{what}: BC unreachable not NaN"
    assert oracle.rel_l1(bc[reach], ref[reach]) <= FP_TOL, f"{what}: BC"
    assert tg.TC(G) == oracle.tc(g), f"{what}: TC"
    assert np.array_equal(tg.SSSP_BF(G, src), oracle.sssp_bf(g, src)), f"{what}: SSSP_BF"


# ------------------------------------------------------------------ generator
@pytest.mark.parametrize("nominal", [False, True])
def test_rmat_device_matches_oracle(nominal):
    n, seed, count = 1 << 20, oracle.aspen_ith_rand(100) & 0xFFFFFFFF, 200000
    say("rmat", nominal, n, count)
    s, d, w = tg.rmat_device(n, seed, 0.5, 0.1, 0.1, count, nominal=nominal)
    assert w == count
    gs, gd = s.download(np.uint32, count), d.download(np.uint32, count)
    os_, od = oracle.rmat_edges(n, seed, 0.5, 0.1, 0.1, count, nominal=nominal)
    got = np.sort((gs.astype(np.uint64) << np.uint64(32)) | gd)
    want = np.sort((os_.astype(np.uint64) << np.uint64(32)) | od)
    assert np.array_equal(got, want)
    # symmetric + loop-free form equals the oracle's key build
    s, d, w = tg.rmat_device(n, seed, 0.5, 0.1, 0.1, count, nominal=nominal, symmetric=True,
                             drop_loops=True)
    keys = (s.download(np.uint32, w).astype(np.uint64) << np.uint64(32)) | d.download(np.uint32, w)
    assert np.array_equal(np.unique(keys), oracle.build_keys(os_, od))


# ------------------------------------------------------------------ golden graphs
@pytest.mark.parametrize("name", GRAPHS)
def test_golden_graph(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    n = int(z["n"][0])
    say("golden", name, n)
    G = tg.Graph(n)
    keys = z["base"]
    tiers = oracle.tier_model(n)
    oracle.tier_insert(tiers, oracle.CSR(np.zeros(0, np.uint64), n), keys)
    s, d = split(keys)
    G.add_edge_batch(s, d)
    for k in range(int(z["nbatches"][0]) + 1):
        if k:
            b = z[f"batch{k - 1}"]
            oracle.tier_insert(tiers, oracle.CSR(keys, n), b)
            s, d = split(b)
            G.add_edge_batch(s, d)
            keys = oracle.set_union(keys, b)
        off, nbr = G.export_adjacency()
        assert np.array_equal((off[1:] - off[:-1]).astype(np.uint32), z[f"deg{k}"]), (name, k)
        assert np.array_equal(nbr, z[f"nbrs{k}"]), (name, k)
        t = G.tiers()
        assert np.array_equal(t == 2, tiers["btree"] == 1), (name, k, "B-tree tier")
    g = oracle.CSR(keys, n)
    src = int(z["src"][0])
    parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
    assert oracle.bfs_check_parents(g, src, parents, levels) == 0
    assert np.array_equal(levels, oracle.bfs_levels(g, src))
    assert oracle.rel_l1(tg.PR_S(G, int(z["iters"][0])), z["pr"]) <= FP_TOL
    assert np.array_equal(tg.CC(G), z["cc"])
    bc = tg.BC(G, src)
    reach = levels != UNREACHED
    assert oracle.rel_l1(bc[reach], z["bc"][reach]) <= FP_TOL
    assert np.all(np.isnan(bc[~reach]))
    assert tg.TC(G) == int(z["tc"][0])
    assert np.array_equal(tg.SSSP_BF(G, src), z["sssp"])  # the reference build's SSSP_BF output
    G.close()


# ------------------------------------------------------------------ updates vs set model
@pytest.mark.parametrize("seed,skew", [(1, False), (2, True)])
def test_mixed_updates_match_set_model(seed, skew):
    n = 1 << 12
    a, b, c = (0.57, 0.19, 0.19) if skew else (0.5, 0.1, 0.1)
    say("mixed", seed, skew)
    rng = np.random.default_rng(seed)
    G = tg.Graph(n)
    keys = oracle.build_keys(*oracle.rmat_edges(n, seed, a, b, c, 40000))
    G.add_edge_batch(*split(keys))
    assert_adjacency(G, keys, n, "static")
    for rnd in range(6):
        ins = oracle.build_keys(*oracle.rmat_edges(n, 100 * seed + rnd, a, b, c, 3000))
        # deletes: existing edges (both directions) plus absent ones
        pick = keys[rng.choice(len(keys), size=min(len(keys), 2500), replace=False)]
        ps, pd = split(pick)
        dels = oracle.sort_unique(np.concatenate([pick, (pd.astype(np.uint64) << np.uint64(32)) | ps,
                                                  oracle.build_keys(*oracle.rmat_edges(n, 999 + rnd, a, b, c, 50))]))
        G.add_edge_batch(*split(ins))
        keys = oracle.set_union(keys, ins)
        assert_adjacency(G, keys, n, f"insert {rnd}")
        G.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(keys, dels)
        assert_adjacency(G, keys, n, f"delete {rnd}")
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(np.argmax(g.degrees())), what="mixed")
    G.close()


def test_random_small_batches_all_tiers():
    """40 symmetric batches of log-uniform size (1..3000 keys, inserts and
    deletes of existing edges, a third of them aimed at hubs) on a graph
    with all three tiers: the in-place tree and in-leaf PMA paths and their
    fallbacks in random mixes.  The adjacency equals the set model after
    every batch; a second graph given the same batches coalesced through
    apply_ops ends equal; every algorithm matches the oracle."""
    n = 1 << 14
    say("random small batches")
    rng = np.random.default_rng(77)

    def sym(k):
        k = np.asarray(k, np.uint64)
        rev = ((k & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (k >> np.uint64(32))
        k = oracle.sort_unique(np.concatenate([k, rev]))
        return k[(k >> np.uint64(32)) != (k & np.uint64(0xFFFFFFFF))]

    keys = oracle.build_keys(*oracle.rmat_edges(n, 17, 0.57, 0.19, 0.19, 300000))
    G, H = tg.Graph(n), tg.Graph(n)
    G.add_edge_batch(*split(keys))
    H.add_edge_batch(*split(keys))
    tiers = G.tiers()
    assert (tiers == 1).any() and (tiers == 2).any()
    hubs = np.nonzero(tiers == 2)[0].astype(np.uint64)
    ops_s, ops_d, ops_o = [], [], []
    st0 = G.stats()
    for rnd in range(40):
        size = int(np.exp(rng.uniform(0, np.log(1500))))
        if rnd % 3 == 0:
            src = hubs[rng.integers(0, len(hubs), size)]
        else:
            src = rng.integers(0, n, size).astype(np.uint64)
        if rnd % 2 == 0:
            batch = sym((src << np.uint64(32)) | rng.integers(0, n, size).astype(np.uint64))
            G.add_edge_batch(*split(batch))
            keys = oracle.set_union(keys, batch)
            op = 0
        else:
            batch = sym(keys[rng.choice(len(keys), min(size, len(keys)), replace=False)])
            G.remove_edge_batch(*split(batch))
            keys = oracle.set_difference(keys, batch)
            op = 1
        bs, bd = split(batch)
        ops_s.append(bs)
        ops_d.append(bd)
        ops_o.append(np.full(len(batch), op, np.uint32))
        assert_adjacency(G, keys, n, f"random batch {rnd} ({len(batch)} keys, op {op})")
    st1 = G.stats()
    assert st1["tree_small_updates"] > st0["tree_small_updates"]
    assert st1["pma_leaf_merges"] > st0["pma_leaf_merges"]
    H.apply_ops(np.concatenate(ops_s), np.concatenate(ops_d), np.concatenate(ops_o))
    assert_adjacency(H, keys, n, "coalesced replay")
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(hubs[0]), what="random small batches")
    G.close()
    H.close()


def test_delete_everything_then_reinsert():
    n = 1 << 11
    say("delete-all")
    hub = np.arange(20, 1600, dtype=np.uint32)
    keys = oracle.sort_unique(np.concatenate([
        oracle.build_keys(*oracle.rmat_edges(n, 5, 0.57, 0.19, 0.19, 30000)),
        oracle.build_keys(np.full(len(hub), 3, np.uint32), hub)]))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    assert (G.tiers() == 2).any()
    half = keys[::2]
    G.remove_edge_batch(*split(half))
    rest = oracle.set_difference(keys, half)
    assert_adjacency(G, rest, n, "half")
    G.remove_edge_batch(*split(rest))
    assert G.get_num_edges() == 0
    assert_adjacency(G, np.zeros(0, np.uint64), n, "empty")
    # B-trees are never demoted (include/graph.h); they keep working empty
    G.add_edge_batch(*split(keys))
    assert_adjacency(G, keys, n, "reinsert")
    G.close()


def test_promotion_and_hub_shrink():
    n = 4096
    say("promotion")
    G = tg.Graph(n)
    e = lambda s, ds: np.array([(s << 32) | d for d in ds], np.uint64)  # noqa: E731
    keys = e(7, range(100, 900))
    G.add_edge_batch(*split(keys))
    assert G.tiers()[7] == 1
    more = e(7, range(900, 1300))
    G.add_edge_batch(*split(more))
    keys = oracle.set_union(keys, more)
    assert G.tiers()[7] == 2
    assert_adjacency(G, keys, n, "promoted")
    drop = e(7, range(100, 1290))
    G.remove_edge_batch(*split(drop))
    keys = oracle.set_difference(keys, drop)
    assert_adjacency(G, keys, n, "shrunk")
    assert G.degree(7) == 10 and G.tiers()[7] == 2
    assert G.neighbors(7).tolist() == list(range(1290, 1300))
    G.close()


# ------------------------------------------------------------------ API edge cases
def test_api_edge_cases():
    say("api")
    G = tg.Graph(100)
    G.add_edge_batch([], [])
    G.remove_edge_batch([], [])
    assert G.get_num_edges() == 0
    # duplicates and self-loops inside one batch
    G.add_edge_batch([1, 1, 1, 5, 5], [2, 2, 1, 6, 6])
    assert G.get_num_edges() == 3 and G.is_edge(1, 1) == 1 and G.is_edge(1, 2) == 1
    assert G.is_edge(2, 1) == 0
    assert G.add_edge(1, 2) == -1 and G.add_edge(2, 1) == 0
    assert G.remove_edge(2, 1) == 0 and G.remove_edge(2, 1) == -1
    assert G.neighbors(1).tolist() == [1, 2]
    with pytest.raises(tg.TerraceError) as ei:
        G.add_edge_batch([1, 200], [2, 3])
    assert ei.value.code == -1
    assert G.get_num_edges() == 3  # rejected batch changed nothing
    with pytest.raises(tg.TerraceError):
        G.degree(100)
    with pytest.raises(tg.TerraceError):
        tg.BFS_with_edge_map(G, 1000)
    # BFS from an isolated vertex
    parents, levels = tg.BFS_with_edge_map(G, 50, with_levels=True)
    assert parents[50] == 50 and (parents >= 0).sum() == 1 and levels[50] == 0
    bc = tg.BC(G, 50)
    assert bc[50] == 0.0 and np.isnan(bc[:50]).all()
    G.close()


def test_mid_size_batches_bucket_sort():
    """Batches of 4097..2^18 raw pairs are sorted by 4096 buckets of the
    key's top bits, each bucket sorted and de-duplicated in LDS.  Unsorted
    pairs with repeats, ids up to n - 1 on a vertex count that is not a
    power of two, inserts and deletes (some of absent edges) match the set
    model after every batch; batches whose keys share one source (one
    bucket beyond its LDS capacity) are sorted again by radix and match
    too; a mid-size batch holding one bad id is rejected without changing
    the graph."""
    n = 100_003
    say("mid-size batches")
    rng = np.random.default_rng(41)
    keys = oracle.build_keys(*oracle.rmat_edges(1 << 16, 23, 0.57, 0.19, 0.19, 200000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    for rnd, size in enumerate([4097, 9000, 16384, 16385, 12000, 5000]):
        hi = np.uint64(n)
        s = rng.integers(0, n, size).astype(np.uint64)
        d = rng.integers(0, n, size).astype(np.uint64)
        s[:50], d[:50] = hi - np.uint64(1), rng.integers(0, n, 50).astype(np.uint64)
        raw = (s << np.uint64(32)) | d
        raw[size // 2:size // 2 + 200] = raw[:200]  # repeats inside the batch
        if rnd % 2 == 0:
            G.add_edge_batch(*split(raw))
            keys = oracle.set_union(keys, oracle.sort_unique(raw))
        else:
            raw[:size // 2] = keys[rng.choice(len(keys), size // 2, replace=False)]  # existing edges
            G.remove_edge_batch(*split(raw))
            keys = oracle.set_difference(keys, oracle.sort_unique(raw))
        assert_adjacency(G, keys, n, f"mid-size batch {rnd} ({size})")
    for rnd, hub in enumerate([7, n - 1]):  # one bucket holds the whole batch
        d = rng.permutation(n)[:12000].astype(np.uint64)
        raw = (np.uint64(hub) << np.uint64(32)) | d
        G.add_edge_batch(*split(raw))
        keys = oracle.set_union(keys, oracle.sort_unique(raw))
        assert_adjacency(G, keys, n, f"one-source insert {rnd}")
        G.remove_edge_batch(*split(raw[::2]))
        keys = oracle.set_difference(keys, oracle.sort_unique(raw[::2]))
        assert_adjacency(G, keys, n, f"one-source delete {rnd}")
    m = G.get_num_edges()
    bs, bd = split(keys[:10000])
    bs = bs.copy()
    bs[7000] = n  # out of range
    with pytest.raises(tg.TerraceError) as ei:
        G.remove_edge_batch(bs, bd)
    assert ei.value.code == -1 and G.get_num_edges() == m
    G.close()


def test_delete_prefix_moves_begin_across_leaves():
    """Vertices of degree ~500 (PMA tier, their runs spanning several
    64-slot leaves) lose their smallest neighbours in one batch: the 13
    inline ones and the next 100, then the next 200, then all but 10.  The
    begin pointer must skip whole emptied leaves; the adjacency (exported
    through the begin pointers) equals the set model after every batch, and
    BFS, PageRank and CC match the oracle."""
    n = 1 << 14
    say("delete prefix across leaves")
    rng = np.random.default_rng(13)
    hubs = np.array([3, 777, 9000, 16000], np.uint64)

    def sym(k):  # both directions (the traversals expect an undirected graph)
        k = np.asarray(k, np.uint64)
        rev = ((k & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (k >> np.uint64(32))
        return oracle.sort_unique(np.concatenate([k, rev]))

    keys = oracle.build_keys(*oracle.rmat_edges(n, 29, 0.5, 0.1, 0.1, 60000))
    extra = []
    for h in hubs:
        nb = np.sort(rng.choice(n, 520, replace=False)).astype(np.uint64)
        nb = nb[np.isin(nb, hubs, invert=True)]
        extra.append((h << np.uint64(32)) | nb)
    keys = oracle.set_union(keys, sym(np.concatenate(extra)))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    assert (G.tiers()[hubs.astype(np.int64)] == 1).all()
    assert_adjacency(G, keys, n, "built")
    for cut in (113, 200, -10):  # from the front of each current row
        dels = []
        for h in hubs:
            row = keys[(keys >> np.uint64(32)) == h]
            dels.append(row[:cut])
        dels = sym(np.concatenate(dels))
        G.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(keys, dels)
        assert_adjacency(G, keys, n, f"prefix delete {cut}")
    assert (G.degrees()[hubs.astype(np.int64)] == 10).all()
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(hubs[0]), what="prefix deletes")
    G.close()


def test_edge_map_frontier_api():
    n = 1 << 10
    say("edge_map api")
    keys = oracle.build_keys(*oracle.rmat_edges(n, 9, 0.5, 0.1, 0.1, 5000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    g = oracle.CSR(keys, n)
    src = int(np.argmax(g.degrees()))
    lv = oracle.bfs_levels(g, src)
    parent = tg.DeviceBuffer(n * 4)
    parent.upload(np.where(np.arange(n) == src, src, -1).astype(np.int32))
    vs = tg.VertexSubset(G, src)
    depth = 0
    for threshold in (20,):
        while vs.not_empty():
            depth += 1
            args = tg.OpArgs(parent=parent.ptr, depth=depth)
            nxt = tg.edgeMap(G, vs, tg.OPS["BFS"], args, True, threshold)
            assert nxt.to_array().tolist() == np.nonzero(lv == depth)[0].tolist(), depth
            vs = nxt
    G.close()


# ------------------------------------------------------------------ larger scale
def test_scale16_static_and_batches():
    n = 1 << 16
    say("scale16")
    keys = oracle.build_keys(*oracle.rmat_edges(n, 77, 0.57, 0.19, 0.19, 600000))
    G = tg.Graph(n)
    s, d, w = tg.rmat_device(n, 77, 0.57, 0.19, 0.19, 600000, symmetric=True, drop_loops=True)
    G.add_edge_batch_dev(s, d, w)
    assert_adjacency(G, keys, n, "static dev")
    for t in range(3):
        b = oracle.build_keys(*oracle.rmat_edges(n, 1000 + t, 0.57, 0.19, 0.19, 20000 * (t + 1)))
        G.add_edge_batch(*split(b))
        keys = oracle.set_union(keys, b)
    assert_adjacency(G, keys, n, "after batches")
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(np.argmax(g.degrees())), what="scale16")
    G.close()


def test_tc_complete_graph_closed_form():
    """K_2100: every vertex in the B-tree tier; degree-ordered out-lists of
    up to 2099 entries exceed the LDS-staged length (2048), so both search
    paths run.  TC = C(n, 3), and each vertex lies in C(n-1, 2) triangles."""
    n = 2100
    print("[gpu-test] tc complete graph", flush=True)
    a, b = np.meshgrid(np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), indexing="ij")
    keep = a != b
    G = tg.Graph(n)
    G.add_edge_batch(a[keep].ravel(), b[keep].ravel())
    assert G.get_num_edges() == n * (n - 1)
    assert tg.TC(G) == n * (n - 1) * (n - 2) // 6
    v = np.array([0, 1, n // 2, n - 1], np.uint32)
    assert np.array_equal(tg.local_triangles(G, v), np.full(4, (n - 1) * (n - 2) // 2, np.uint64))
    G.close()


def test_skew_all_tiers_hub_targeted_batches():
    """Skewed rMAT (a=.57 b=.19 c=.19) with all three tiers populated, then
    mixed batches whose sources are the top-1% degree vertices (B-tree
    splits, PMA windows around hubs, promotions), checked against the set
    model after every batch and by every algorithm at the end."""
    n = 1 << 14
    say("skew hub-targeted")
    rng = np.random.default_rng(23)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 31, 0.57, 0.19, 0.19, 400000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    tiers = G.tiers()
    assert (tiers == 0).any() and (tiers == 1).any() and (tiers == 2).any()
    assert_adjacency(G, keys, n, "static")
    for rnd in range(4):
        deg = oracle.CSR(keys, n).degrees()
        active = np.nonzero(deg)[0]
        top = np.argsort(deg)[-max(1, n // 100):]
        s = top[rng.integers(0, len(top), 20000)].astype(np.uint32)
        d = active[rng.integers(0, len(active), 20000)].astype(np.uint32)
        ins = oracle.build_keys(s, d)
        pick = keys[rng.choice(len(keys), 3000, replace=False)]
        ps, pd = split(pick)
        dels = oracle.sort_unique(np.concatenate([pick, (pd.astype(np.uint64) << np.uint64(32)) | ps]))
        G.add_edge_batch(*split(ins))
        keys = oracle.set_union(keys, ins)
        assert_adjacency(G, keys, n, f"hub insert {rnd}")
        G.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(keys, dels)
        assert_adjacency(G, keys, n, f"hub delete {rnd}")
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(np.argmax(g.degrees())), what="skew")
    G.close()


def test_small_tree_updates_in_place_and_fallback():
    """Small batches into B-tree vertices take the in-place path (every
    touched leaf block keeps 1..256 keys: block counts and first keys change,
    the index does not); a batch that overfills a block or empties one takes
    the copy-on-write pass.  Their reverse edges go to PMA leaves that take
    them in place (in-leaf merge).  The adjacency equals the set model after every
    batch, both paths are seen in the counters, and every algorithm matches
    the oracle at the end (batches are symmetric)."""
    n = 1 << 12
    say("small tree updates")
    rng = np.random.default_rng(41)
    hubs = np.array([3, 100, 2000], np.uint64)

    def sym(k):
        k = np.asarray(k, np.uint64)
        rev = ((k & np.uint64(0xFFFFFFFF)) << np.uint64(32)) | (k >> np.uint64(32))
        return oracle.sort_unique(np.concatenate([k, rev]))

    parts = []
    for h in hubs:
        nb = rng.choice(np.arange(14, n, dtype=np.uint64), 2500, replace=False)
        parts.append((h << np.uint64(32)) | nb[nb != h])
    bg = oracle.build_keys(*oracle.rmat_edges(n, 5, 0.5, 0.1, 0.1, 20000))
    keys = oracle.sort_unique(np.concatenate([sym(np.concatenate(parts)), bg]))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    assert np.all(G.tiers()[hubs.astype(np.int64)] == 2)
    st0 = G.stats()
    for rnd in range(24):
        h = hubs[rnd % 3]
        cnt = int(rng.integers(1, 120))
        if rnd % 2 == 0:
            nb = rng.choice(np.arange(0, n, dtype=np.uint64), cnt, replace=False)
            ins = sym((h << np.uint64(32)) | nb[nb != h])
            G.add_edge_batch(*split(ins))
            keys = oracle.set_union(keys, ins)
        else:
            mine = keys[(keys >> np.uint64(32)) == h]
            dels = sym(mine[rng.choice(len(mine), min(cnt, len(mine)), replace=False)])
            G.remove_edge_batch(*split(dels))
            keys = oracle.set_difference(keys, dels)
        assert_adjacency(G, keys, n, f"small tree batch {rnd}")
    st1 = G.stats()
    assert st1["tree_small_updates"] > st0["tree_small_updates"]
    assert st1["pma_leaf_merges"] > st0["pma_leaf_merges"]  # the reverse edges land in PMA leaves
    # 300 new neighbours of one hub in one id range overfill a block
    h = hubs[0]
    have = set((keys[(keys >> np.uint64(32)) == h] & np.uint64(0xFFFFFFFF)).tolist())
    fresh = np.array([v for v in range(n) if v not in have and v != h][:300], np.uint64)
    ins = sym((h << np.uint64(32)) | fresh)
    G.add_edge_batch(*split(ins))
    keys = oracle.set_union(keys, ins)
    assert_adjacency(G, keys, n, "overfilled block")
    st2 = G.stats()
    assert st2["tree_blocks_written"] > st1["tree_blocks_written"]
    # deleting a hub's 400 largest neighbours empties its last block
    mine = keys[(keys >> np.uint64(32)) == hubs[1]]
    dels = sym(mine[-400:])
    G.remove_edge_batch(*split(dels))
    keys = oracle.set_difference(keys, dels)
    assert_adjacency(G, keys, n, "emptied block")
    assert G.stats()["tree_blocks_written"] >= st2["tree_blocks_written"]
    # inline slots refilled from a tree whose in-place delete did not fit
    # (an emptied block): the refill is redone after the copy-on-write pass
    mine = keys[(keys >> np.uint64(32)) == hubs[2]]
    dels = sym(np.concatenate([mine[:13], mine[-400:]]))
    G.remove_edge_batch(*split(dels))
    keys = oracle.set_difference(keys, dels)
    assert_adjacency(G, keys, n, "refill after fallback")
    g = oracle.CSR(keys, n)
    check_algorithms(G, g, int(hubs[2]), what="small tree")
    G.close()


def test_cfg_cpu_small_full_size():
    """The reference's CPU-runnable workload at full size: rMAT scale 20,
    8M generated edges (a=.5 b=.1 c=.1, nominal generator, static seed
    ith_rand(100)), symmetrized; static build, then 10 insert batches of 1e5
    generated edges (seeds ith_rand(1000 + t)); BFS from 9, PageRank 10
    iterations, CC, BC from 9 and TC, all against the oracle."""
    n = 1 << 20
    say("cfg cpu-small full size")
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    keys = oracle.build_keys(*oracle.rmat_edges(n, seed, 0.5, 0.1, 0.1, 8_000_000))
    assert len(keys) == 15_706_884  # the survey's measured count for this workload
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    assert_adjacency(G, keys, n, "static")
    for t in range(10):
        ins = oracle.build_keys(*oracle.rmat_edges(n, oracle.aspen_ith_rand(1000 + t) & 0xFFFFFFFF,
                                                   0.5, 0.1, 0.1, 50_000))
        G.add_edge_batch(*split(ins))
        keys = oracle.set_union(keys, ins)
    assert_adjacency(G, keys, n, "after 10 batches")
    g = oracle.CSR(keys, n)
    assert g.degrees()[9] > 0
    check_algorithms(G, g, 9, what="cfg cpu-small")
    G.close()
