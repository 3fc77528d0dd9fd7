"""GPU tests at the benchmark workloads' full sizes (rMAT scale 22).

The graphs come from the device generator (bit-exact with the oracle's rMAT
at smaller sizes, tests/test_gpu_parity.py) and are pinned at full size by
the counts measured from the reference's generator: directed edges, maximum
degree, one vertex's degree and the vertices per tier.  Every update batch is
checked against the set model, and BFS, PageRank, CC, BC and TC against the
oracle.  Each workload keeps one graph for its module (fixtures), and each
check is its own test."""
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


def mixed_batch(G, keys, n, seed, size, rng, probs, sources=None):
    """One update batch of `size` directed edges, 90% inserts (fresh rMAT
    edges, both directions; or, with `sources`, edges between those vertices
    and random active ones, both directions) and 10% deletes of existing edges (both directions),
    applied as the benchmark applies it: one mixed call (the inserts, then
    the deletes).  Returns the new key set (the set model)."""
    if sources is None:
        ins = oracle.build_keys(*oracle.rmat_edges(n, seed, *probs, max(1, int(size * 0.9) // 2)))
    else:
        active = np.nonzero(oracle.CSR(keys, n).degrees())[0]
        cnt = int(size * 0.9)
        bs = sources[rng.integers(0, len(sources), cnt)].astype(np.uint64)
        bd = active[rng.integers(0, len(active), cnt)].astype(np.uint64)
        ins = sym((bs << np.uint64(32)) | bd)  # both directions: the graph stays symmetric
        ins = ins[(ins >> np.uint64(32)) != (ins & np.uint64(0xFFFFFFFF))]
    dels = sym(keys[rng.choice(len(keys), max(1, int(size * 0.1) // 2), replace=False)])
    G.update_batch(*split(ins), *split(dels))
    return oracle.set_difference(oracle.set_union(keys, ins), dels)


def check_queries(G, g, src, what, bc=True):
    parents, levels = tg.BFS_with_edge_map(G, src, with_levels=True)
    lv = oracle.bfs_levels(g, src)
    assert np.array_equal(levels, lv), f"{what}: BFS levels"
    assert oracle.bfs_check_parents(g, src, parents, lv) == 0, f"{what}: BFS parents"
    assert oracle.rel_l1(tg.PR_S(G, 10), oracle.pagerank(g, 10)) <= FP_TOL, f"{what}: PR"
    assert np.array_equal(tg.CC(G), oracle.cc(g)), f"{what}: CC"
    if bc:
        got = tg.BC(G, src)
        ref = oracle.bc(g, src)
        reach = lv != UNREACHED
        assert np.all(np.isnan(got[~reach])), f"{what}: BC unreachable not NaN"
        assert oracle.rel_l1(got[reach], ref[reach]) <= FP_TOL, f"{what}: BC"


# ------------------------------------------------------------- plain workload
@pytest.fixture(scope="module")
def plain():
    """rMAT scale 22, 69M generated edges (a=.5 b=.1 c=.1), built on the
    device from the generator's output in HBM."""
    n = 1 << 22
    say("plain workload: build")
    s, d, w, keys = device_graph(n, (0.5, 0.1, 0.1), 69_000_000)
    G = tg.Graph(n)
    G.add_edge_batch_dev(s, d, w)
    del s, d
    st = {"G": G, "keys": keys, "n": n, "rng": np.random.default_rng(5)}
    yield st
    G.close()


def test_plain_static_graph(plain):
    """135,910,260 directed edges, maximum degree 1,441, deg(9) = 707,
    17 B-tree vertices (the reference generator's counts)."""
    G, keys, n = plain["G"], plain["keys"], plain["n"]
    assert len(keys) == 135_910_260
    deg = G.degrees()
    assert int(deg.max()) == 1441 and int(deg[9]) == 707
    assert int((G.tiers() == 2).sum()) == 17
    assert_adjacency(G, keys, n, "static")


@pytest.mark.parametrize("size", [100_000, 100_000, 1_000, 10_000_000])
def test_plain_mixed_batch(plain, size):
    """Mixed batches of 1e5 (twice), 1e3 and 1e7 update edges (90% inserts
    of fresh rMAT edges generated with ith_rand(1000 + t), 10% deletes of
    existing edges in both directions): the adjacency equals the set model
    after each."""
    plain["t"] = t = plain.get("t", -1) + 1
    say("plain workload: mixed batch", t, size)
    seed = oracle.aspen_ith_rand(1000 + t) & 0xFFFFFFFF
    plain["keys"] = mixed_batch(plain["G"], plain["keys"], plain["n"], seed, size, plain["rng"], (0.5, 0.1, 0.1))
    assert_adjacency(plain["G"], plain["keys"], plain["n"], f"batch {t} of {size}")


def test_plain_queries(plain):
    """BFS from 9, PageRank 10 iterations, CC and BC from 9 on the graph
    after the batches equal the oracle."""
    say("plain workload: queries")
    g = oracle.CSR(plain["keys"], plain["n"])
    check_queries(plain["G"], g, 9, "plain full size")


# ------------------------------------------------------------- skew workload
@pytest.fixture(scope="module")
def skew():
    """Skewed rMAT scale 22, 120M generated edges (a=.57 b=.19 c=.19)."""
    n = 1 << 22
    say("skew workload: build")
    s, d, w, keys = device_graph(n, (0.57, 0.19, 0.19), 120_000_000)
    G = tg.Graph(n)
    G.add_edge_batch_dev(s, d, w)
    del s, d
    st = {"G": G, "keys": keys, "n": n, "rng": np.random.default_rng(9)}
    yield st
    G.close()


def test_skew_static_graph(skew):
    """224,914,926 directed edges, maximum degree 243,995; vertices with
    edges by tier: 1,685,167 inline only, 989,920 PMA, 35,443 B-tree."""
    G, keys, n = skew["G"], skew["keys"], skew["n"]
    assert len(keys) == 224_914_926
    deg = G.degrees()
    assert int(deg.max()) == 243_995
    tiers = G.tiers()
    assert int(((tiers == 0) & (deg > 0)).sum()) == 1_685_167
    assert int((tiers == 1).sum()) == 989_920
    assert int((tiers == 2).sum()) == 35_443
    assert_adjacency(G, keys, n, "static")


def test_skew_hub_round_trip(skew):
    """A 1e6-edge batch whose sources are the top-1% degree vertices,
    inserted and then its new edges deleted, leaves the adjacency as it
    was."""
    G, keys, n, rng = skew["G"], skew["keys"], skew["n"], skew["rng"]
    say("skew workload: hub round trip")
    deg = G.degrees()
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


def test_skew_hub_mixed_batch(skew):
    """A 1e6-edge mixed batch skewed toward the top-1% degree vertices (90%
    inserts from them, 10% deletes of existing edges) equals the set
    model."""
    G, n, rng = skew["G"], skew["n"], skew["rng"]
    say("skew workload: hub-targeted mixed batch")
    deg = G.degrees()
    top = np.argsort(deg)[-(n // 100):]
    skew["keys"] = mixed_batch(G, skew["keys"], n, 0, 1_000_000, rng, None, sources=top)
    assert_adjacency(G, skew["keys"], n, "hub mixed batch")


def test_skew_queries(skew):
    """BFS from the largest hub, PageRank 10 iterations, CC and BC equal the
    oracle."""
    say("skew workload: queries")
    g = oracle.CSR(skew["keys"], skew["n"])
    check_queries(skew["G"], g, int(np.argmax(g.degrees())), "skew full size")


def test_skew_triangle_count(skew):
    """TC (the degree-ordered device count) and TC.h's edgeMap + countF
    form both equal the oracle's count_common sum over every edge."""
    say("skew workload: TC")
    g = oracle.CSR(skew["keys"], skew["n"])
    want = oracle.tc(g)
    assert tg.TC(skew["G"]) == want
    cnt = tg.DeviceBuffer(8)
    cnt.upload(np.zeros(1, np.uint64))
    vs = tg.VertexSubset(None, 0, all=True, max_el=skew["n"])
    tg.edgeMap(skew["G"], vs, tg.OP_TC, tg.OpArgs(count=cnt.ptr), output=False)
    assert int(cnt.download(np.uint64, 1)[0]) == want
    cnt.close()
