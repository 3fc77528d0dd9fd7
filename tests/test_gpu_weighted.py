"""GPU tests of weighted graphs (the reference's WEIGHTED build): weighted
batches, deletes and lookups against the weighted set model, and weighted
SSSP_BF (include/BellmanFord.h) against the oracle."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
MISSING = 0xFFFFFFFF


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def sym_weight(keys, salt=0):
    """Symmetric weights in [1, 1000]: w(u, v) = w(v, u)."""
    s, d = split(keys)
    a = np.minimum(s, d).astype(np.uint64)
    b = np.maximum(s, d).astype(np.uint64)
    h = (a * np.uint64(2654435761) + b * np.uint64(40503) + np.uint64(salt)) >> np.uint64(7)
    return (h % np.uint64(1000) + np.uint64(1)).astype(np.uint32)


def check(G, keys, wts, n, what):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off, ref.off) and np.array_equal(nbr, ref.nbr[:ref.m]), what
    assert np.array_equal(G.edge_weights(*split(keys)), wts), what


def test_weighted_updates_and_sssp():
    n = 1 << 12
    print("[gpu-test] weighted updates", flush=True)
    rng = np.random.default_rng(31)
    base = oracle.build_keys(*oracle.rmat_edges(n, 101, 0.57, 0.19, 0.19, 50000))
    G = tg.Graph(n, weighted=True)
    assert G.weighted
    G.add_edge_batch_weighted(*split(base), sym_weight(base))
    keys, wts = oracle.weighted_insert(np.zeros(0, np.uint64), np.zeros(0, np.uint32), base, sym_weight(base))
    check(G, keys, wts, n, "static")
    assert (G.tiers() == 2).any()  # B-tree vertices carry weights too
    for rnd in range(3):
        fresh = oracle.build_keys(*oracle.rmat_edges(n, 200 + rnd, 0.57, 0.19, 0.19, 2000))
        old = keys[rng.choice(len(keys), 300, replace=False)]
        # existing keys with other weights (kept), fresh keys twice with two weights (first kept)
        ins = np.concatenate([fresh, old, fresh])
        w = np.concatenate([sym_weight(fresh, rnd), sym_weight(old, 77), sym_weight(fresh, 999)])
        G.add_edge_batch_weighted(*split(ins), w)
        keys, wts = oracle.weighted_insert(keys, wts, ins, w)
        check(G, keys, wts, n, f"insert {rnd}")
        pick = keys[rng.choice(len(keys), 1500, replace=False)]
        ps, pd = split(pick)
        dels = oracle.sort_unique(np.concatenate([pick, (pd.astype(np.uint64) << np.uint64(32)) | ps]))
        G.remove_edge_batch(*split(dels))
        keys, wts = oracle.weighted_delete(keys, wts, dels)
        check(G, keys, wts, n, f"delete {rnd}")
        assert np.all(G.edge_weights(*split(dels)) == MISSING)
    g = oracle.CSR(keys, n)
    for src in (0, int(np.argmax(g.degrees())), 17):
        assert np.array_equal(tg.SSSP_BF(G, src), oracle.sssp_bf_weighted(g, wts, src)), src
    G.close()


def test_weighted_errors():
    n = 64
    print("[gpu-test] weighted errors", flush=True)
    G = tg.Graph(n, weighted=True)
    with pytest.raises(Exception):
        G.add_edge_batch([1], [2])  # no weights
    with pytest.raises(Exception):
        G.add_edge_batch_weighted([1], [2], [1 << 31])  # length must fit int32
    with pytest.raises(Exception):
        G.add_edge(1, 2)
    G.add_edge_batch_weighted([1, 2], [2, 1], [5, 5])
    assert G.edge_weights([1, 2, 3], [2, 1, 4]).tolist() == [5, 5, MISSING]
    assert tg.SSSP_BF(G, 1).tolist()[:3] == [(2**31 - 1) // 2, 0, 5]
    G.close()
    U = tg.Graph(n)
    assert not U.weighted
    with pytest.raises(Exception):
        U.edge_weights([1], [2])
    U.close()
