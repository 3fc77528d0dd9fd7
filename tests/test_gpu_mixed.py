"""GPU tests of the mixed update batch (tg_update_batch: the inserts, then
the deletes, in one call with one sort and preparation) against the set
model (K ∪ ins) \\ del, and against the same halves applied as two calls on
a second graph.  Sizes cover the one-workgroup sort (at most 4096 edges in
all), the bucket sort (up to 2^18) with its overflow fallback, and the radix
sort."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import TerraceError

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def assert_set_model(G, keys, n, what):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off, ref.off), what
    assert np.array_equal(nbr, ref.nbr[:ref.m]), what
    assert G.stats()["pma_elements"] == G.tier_counts()["pma_edges"], what


def mixed(rng, keys, n, n_ins, n_del, hubs):
    """Insert keys: fresh random edges, a third from hubs, some present
    already, some repeated; delete keys: present edges, absent ones, some
    repeated, and some of the inserted ones (those end absent)."""
    src = np.where(rng.random(n_ins) < 0.33, hubs[rng.integers(0, len(hubs), n_ins)], rng.integers(0, n, n_ins))
    ins = (src.astype(np.uint64) << np.uint64(32)) | rng.integers(0, n, n_ins).astype(np.uint64)
    ins[: n_ins // 20] = keys[rng.integers(0, len(keys), n_ins // 20)]
    ins[-(n_ins // 50) - 1:] = ins[: n_ins // 50 + 1]
    parts = [keys[rng.integers(0, len(keys), max(0, n_del - n_del // 4))],
             (rng.integers(0, n, n_del // 8).astype(np.uint64) << np.uint64(32))
             | rng.integers(0, n, n_del // 8).astype(np.uint64),
             ins[rng.integers(0, n_ins, n_del // 8)] if n_ins else np.zeros(0, np.uint64)]
    dels = np.concatenate(parts)[:n_del]
    rng.shuffle(ins)
    rng.shuffle(dels)
    return ins, dels


@pytest.mark.parametrize("n_ins,n_del", [(900, 100), (3000, 1096), (4096, 0), (0, 3000), (90_000, 10_000),
                                         (200_000, 62_144), (400_000, 60_000)])
def test_update_batch_matches_set_model(n_ins, n_del):
    n = 1 << 15
    say("update_batch", n_ins, n_del)
    rng = np.random.default_rng(n_ins * 7 + n_del)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 19, 0.57, 0.19, 0.19, 400_000))
    G, H = tg.Graph(n), tg.Graph(n)
    G.add_edge_batch(*split(keys))
    H.add_edge_batch(*split(keys))
    tiers = G.tiers()
    assert (tiers == 1).any() and (tiers == 2).any()
    hubs = np.nonzero(tiers == 2)[0]
    for rnd in range(3):
        ins, dels = mixed(rng, keys, n, n_ins, n_del, hubs)
        G.update_batch(*split(ins), *split(dels))
        if len(ins):
            H.add_edge_batch(*split(ins))
        if len(dels):
            H.remove_edge_batch(*split(dels))
        keys = oracle.set_difference(oracle.set_union(keys, oracle.sort_unique(ins)), oracle.sort_unique(dels))
        assert_set_model(G, keys, n, f"round {rnd}")
        o1, n1 = G.export_adjacency()
        o2, n2 = H.export_adjacency()
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2), f"round {rnd}: two calls differ"
    G.close()
    H.close()


def test_update_batch_bucket_overflow_and_errors():
    """12,000 inserts from one source overflow a sort bucket (the batch is
    sorted again by radix); an id out of range rejects the whole batch."""
    n = 1 << 14
    say("update_batch overflow")
    rng = np.random.default_rng(3)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 23, 0.5, 0.1, 0.1, 60_000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    ins = (np.uint64(77) << np.uint64(32)) | rng.permutation(n)[:12_000].astype(np.uint64)
    dels = keys[rng.integers(0, len(keys), 3000)]
    G.update_batch(*split(ins), *split(dels))
    keys = oracle.set_difference(oracle.set_union(keys, oracle.sort_unique(ins)), oracle.sort_unique(dels))
    assert_set_model(G, keys, n, "overflow")
    bad_s, bad_d = split(keys[:10])
    bad_d = bad_d.copy()
    bad_d[5] = n
    with pytest.raises(TerraceError) as e:
        G.update_batch(bad_s, bad_d, *split(keys[10:20]))
    assert e.value.code == -1
    assert_set_model(G, keys, n, "after the rejected batch")
    G.close()


@pytest.mark.parametrize("n_ins,n_del", [(1500, 500), (60_000, 20_000)])
def test_insert_fallbacks_rerun_the_delete_half(n_ins, n_del):
    """Mixed batches whose insert half needs the host's general passes: new
    B-trees (promotions of vertices pushed past 1024 neighbours) and PMA
    leaves that overflow their bound (the window pass).  The delete half,
    launched behind the insert half, must then have written nothing and run
    again after those passes; the adjacency equals the set model after
    every batch, and the deletes' own counts are reported."""
    n = 1 << 13
    say("insert fallbacks", n_ins, n_del)
    rng = np.random.default_rng(n_ins + 5)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 29, 0.5, 0.1, 0.1, 20_000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    st0 = G.stats()
    for rnd in range(3):
        # a few sources get 1100 new neighbours each (promotions), the rest
        # random edges crowding the PMA's leaves; deletes of present edges
        hub = rng.choice(np.nonzero(G.degrees() < 200)[0], 3, replace=False).astype(np.uint64)
        heavy = np.concatenate([(h << np.uint64(32)) | rng.permutation(n)[:1100].astype(np.uint64) for h in hub])
        light = (rng.integers(0, n, n_ins).astype(np.uint64) << np.uint64(32)) | rng.integers(0, n, n_ins).astype(
            np.uint64)
        ins = np.concatenate([heavy, light])
        dels = keys[rng.choice(len(keys), min(n_del, len(keys)), replace=False)]
        G.update_batch(*split(ins), *split(dels))
        keys = oracle.set_difference(oracle.set_union(keys, oracle.sort_unique(ins)), oracle.sort_unique(dels))
        assert_set_model(G, keys, n, f"round {rnd}")
        st = G.stats()
        assert st["last_del_keys"] == len(oracle.sort_unique(dels)), f"round {rnd}"
    st = G.stats()
    assert st["promotions"] > st0["promotions"]
    assert st["pma_windows"] + st["pma_rebuilds"] > st0["pma_windows"] + st0["pma_rebuilds"]
    G.close()
