"""GPU tests of the update path's error handling and of the batch sizes where
its kernels switch form.

* A failure in a batch's preparation (a look-back wait that gave up) is
  reported before the batch writes anything: the call fails with
  TG_EINTERNAL, the adjacency is unchanged and later batches work.
* A failure after the batch began writing (an allocation for the graph's
  growth) marks the graph invalid: that call and every later call on the
  graph fail with TG_EINTERNAL.
Both failures are injected from the host (tg_debug_fail_next); every kernel
runs to its normal end.

* Batches of exactly 2^18 unique keys (1024 preparation blocks: the last
  launch size whose look-back orders blocks by index, and the bucket sort's
  largest batch) and 2^18 + 1 (ticket order, radix sort, separate lookups),
  inserted and deleted, against the set model.
"""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import TerraceError

pytestmark = pytest.mark.gpu
TG_EINTERNAL = -4


def say(*a):
    print("[gpu-test]", *a, flush=True)


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def adjacency(G):
    off, nbr = G.export_adjacency()
    return off.copy(), nbr.copy()


def assert_set_model(G, keys, n, what):
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off[1:] - off[:-1], ref.degrees().astype(off.dtype)), what
    assert np.array_equal(nbr, ref.nbr[:ref.m]), what
    assert G.stats()["pma_elements"] == G.tier_counts()["pma_edges"], what


def base_graph(n, m0, seed):
    keys = oracle.build_keys(*oracle.rmat_edges(n, seed, 0.57, 0.19, 0.19, m0))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    return G, keys


@pytest.mark.parametrize("size", [1000, 50_000, 300_000])
def test_prep_failure_leaves_graph_unchanged(size):
    """Injected look-back error in the preparation of an insert, a delete
    and a mixed batch (small, bucket and radix sizes; up to 2^18 edges the
    batch's update kernels are already launched, behind the device-side
    gate): TG_EINTERNAL, no change, and the same batches then apply
    normally."""
    n = 1 << 14
    say("prep failure", size)
    G, keys = base_graph(n, 150_000, 11)
    before = adjacency(G)
    rng = np.random.default_rng(size)
    ins = oracle.sort_unique((rng.integers(0, n, size).astype(np.uint64) << np.uint64(32))
                             | rng.integers(0, n, size).astype(np.uint64))
    dels = oracle.sort_unique(keys[rng.choice(len(keys), min(size, len(keys) // 2), replace=False)])
    calls = (("insert", lambda: G.add_edge_batch(*split(ins))),
             ("delete", lambda: G.remove_edge_batch(*split(dels))),
             ("mixed", lambda: G.update_batch(*split(ins[: size // 2]), *split(dels[: size // 2]))))
    for kind, call in calls:
        tg.lib().tg_debug_fail_next(G.handle, 1)
        with pytest.raises(TerraceError) as e:
            call()
        assert e.value.code == TG_EINTERNAL, kind
        after = adjacency(G)
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1]), kind
    G.add_edge_batch(*split(ins))
    keys = oracle.set_union(keys, ins)
    G.remove_edge_batch(*split(dels))
    keys = oracle.set_difference(keys, dels)
    assert_set_model(G, keys, n, "after the failed attempts")
    G.close()


@pytest.mark.parametrize("kind", ["insert", "delete", "mixed"])
def test_failure_after_first_write_marks_graph_invalid(kind):
    n = 1 << 12
    say("mid-batch failure", kind)
    G, keys = base_graph(n, 40_000, 5)
    batch = keys[:5000] if kind == "delete" else oracle.build_keys(
        *oracle.rmat_edges(n, 77, 0.5, 0.1, 0.1, 5000))
    tg.lib().tg_debug_fail_next(G.handle, 2)
    with pytest.raises(TerraceError) as e:
        if kind == "mixed":
            G.update_batch(*split(batch), *split(keys[:500]))
        else:
            (G.add_edge_batch if kind == "insert" else G.remove_edge_batch)(*split(batch))
    assert e.value.code == TG_EINTERNAL
    for call in (lambda: G.add_edge_batch(*split(batch)),
                 lambda: G.export_adjacency(),
                 lambda: tg.BFS_with_edge_map(G, 0)):
        with pytest.raises(TerraceError) as e:
            call()
        assert e.value.code == TG_EINTERNAL
    G.close()


@pytest.mark.parametrize("count", [1 << 18, (1 << 18) + 1])
def test_batches_at_the_block_order_boundary(count):
    """Exactly `count` unique keys per batch, inserted (mostly new keys, some
    present) and deleted (mostly present, some absent)."""
    n = 1 << 16
    say("boundary batch", count)
    G, keys = base_graph(n, 300_000, 3)
    rng = np.random.default_rng(count)
    fresh = oracle.sort_unique((rng.integers(0, n, 2 * count).astype(np.uint64) << np.uint64(32))
                               | rng.integers(0, n, 2 * count).astype(np.uint64))
    old = keys[rng.choice(len(keys), 1000, replace=False)]
    ins = oracle.sort_unique(np.concatenate([fresh[:count], old]))[:count]
    assert len(ins) == count
    shuffled = ins.copy()
    rng.shuffle(shuffled)  # the batch in any order; the set model takes it sorted
    G.add_edge_batch(*split(shuffled))
    keys = oracle.set_union(keys, ins)
    assert_set_model(G, keys, n, f"insert {count}")
    absent = oracle.set_difference(fresh[count:], keys)[:1000]
    dels = oracle.sort_unique(np.concatenate([keys[rng.choice(len(keys), count - len(absent), replace=False)],
                                              absent]))
    assert len(dels) == count
    shuffled = dels.copy()
    rng.shuffle(shuffled)
    G.remove_edge_batch(*split(shuffled))
    keys = oracle.set_difference(keys, dels)
    assert_set_model(G, keys, n, f"delete {count}")
    G.close()
