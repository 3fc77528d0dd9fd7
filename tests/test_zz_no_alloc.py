"""GPU test of tg_count_common_noalloc (staging in a static device array, no
device allocation) against tg_count_common, through the C ABI: two calls in
a row, the second with fewer pairs, results bit-exact; then the limit and
the id check."""
import ctypes

import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import lib, ptr

pytestmark = pytest.mark.gpu
LIMIT = 1 << 20
TG_EINVAL = -1


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def noalloc(G, s, d):
    s, d = np.ascontiguousarray(s, np.uint32), np.ascontiguousarray(d, np.uint32)
    out = np.full(len(s), 0xDEADBEEF, np.uint32)
    rc = lib().tg_count_common_noalloc(G.handle, ptr(s), ptr(d), len(s), ptr(out))
    return rc, out


def test_count_common_noalloc_matches_allocating_entry():
    print("[gpu-test] count_common_noalloc", flush=True)
    n = 1 << 14
    keys = oracle.build_keys(*oracle.rmat_edges(n, 17, 0.57, 0.19, 0.19, 300_000))
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    tiers = G.tiers()
    assert (tiers == 1).any() and (tiers == 2).any()
    rng = np.random.default_rng(11)
    for q in (200_000, 3_000):  # the second call smaller: stale staging must not leak in
        e = rng.integers(0, len(keys), q // 2)
        es, ed = split(keys[e])
        s = np.concatenate([es, rng.integers(0, n, q - len(es))]).astype(np.uint32)
        d = np.concatenate([ed, rng.integers(0, n, q - len(ed))]).astype(np.uint32)
        want = G.count_common_batch(s, d)
        rc, got = noalloc(G, s, d)
        assert rc == 0
        assert np.array_equal(got, want), q
        assert want.max() > 0
    # the limit, checked before any copy (arrays of one pair, q past the limit)
    one = np.zeros(1, np.uint32)
    assert lib().tg_count_common_noalloc(G.handle, ptr(one), ptr(one), LIMIT + 1, ptr(one)) == TG_EINVAL
    # an id out of range is reported, and the next call starts clean
    rc, _ = noalloc(G, [0, n], [1, 2])
    assert rc == TG_EINVAL
    rc, got = noalloc(G, [5, 6], [7, 8])
    assert rc == 0 and np.array_equal(got, G.count_common_batch([5, 6], [7, 8]))
    G.close()
