"""Write the golden fixtures under tests/golden/ from the reference build.

Runs ``oracle/_ref/ref_harness`` (the reference's own graph.h / PMA.cc /
btree.h / BFS.h / Pagerank.h / Components.h / BC.h / TC.h, compiled by
``oracle/Makefile``) on small seeded inputs and stores inputs and outputs as
``.npz`` data.  Run it in a container where the reference is mounted:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures:
  rmat_kat.npz     rMat<uint32_t>(i) pairs from the reference generator
                   (rmat_util.h:81-115, divisor INT32_MAX) for several
                   (n, seed, a, b, c)
  graph_*.npz      a static build plus insert batches: inputs (sorted unique
                   directed keys), the adjacency after every step (degree
                   array + neighbours in NeighborIterator order) and the
                   B-tree tier membership after every step, then BFS
                   parents, PR (10 iterations), CC labels, BC and the
                   triangle count on the final graph.

Deletes are not taken from the reference: its remove_edge shifts inline
slots with an overlapping memcpy (graph.h:700,735), so the post-delete state
is pinned by the set model instead.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

HARNESS = oracle.REF_HARNESS


def write_edges(path, n, keys):
    keys = np.asarray(keys, np.uint64)
    pairs = np.empty((len(keys), 2), np.uint32)
    pairs[:, 0] = (keys >> np.uint64(32)).astype(np.uint32)
    pairs[:, 1] = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    with open(path, "wb") as f:
        np.array([n, len(keys)], np.uint64).tofile(f)
        pairs.tofile(f)


def rmat_keys(n, seed, a, b, c, count, symmetric=True):
    s, d = oracle.rmat_edges(n, seed, a, b, c, count, nominal=True)
    return oracle.build_keys(s, d, symmetric=symmetric)


def make_kat():
    out = {}
    cases = [
        (1 << 20, oracle.aspen_ith_rand(100) & 0xFFFFFFFF, 0.5, 0.1, 0.1, 2000),
        (1 << 12, 7, 0.57, 0.19, 0.19, 2000),
        (1 << 16, 12345, 0.25, 0.25, 0.25, 2000),
        (1 << 22, 3141592653, 0.45, 0.15, 0.15, 2000),
    ]
    with tempfile.TemporaryDirectory() as td:
        for k, (n, seed, a, b, c, count) in enumerate(cases):
            path = os.path.join(td, "r.bin")
            subprocess.run([HARNESS, "rmat", str(n), str(seed), repr(a), repr(b), repr(c),
                            str(count), path], check=True)
            pairs = np.fromfile(path, np.uint32).reshape(-1, 2)
            out[f"params{k}"] = np.array([n, seed, count], np.uint64)
            out[f"probs{k}"] = np.array([a, b, c], np.float64)
            out[f"pairs{k}"] = pairs
    np.savez_compressed(os.path.join(HERE, "rmat_kat.npz"), **out)


def make_graph(name, n, base_keys, batches, src, iters=10):
    with tempfile.TemporaryDirectory() as td:
        gpath = os.path.join(td, "g.bin")
        write_edges(gpath, n, base_keys)
        bpaths = []
        for k, b in enumerate(batches):
            p = os.path.join(td, f"b{k}.bin")
            write_edges(p, n, b)
            bpaths.append(p)
        prefix = os.path.join(td, "out")
        env = dict(os.environ, OMP_NUM_THREADS="4")
        subprocess.run([HARNESS, "run", gpath, prefix, str(src), str(iters)] + bpaths,
                       check=True, env=env, stdout=subprocess.DEVNULL)
        data = {"n": np.array([n], np.uint64), "src": np.array([src], np.uint32),
                "iters": np.array([iters], np.uint32),
                "base": np.asarray(base_keys, np.uint64),
                "nbatches": np.array([len(batches)], np.uint32)}
        for k, b in enumerate(batches):
            data[f"batch{k}"] = np.asarray(b, np.uint64)
        for k in range(len(batches) + 1):
            raw = np.fromfile(prefix + f".adj{k}", np.uint32)
            data[f"deg{k}"] = raw[:n]
            data[f"nbrs{k}"] = raw[n:]
            data[f"btree{k}"] = np.fromfile(prefix + f".btree{k}", np.uint8)
        data["bfs_parents"] = np.fromfile(prefix + ".bfs", np.int32)
        data["pr"] = np.fromfile(prefix + ".pr", np.float64)
        data["cc"] = np.fromfile(prefix + ".cc", np.uint32)
        data["bc"] = np.fromfile(prefix + ".bc", np.float64)
        data["tc"] = np.fromfile(prefix + ".tc", np.uint64)
        data["sssp"] = np.fromfile(prefix + ".sssp", np.int32)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **data)
    print(name, "n", n, "m", len(base_keys), "tc", int(data["tc"][0]))


def main():
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference first: make -C oracle ref")
    make_kat()
    # inline + PMA tiers: scale 10, .5/.1/.1
    n = 1 << 10
    base = rmat_keys(n, 11, 0.5, 0.1, 0.1, 6000)
    b1 = rmat_keys(n, 12, 0.5, 0.1, 0.1, 800)
    b2 = rmat_keys(n, 13, 0.5, 0.1, 0.1, 3000)
    make_graph("graph_s10", n, base, [b1, b2], src=9)
    # all three tiers: scale 12 skewed .57/.19/.19 (hubs above degree 1024),
    # batches skewed toward the hubs by reusing the same distribution
    n = 1 << 12
    base = rmat_keys(n, 21, 0.57, 0.19, 0.19, 60000)
    b1 = rmat_keys(n, 22, 0.57, 0.19, 0.19, 3000)
    b2 = rmat_keys(n, 23, 0.57, 0.19, 0.19, 20000)
    b3 = rmat_keys(n, 24, 0.57, 0.19, 0.19, 500)
    make_graph("graph_skew12", n, base, [b1, b2, b3], src=0)
    # small vertex gaining many neighbours in one batch: PMA -> B-tree
    # promotion mid-life, plus an isolated BFS source elsewhere
    n = 1 << 11
    star = np.array([(np.uint64(5) << np.uint64(32)) | np.uint64(v) for v in range(6, 40)]
                    + [(np.uint64(v) << np.uint64(32)) | np.uint64(5) for v in range(6, 40)],
                    np.uint64)
    base = np.sort(np.unique(np.concatenate([star, rmat_keys(n, 31, 0.5, 0.1, 0.1, 4000)])))
    big = np.array([(np.uint64(5) << np.uint64(32)) | np.uint64(v) for v in range(40, 1500)]
                   + [(np.uint64(v) << np.uint64(32)) | np.uint64(5) for v in range(40, 1500)],
                   np.uint64)
    b1 = np.sort(np.unique(big))
    b2 = rmat_keys(n, 32, 0.5, 0.1, 0.1, 1000)
    make_graph("graph_promote11", n, base, [b1, b2], src=5)


if __name__ == "__main__":
    main()
