"""GPU test of the C++ drop-in: the graph_bm driver, written against the
graphstore shim (include/terrace/graphstore.hpp) with the algorithm
restatements of terrace_amd/host/algorithms.hpp, checked against the
oracle on the same rMAT graph."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "terrace_amd", "host", "graph_bm")


@pytest.mark.parametrize("scale,m0,probs", [(12, 40000, (0.5, 0.1, 0.1)), (12, 120000, (0.57, 0.19, 0.19))])
def test_graph_bm_driver_matches_oracle(tmp_path, scale, m0, probs):
    assert os.path.exists(DRIVER), "build the driver: python -m terrace_amd.build"
    a, b, c = probs
    prefix = str(tmp_path / "out")
    print(f"[gpu-test] graph_bm scale={scale} m0={m0} probs={probs}", flush=True)
    r = subprocess.run([DRIVER, "-s", str(scale), "-m", str(m0), "-a", str(a), "-b", str(b), "-c", str(c),
                        "-src", "9", "-rounds", "1", "-TC", "1", "-SSSP", "1", "-batches", "10,1000", "-trials", "2",
                        "-o", prefix], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert re.search(r"AVG\ttest=BFS\ttime=", r.stdout)
    assert re.search(r"batch_size = 1000, average insert: [0-9.]+, throughput", r.stdout)
    n = 1 << scale
    keys = oracle.build_keys(*oracle.rmat_edges(n, oracle.aspen_ith_rand(100) & 0xFFFFFFFF, a, b, c, m0))
    g = oracle.CSR(keys, n)
    lv = oracle.bfs_levels(g, 9)
    parents = np.fromfile(prefix + ".bfs", np.int32)
    assert oracle.bfs_check_parents(g, 9, parents, lv) == 0
    assert oracle.rel_l1(np.fromfile(prefix + ".pr", np.float64), oracle.pagerank(g, 10)) <= 1e-9
    assert np.array_equal(np.fromfile(prefix + ".cc", np.uint32), oracle.cc(g))
    bc = np.fromfile(prefix + ".bc", np.float64)
    ref = oracle.bc(g, 9)
    reach = lv != np.iinfo(np.uint32).max
    assert oracle.rel_l1(bc[reach], ref[reach]) <= 1e-9 and np.all(np.isnan(bc[~reach]))
    assert int(np.fromfile(prefix + ".tc", np.uint64)[0]) == oracle.tc(g)
    assert np.array_equal(np.fromfile(prefix + ".sssp", np.int32), oracle.sssp_bf(g, 9))


def sym_weight(keys):
    s = (keys >> np.uint64(32)).astype(np.uint64)
    d = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    a, b = np.minimum(s, d), np.maximum(s, d)
    h = (a * np.uint64(2654435761) + b * np.uint64(40503)) >> np.uint64(7)
    return (h % np.uint64(1000) + np.uint64(1)).astype(np.uint32)


def test_graph_bm_driver_weighted(tmp_path):
    """The driver on a weighted graph (the WEIGHTED build): the shim's
    weighted add_edge_batch, then SSSP_BF with the weights as lengths."""
    assert os.path.exists(DRIVER), "build the driver: python -m terrace_amd.build"
    scale, m0 = 12, 60000
    prefix = str(tmp_path / "w")
    print("[gpu-test] graph_bm weighted", flush=True)
    r = subprocess.run([DRIVER, "-s", str(scale), "-m", str(m0), "-a", "0.57", "-b", "0.19", "-c", "0.19",
                        "-src", "9", "-rounds", "1", "-SSSP", "1", "-weighted", "1", "-o", prefix],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    n = 1 << scale
    keys = oracle.build_keys(*oracle.rmat_edges(n, oracle.aspen_ith_rand(100) & 0xFFFFFFFF, 0.57, 0.19, 0.19, m0))
    g = oracle.CSR(keys, n)
    want = oracle.sssp_bf_weighted(g, sym_weight(keys), 9)
    assert np.array_equal(np.fromfile(prefix + ".sssp", np.int32), want)
    # the unweighted algorithms ignore the weights
    assert np.array_equal(np.fromfile(prefix + ".cc", np.uint32), oracle.cc(g))
