"""AdjacencyGraph input (include/io_util.h:112-145) through the C ABI.  The
reader is host code, so these run without a GPU; the load-into-graph test
is marked gpu."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg
from terrace_amd._lib import lib


def test_reference_style_file(tmp_path):
    # vertex 0 -> {1, 2}, vertex 1 -> {0}, vertex 2 -> {0}
    p = tmp_path / "tiny.adj"
    p.write_text("AdjacencyGraph\n3\n4\n0\n2\n3\n1\n2\n0\n0\n")
    n, s, d = tg.read_adjacency_graph(p)
    assert n == 3
    assert s.tolist() == [0, 0, 1, 2] and d.tolist() == [1, 2, 0, 0]


def test_round_trip_rmat(tmp_path):
    n = 1 << 10
    keys = oracle.build_keys(*oracle.rmat_edges(n, 77, 0.57, 0.19, 0.19, 5000))
    src = (keys >> np.uint64(32)).astype(np.uint32)
    dst = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    p = tmp_path / "g.adj"
    tg.write_adjacency_graph(p, n, src, dst)
    n2, s, d = tg.read_adjacency_graph(p)
    assert n2 == n and np.array_equal(s, src) and np.array_equal(d, dst)
    # isolated trailing vertices and an empty graph
    p2 = tmp_path / "empty.adj"
    tg.write_adjacency_graph(p2, 5, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    n3, s3, d3 = tg.read_adjacency_graph(p2)
    assert n3 == 5 and len(s3) == 0 and len(d3) == 0


@pytest.mark.parametrize("text, why", [
    ("WeightedAdjacencyGraph\n1\n0\n0\n", "missing header"),
    ("AdjacencyGraph\n3\n4\n0\n2\n3\n1\n2\n0\n", "length"),          # one id short
    ("AdjacencyGraph\n2\n1\n0\n1\n1\n7\n", "more than"),              # extra entry
    ("AdjacencyGraph\n2\n1\n0\n1\n5\n", "out of range"),              # id >= n
    ("AdjacencyGraph\n2\n2\n0\n3\n1\n0\n", "offsets"),                # offset beyond m
    ("AdjacencyGraph\n2\nx\n", "missing n or m"),
])
def test_malformed_files_are_rejected(tmp_path, text, why):
    p = tmp_path / "bad.adj"
    p.write_text(text)
    with pytest.raises(RuntimeError) as e:
        tg.read_adjacency_graph(p)
    assert why in str(e.value)


def test_missing_file_is_rejected(tmp_path):
    with pytest.raises(RuntimeError):
        tg.read_adjacency_graph(tmp_path / "nope.adj")


@pytest.mark.gpu
def test_load_into_graph(tmp_path):
    n = 1 << 12
    print("[gpu-test] adjacency load", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 79, 0.57, 0.19, 0.19, 60000))
    src = (keys >> np.uint64(32)).astype(np.uint32)
    dst = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    p = tmp_path / "g.adj"
    tg.write_adjacency_graph(p, n, src, dst)
    G = tg.Graph.from_adjacency_graph(p)
    assert G.get_num_vertices() == n
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(keys, n)
    assert np.array_equal(off, ref.off) and np.array_equal(nbr, ref.nbr[:ref.m])
    G.close()
