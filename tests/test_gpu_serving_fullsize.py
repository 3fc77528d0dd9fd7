"""GPU tests of the serving workload at full size: skewed rMAT scale 26
(67,108,864 vertex ids, a=.57 b=.19 c=.19), 775M generated edges,
symmetrized (about 1.52B directed edges) resident in HBM.

* The device-built graph's adjacency equals the oracle generator's key set.
* A 40-request Poisson stream (update batches of 10..1e6 edges, log-uniform,
  90% inserts of fresh rMAT edges / 10% deletes of existing edges, both
  directions; queries 60% BFS, 20% CC, 10% PageRank 10 iterations, 10% local
  triangle count) runs through the admission-mode server (queries pooled and
  run on the version current at admission).  Every result equals the same
  request run alone, one at a time, on the version it reports (the plain
  server replaying the admission server's order on a second graph).
* Three BFS results and one CC result equal the oracle on their versions
  (the set model of the update batches applied in arrival order)."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
N = 1 << 26
M0 = 775_000_000
PROBS = (0.57, 0.19, 0.19)
REQUESTS = 40


def say(*a):
    print("[gpu-test]", *a, flush=True)


def device_graph():
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    s, d, w = tg.rmat_device(N, seed, *PROBS, M0, nominal=True, symmetric=True, drop_loops=True)
    G = tg.Graph(N)
    G.add_edge_batch_dev(s, d, w)
    s.close(), d.close()
    return G


def as_keys(s, d):
    return (np.asarray(s, np.uint64) << np.uint64(32)) | np.asarray(d, np.uint64)


@pytest.fixture(scope="module")
def serving():
    seed = oracle.aspen_ith_rand(100) & 0xFFFFFFFF
    say("serving workload: oracle graph")
    s, d = oracle.rmat_edges(N, seed, *PROBS, M0)
    g0 = oracle.CSR.from_edges(N, s, d)  # symmetrized, self-loops dropped
    del s, d
    say("serving workload: device build", g0.m)
    st = {"g0": g0, "G1": device_graph()}
    yield st
    for k in ("G1", "G2"):
        if k in st:
            st[k].close()


def test_serving_graph_equals_oracle(serving):
    g0 = serving["g0"]
    off, nbr = serving["G1"].export_adjacency()
    assert len(nbr) == g0.m > 1_400_000_000
    assert np.array_equal(off, g0.off), "offsets"
    assert np.array_equal(nbr, g0.nbr[:g0.m]), "neighbours"
    serving["active"] = np.nonzero(g0.degrees())[0].astype(np.uint32)


def build_stream(g0, active, rng):
    gaps = rng.exponential(1.0 / 100.0, REQUESTS)
    arrivals = np.cumsum(gaps)
    kinds = ["U"] * 12 + ["BFS"] * 16 + ["CC"] * 5 + ["PR"] * 3 + ["LTC"] * 4
    rng.shuffle(kinds)
    reqs, t_gen = [], 0
    for i, k in enumerate(kinds):
        if k == "U":
            size = int(round(10 ** rng.uniform(1, 6)))
            s, d = oracle.rmat_edges(N, oracle.aspen_ith_rand(7000 + t_gen) & 0xFFFFFFFF, *PROBS,
                                     max(1, int(size * 0.9) // 2))
            t_gen += 1
            keep = s != d
            s, d = s[keep], d[keep]
            ins = (np.concatenate([s, d]), np.concatenate([d, s]))
            e = rng.integers(0, g0.m, max(1, int(size * 0.1) // 2)).astype(np.uint64)
            ps = (np.searchsorted(g0.off, e, side="right") - 1).astype(np.uint32)  # existing edges
            pd = g0.nbr[e]
            dels = (np.concatenate([ps, pd]), np.concatenate([pd, ps]))
            reqs.append(("U", arrivals[i], ins, dels))
        elif k in ("BFS", "LTC"):
            reqs.append((k, arrivals[i], int(active[rng.integers(0, len(active))])))
        else:
            reqs.append((k, arrivals[i]))
    return reqs


def submit(S, r, arrival):
    if r[0] == "U":
        return S.submit(tg.REQ_UPDATE, arrival, ins=r[2], dels=r[3])
    if r[0] == "BFS":
        return S.submit(tg.REQ_BFS, arrival, src=r[2])
    if r[0] == "CC":
        return S.submit(tg.REQ_CC, arrival)
    if r[0] == "PR":
        return S.submit(tg.REQ_PR, arrival, iters=10)
    return S.submit(tg.REQ_LTC, arrival, src=r[2])


def test_serving_stream_alone_vs_mix(serving):
    rng = np.random.default_rng(2024)
    reqs = build_stream(serving["g0"], serving["active"], rng)
    serving["reqs"] = reqs
    say("serving workload: admission-mode stream")
    S = tg.Server(serving["G1"], mode="admission", mem_budget=160 << 30, keep_results=True)
    ids = [submit(S, r, r[1]) for r in reqs]
    S.run()
    res = [S.result(i) for i in ids]
    assert all(r["status"] == 0 for r in res)
    # keep the data of the queries checked against the oracle below
    bfs = [i for i, r in enumerate(reqs) if r[0] == "BFS"]
    cc = [i for i, r in enumerate(reqs) if r[0] == "CC"]
    serving["bfs_data"] = {i: S.result_data(ids[i], np.uint32) for i in bfs[:3]}
    serving["cc_data"] = {cc[0]: S.result_data(ids[cc[0]], np.uint32)}
    serving["res"] = res
    summ = S.summary()
    S.close()
    assert summ["groups"] < REQUESTS  # queries were pooled
    # updates apply in arrival order: the k-th update request makes version k
    nu = 0
    for r, x in zip(reqs, res):
        if r[0] == "U":
            nu += 1
            assert x["version"] == nu
    say("serving workload: replay one request at a time on each version")
    order = sorted(range(REQUESTS), key=lambda i: (res[i]["version"], 0 if reqs[i][0] == "U" else 1, i))
    serving.pop("G1").close()  # HBM for the replay's graph
    serving["G2"] = device_graph()
    P = tg.Server(serving["G2"], mode="plain", mem_budget=160 << 30)
    pids = [submit(P, reqs[i], 0.0) for i in order]
    P.run()
    for pos, i in enumerate(order):
        a, b = P.result(pids[pos]), res[i]
        kind = reqs[i][0]
        assert a["status"] == 0 and a["version"] == b["version"], (i, kind)
        if kind in ("BFS", "CC"):
            assert a["digest"] == b["digest"], (i, kind)
        elif kind == "LTC":
            assert a["value"] == b["value"], (i, kind)
        elif kind == "PR":
            assert abs(a["fvalue"] - b["fvalue"]) <= 1e-9 * abs(a["fvalue"]), (i, kind)
    P.close()


def state_at(g0, reqs, version):
    """The set model after the first `version` update requests (each one's
    inserts, then its deletes): the last op per edge decides."""
    ks, ops = [], []
    seen = 0
    for r in reqs:
        if r[0] != "U":
            continue
        if seen == version:
            break
        seen += 1
        ks += [as_keys(*r[2]), as_keys(*r[3])]
        ops += [np.zeros(len(r[2][0]), np.uint8), np.ones(len(r[3][0]), np.uint8)]
    if not ks:
        return g0
    k, o = np.concatenate(ks), np.concatenate(ops)
    order = np.argsort(k, kind="stable")
    k, o = k[order], o[order]
    last = np.ones(len(k), bool)
    last[:-1] = k[1:] != k[:-1]
    return g0.updated(k[last & (o == 0)], k[last & (o == 1)])


def oracle_graph_at(serving, v):
    if serving.get("version") != v:
        serving.pop("g", None)
        serving["g"] = state_at(serving["g0"], serving["reqs"], v)
        serving["version"] = v
    return serving["g"]


@pytest.mark.parametrize("j", [0, 1, 2])
def test_serving_bfs_matches_oracle_on_version(serving, j):
    reqs, res = serving["reqs"], serving["res"]
    i = sorted(serving["bfs_data"])[j]
    v = res[i]["version"]
    say("serving workload: oracle BFS at version", v)
    g = oracle_graph_at(serving, v)
    assert np.array_equal(serving["bfs_data"][i], oracle.bfs_levels(g, reqs[i][2])), (i, v)


def test_serving_cc_matches_oracle_on_version(serving):
    reqs, res = serving["reqs"], serving["res"]
    for i, labels in serving["cc_data"].items():
        v = res[i]["version"]
        say("serving workload: oracle CC at version", v)
        assert np.array_equal(labels, oracle.cc(oracle_graph_at(serving, v))), (i, v)
