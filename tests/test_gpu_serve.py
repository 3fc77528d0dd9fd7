"""GPU tests of the batching / serving layer: multi-source BFS equals
single-source BFS, batched local triangle counts equal the oracle, and every
request's result is the same alone (plain server) and in the mix (batched
server), on the same version of the graph."""
import numpy as np
import pytest

import oracle
from terrace_amd import graph as tg

pytestmark = pytest.mark.gpu
UNREACHED = np.iinfo(np.uint32).max


def split(keys):
    keys = np.asarray(keys, np.uint64)
    return ((keys >> np.uint64(32)).astype(np.uint32),
            (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def make_graph(n, keys):
    G = tg.Graph(n)
    G.add_edge_batch(*split(keys))
    return G


@pytest.mark.parametrize("skew", [False, True])
def test_bfs_multi_matches_single(skew):
    n = 1 << 12
    a, b, c = (0.57, 0.19, 0.19) if skew else (0.5, 0.1, 0.1)
    print("[gpu-test] bfs_multi", skew, flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 41, a, b, c, 60000))
    G = make_graph(n, keys)
    g = oracle.CSR(keys, n)
    rng = np.random.default_rng(5)
    srcs = rng.integers(0, n, 64).astype(np.uint32)
    srcs[1] = srcs[0]  # duplicate source in two lanes
    srcs[2] = int(np.argmax(g.degrees()))
    lv = tg.bfs_multi(G, srcs)
    for j in range(len(srcs)):
        want = oracle.bfs_levels(g, int(srcs[j]))
        assert np.array_equal(lv[j], want), j
    # fewer lanes than 64 and a single lane
    lv3 = tg.bfs_multi(G, srcs[:3])
    assert np.array_equal(lv3, lv[:3])
    _, single = tg.BFS_with_edge_map(G, int(srcs[5]), with_levels=True)
    assert np.array_equal(tg.bfs_multi(G, srcs[5:6])[0], single)
    G.close()


@pytest.mark.parametrize("skew", [False, True])
def test_bc_multi_matches_single(skew):
    """64 BC sources in one traversal: every lane equals the single-source
    restatement bit for bit (sums in sorted neighbour order, NaN where
    unreachable) and the single-source GPU BC within 1e-9."""
    n = 1 << 12
    a, b, c = (0.57, 0.19, 0.19) if skew else (0.5, 0.1, 0.1)
    print("[gpu-test] bc_multi", skew, flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 59, a, b, c, 60000))
    G = make_graph(n, keys)
    g = oracle.CSR(keys, n)
    rng = np.random.default_rng(17)
    srcs = rng.integers(0, n, 64).astype(np.uint32)
    srcs[1] = srcs[0]
    srcs[2] = int(np.argmax(g.degrees()))
    srcs[3] = int(np.argmin(g.degrees()))  # possibly isolated
    got = tg.bc_multi(G, srcs)
    for j in range(len(srcs)):
        want = oracle.bc(g, int(srcs[j]))
        np.testing.assert_array_equal(got[j], want, err_msg=f"lane {j}")
    for j in (0, 2, 5):
        single = tg.BC(G, int(srcs[j]))
        fin = np.isfinite(single)
        assert np.array_equal(fin, np.isfinite(got[j]))
        assert oracle.rel_l1(got[j][fin], single[fin]) <= 1e-9
    # fewer lanes
    np.testing.assert_array_equal(tg.bc_multi(G, srcs[:5]), got[:5])
    G.close()


def test_local_triangles_match_oracle():
    n = 1 << 12
    print("[gpu-test] ltc", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 43, 0.57, 0.19, 0.19, 80000))
    G = make_graph(n, keys)
    g = oracle.CSR(keys, n)
    deg = g.degrees()
    verts = np.concatenate([np.argsort(deg)[-5:], np.arange(0, n, 97), [np.argmin(deg)]]).astype(np.uint32)
    got = tg.local_triangles(G, verts)
    want = np.array([oracle.ltc(g, int(v)) for v in verts], np.uint64)
    assert np.array_equal(got, want)
    # the per-vertex counts sum to three times the triangle count
    allv = np.arange(n, dtype=np.uint32)
    assert int(tg.local_triangles(G, allv).sum()) == 3 * oracle.tc(g)
    G.close()


def build_stream(n, keys, rng):
    """Interleaved updates and queries; conflicting ops on one edge across
    batches (insert, delete, re-insert)."""
    e = keys[rng.choice(len(keys), 30, replace=False)]
    reqs = []
    t = 0.0

    def upd(ins_keys, del_keys):
        i_s, i_d = split(ins_keys)
        d_s, d_d = split(del_keys)
        return ("U", (np.concatenate([i_s, i_d]), np.concatenate([i_d, i_s])),
                (np.concatenate([d_s, d_d]), np.concatenate([d_d, d_s])))

    for r in range(5):
        fresh = oracle.build_keys(*oracle.rmat_edges(n, 500 + r, 0.57, 0.19, 0.19, 300))
        reqs.append(upd(fresh, e[r * 5:(r + 1) * 5]))
        # same edges deleted by one batch and re-inserted by the next
        reqs.append(upd(e[r * 5:(r + 1) * 5][:2], np.zeros(0, np.uint64)))
        for q in range(12):
            kind = rng.choice(["BFS", "BFS", "BFS", "CC", "PR", "LTC"])
            reqs.append((kind, int(rng.integers(0, n))))
    out = []
    for req in reqs:
        t += 1e-5
        out.append((t, req))
    return out


def submit_all(S, stream):
    ids = []
    for t, req in stream:
        if req[0] == "U":
            ids.append(S.submit(tg.REQ_UPDATE, t, ins=req[1], dels=req[2]))
        elif req[0] == "BFS":
            ids.append(S.submit(tg.REQ_BFS, t, src=req[1]))
        elif req[0] == "CC":
            ids.append(S.submit(tg.REQ_CC, t))
        elif req[0] == "PR":
            ids.append(S.submit(tg.REQ_PR, t, iters=10))
        else:
            ids.append(S.submit(tg.REQ_LTC, t, src=req[1]))
    return ids


def test_server_alone_vs_mix():
    n = 1 << 12
    print("[gpu-test] server alone vs mix", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 47, 0.57, 0.19, 0.19, 50000))
    rng = np.random.default_rng(11)
    stream = build_stream(n, keys, rng)
    G1, G2 = make_graph(n, keys), make_graph(n, keys)
    plain = tg.Server(G1, batched=False, keep_results=True)
    batched = tg.Server(G2, batched=True, keep_results=True)
    ids1, ids2 = submit_all(plain, stream), submit_all(batched, stream)
    plain.run()
    batched.run()
    # versions: the oracle state after each update batch, in order
    states = [keys]
    for t, req in stream:
        if req[0] == "U":
            k = states[-1]
            ins = oracle.sort_unique((req[1][0].astype(np.uint64) << np.uint64(32)) | req[1][1])
            dels = oracle.sort_unique((req[2][0].astype(np.uint64) << np.uint64(32)) | req[2][1])
            k = oracle.set_difference(oracle.set_union(k, ins), dels)
            states.append(k)
    groups = set()
    for (t, req), a, b in zip(stream, ids1, ids2):
        ra, rb = plain.result(a), batched.result(b)
        assert ra["status"] == 0 and rb["status"] == 0
        assert ra["version"] == rb["version"], (req[0], ra, rb)
        groups.add(rb["group"])
        if req[0] == "U":
            continue
        g = oracle.CSR(states[ra["version"]], n)
        if req[0] == "BFS":
            la, lb = plain.result_data(a, np.uint32), batched.result_data(b, np.uint32)
            assert np.array_equal(la, lb) and np.array_equal(la, oracle.bfs_levels(g, req[1]))
            assert ra["digest"] == rb["digest"]
        elif req[0] == "CC":
            la, lb = plain.result_data(a, np.uint32), batched.result_data(b, np.uint32)
            assert np.array_equal(la, lb) and np.array_equal(la, oracle.cc(g))
        elif req[0] == "PR":
            pa, pb = plain.result_data(a, np.float64), batched.result_data(b, np.float64)
            assert oracle.rel_l1(pb, pa) <= 1e-9 and oracle.rel_l1(pa, oracle.pagerank(g, 10)) <= 1e-9
        else:
            assert ra["value"] == rb["value"] == oracle.ltc(g, req[1])
    # batching actually grouped requests
    assert len(groups) < len(stream)
    # both graphs end in the same state, equal to the oracle
    for G in (G1, G2):
        off, nbr = G.export_adjacency()
        ref = oracle.CSR(states[-1], n)
        assert np.array_equal(nbr, ref.nbr[:ref.m]) and np.array_equal(off, ref.off)
    s1, s2 = plain.summary(), batched.summary()
    assert s1["rejected"] == 0 and s2["rejected"] == 0
    plain.close(), batched.close(), G1.close(), G2.close()


def oracle_states(keys, stream):
    """The oracle edge set after each update request, in arrival order."""
    states = [keys]
    for t, req in stream:
        if req[0] == "U":
            ins = oracle.sort_unique((req[1][0].astype(np.uint64) << np.uint64(32)) | req[1][1])
            dels = oracle.sort_unique((req[2][0].astype(np.uint64) << np.uint64(32)) | req[2][1])
            states.append(oracle.set_difference(oracle.set_union(states[-1], ins), dels))
    return states


def test_server_snapshot_at_admission():
    """Admission mode: every query's result equals the query run alone on
    the version it reports, and that version holds at least every update
    that arrived before the query."""
    n = 1 << 12
    print("[gpu-test] server snapshot-at-admission", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 53, 0.57, 0.19, 0.19, 50000))
    rng = np.random.default_rng(13)
    stream = build_stream(n, keys, rng)
    # spread arrivals so that groups form at several versions
    stream = [(i * 4e-4, req) for i, (t, req) in enumerate(stream)]
    G = make_graph(n, keys)
    S = tg.Server(G, mode="admission", keep_results=True)
    ids = submit_all(S, stream)
    S.run()
    states = oracle_states(keys, stream)
    seen_updates = 0
    versions, groups = set(), {}
    for (t, req), i in zip(stream, ids):
        r = S.result(i)
        assert r["status"] == 0
        if req[0] == "U":
            seen_updates += 1
            assert r["version"] == seen_updates  # updates apply in arrival order
            continue
        assert seen_updates <= r["version"] <= len(states) - 1, (req, r, seen_updates)
        versions.add(r["version"])
        groups.setdefault(r["group"], set()).add(r["version"])
        g = oracle.CSR(states[r["version"]], n)
        if req[0] == "BFS":
            assert np.array_equal(S.result_data(i, np.uint32), oracle.bfs_levels(g, req[1]))
        elif req[0] == "CC":
            assert np.array_equal(S.result_data(i, np.uint32), oracle.cc(g))
        elif req[0] == "PR":
            assert oracle.rel_l1(S.result_data(i, np.float64), oracle.pagerank(g, 10)) <= 1e-9
        else:
            assert r["value"] == oracle.ltc(g, req[1])
    # one version per group, and queries were pooled
    assert all(len(v) == 1 for v in groups.values())
    assert len(groups) < sum(1 for _, req in stream if req[0] != "U")
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(states[-1], n)
    assert np.array_equal(nbr, ref.nbr[:ref.m]) and np.array_equal(off, ref.off)
    S.close(), G.close()


def test_server_admission_rejects_oversized():
    n = 1 << 12
    print("[gpu-test] server admission", flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 49, 0.5, 0.1, 0.1, 20000))
    G = make_graph(n, keys)
    # budget fits a BFS lane but not a PageRank (24 n bytes)
    S = tg.Server(G, batched=True, mem_budget=20 * n)
    a = S.submit(tg.REQ_BFS, 0.0, src=1)
    b = S.submit(tg.REQ_PR, 0.0, iters=3)
    c = S.submit(tg.REQ_BFS, 0.0, src=2)
    S.run()
    assert S.result(a)["status"] == 0 and S.result(c)["status"] == 0
    assert S.result(b)["status"] == -2
    assert S.summary()["rejected"] == 1
    S.close(), G.close()


def submit_one(S, t, req):
    return submit_all(S, [(t, req)])[0]


@pytest.mark.parametrize("mode", ["plain", "strict", "admission"])
def test_server_live_submission(mode):
    """Requests submitted while the loop runs (tg_server_start ...
    tg_server_finish): a request completes while the input is still open,
    every live request arrives at its submission time, the update requests
    apply in submission order, and every query equals the oracle on the
    version it reports -- exactly the updates submitted before it (plain,
    strict), at least those (admission).  A request with a later arrival
    offset stays pending until then, and finish waits for it."""
    import time
    n = 1 << 12
    print("[gpu-test] server live submission", mode, flush=True)
    keys = oracle.build_keys(*oracle.rmat_edges(n, 61, 0.57, 0.19, 0.19, 50000))
    rng = np.random.default_rng(29)
    stream = build_stream(n, keys, rng)
    G = make_graph(n, keys)
    S = tg.Server(G, mode=mode, keep_results=True)
    pre = 4  # submitted before the loop starts, arriving at its start
    ids = [submit_one(S, 0.0, req) for _, req in stream[:pre]]
    S.start()
    with pytest.raises(tg.TerraceError):
        S.run()  # the loop is already running
    assert S.wait(ids[0], 60.0) == 0  # done while the input is open
    open_done = 0
    for k, (_, req) in enumerate(stream[pre:]):
        ids.append(submit_one(S, 0.0, req))
        if k % 9 == 4:
            time.sleep(0.003)
        if k % 20 == 10:
            assert S.wait(ids[-1], 60.0) == 0
            open_done += 1
    # a request that arrives 0.3 s after its submission
    late_t = S.result(ids[-1])["arrival_s"] + 0.3
    late = S.submit(tg.REQ_CC, late_t)
    assert S.wait(late, 0.0) == 1 and S.result(late)["status"] == 1
    S.finish()
    assert open_done >= 2
    r_late = S.result(late)
    assert r_late["status"] == 0 and r_late["start_s"] >= late_t
    states = oracle_states(keys, stream)
    seen_updates, last_arrival = 0, 0.0
    for (_, req), i in zip(stream, ids):
        r = S.result(i)
        assert r["status"] == 0, (req[0], r)
        assert r["arrival_s"] >= last_arrival  # submission order is arrival order
        last_arrival = r["arrival_s"]
        if req[0] == "U":
            seen_updates += 1
            assert r["version"] == seen_updates
            continue
        if mode == "admission":
            assert seen_updates <= r["version"] <= len(states) - 1, (req, r, seen_updates)
        else:
            assert r["version"] == seen_updates, (req, r, seen_updates)
        g = oracle.CSR(states[r["version"]], n)
        if req[0] == "BFS":
            assert np.array_equal(S.result_data(i, np.uint32), oracle.bfs_levels(g, req[1]))
        elif req[0] == "CC":
            assert np.array_equal(S.result_data(i, np.uint32), oracle.cc(g))
        elif req[0] == "PR":
            assert oracle.rel_l1(S.result_data(i, np.float64), oracle.pagerank(g, 10)) <= 1e-9
        else:
            assert r["value"] == oracle.ltc(g, req[1])
    assert np.array_equal(S.result_data(late, np.uint32), oracle.cc(oracle.CSR(states[-1], n)))
    off, nbr = G.export_adjacency()
    ref = oracle.CSR(states[-1], n)
    assert np.array_equal(nbr, ref.nbr[:ref.m]) and np.array_equal(off, ref.off)
    assert S.summary()["rejected"] == 0
    S.close(), G.close()
