"""Benchmark: batched edge updates on the scale-22 rMAT graph, one MI355X.

Workload: rMAT scale 22 (4,194,304 vertex ids), 69M generated edges with
a=0.5 b=0.1 c=0.1 (divisor 2^32), self-loops dropped, symmetrized and
deduplicated (135.9M directed edges), built on the device.  A step is one
mixed update batch of --batch directed edge updates: 90% inserts (fresh
rMAT edges, both directions) and 10% deletes of edges of the graph (both
directions), applied as one mixed call (the inserts, then the deletes).
All batch inputs are generated into HBM before the timed region.

After the timed steps the run also times BFS, PageRank (10 iterations,
FP64), CC, insert batches of 10^3 and 10^7 edges, and reports them under
"extra".  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per mixed update step of 1e5 edges at the default workload,
# summed over the step's kernels from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs scaled by a 1 GiB copy kernel (FETCH_SIZE x 2, WRITE_SIZE
# x 1): profiles/update_step_scale22_hbm_traffic.md.
UPDATE_STEP_TRAFFIC_DEFAULT = 2.051e8
# HBM bytes per PageRank iteration (k_prs_pull + k_prs_fix + k_prs_reduce)
# at the default workload (scale 22, 69M generated edges), from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy kernel
# (FETCH_SIZE x 2, WRITE_SIZE x 1).
PR_ITER_TRAFFIC_DEFAULT = 2.228e9


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def aspen(i: int) -> int:
    # random_aspen().ith_rand(i) = hash64(i) (include/rmat_util.h:57-70)
    u = (i * 3935559000370003845 + 2691343689449507681) & 0xFFFFFFFFFFFFFFFF
    u ^= u >> 21
    u ^= (u << 37) & 0xFFFFFFFFFFFFFFFF
    u ^= u >> 4
    u = (u * 4768777513237032717) & 0xFFFFFFFFFFFFFFFF
    u ^= (u << 20) & 0xFFFFFFFFFFFFFFFF
    u ^= u >> 41
    u ^= (u << 5) & 0xFFFFFFFFFFFFFFFF
    return u


def make_batches(tg, G, n, scale_a, batch, count, seed0, rng, off, nbr):
    """Pre-generate `count` mixed batches in HBM: (ins_src, ins_dst, n_ins,
    del_src, del_dst, n_del)."""
    a, b, c = scale_a
    out = []
    m = len(nbr)
    for t in range(count):
        n_ins_pairs = max(1, int(batch * 0.9) // 2)
        s, d, w = tg.rmat_device(n, aspen(seed0 + t) & 0xFFFFFFFF, a, b, c, n_ins_pairs,
                                 nominal=True, symmetric=True, drop_loops=True)
        n_del_pairs = max(1, int(batch * 0.1) // 2)
        idx = rng.integers(0, m, n_del_pairs)
        ds = (np.searchsorted(off, idx, side="right") - 1).astype(np.uint32)
        dd = nbr[idx]
        del_s = np.concatenate([ds, dd]).astype(np.uint32)
        del_d = np.concatenate([dd, ds]).astype(np.uint32)
        bs, bd = tg.DeviceBuffer(del_s.nbytes), tg.DeviceBuffer(del_d.nbytes)
        bs.upload(del_s)
        bd.upload(del_d)
        out.append((s, d, w, bs, bd, len(del_s)))
    return out


def apply_batch(G, bt):
    """One mixed step: the inserts, then the deletes, in one call
    (tg_update_batch_dev)."""
    s, d, w, bs, bd, nd = bt
    G.update_batch_dev(s, d, w, bs, bd, nd)


def step_bytes(st1, st0, b_ins, b_del):
    """Algorithmic HBM bytes of mixed steps: inserts 8b + 4b + 16u (read
    the pairs, write the ids, read and write each source's metadata),
    deletes 8b + 16u, with u the batches' distinct sources as counted on
    the device (tg_stats after minus before the steps)."""
    u_ins = st1["sum_ins_sources"] - st0["sum_ins_sources"]
    u_del = st1["sum_del_sources"] - st0["sum_del_sources"]
    return 12 * b_ins + 16 * u_ins + 8 * b_del + 16 * u_del


def copy_probe():
    """Three copies of 1 GiB between device buffers by the library's plain
    copy kernel (k_copy16): a kernel of known HBM traffic (1 GiB read +
    1 GiB written per launch) that scales the FETCH_SIZE / WRITE_SIZE
    counters."""
    from terrace_amd._lib import DeviceBuffer, lib, check
    nb = 1 << 30
    a, b = DeviceBuffer(nb), DeviceBuffer(nb)
    for _ in range(3):
        check(lib().tg_memcpy_d2d(b.ptr, a.ptr, nb))
    a.close()
    b.close()
    log(f"copy probe: 3 x {nb} bytes")


def cpu_baseline(quick: bool, batch: int):
    """The reference's own update path (oracle/_ref, built from the
    reference's sources) timed on this host on the benchmark's workload: the
    static graph (rMAT scale 22, 69M generated edges, a=.5 b=.1 c=.1,
    symmetrized) and 20 mixed batches of `batch` edges, 90% inserts of fresh
    rMAT edges (both directions, seeds ith_rand(1000 + t) as the GPU
    batches) and 10% deletes of existing edges (both directions).  Timed as
    src/graph_bm.cc:354-378 times them: add_edge_batch (sort outside the
    region) plus the parallel remove_edge loop.  Delete endpoints are used
    once across the batches (remove_edge's overlapping memcpy, graph.h:700,
    corrupts a vertex it has touched before).  Falls back to the oracle's
    set-model update when the reference build is absent."""
    import oracle

    scale, m0, nb = (18, 4_000_000, 5) if quick else (22, 69_000_000, 20)
    n = 1 << scale
    threads = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(31)
    s, d = oracle.rmat_edges(n, aspen(100) & 0xFFFFFFFF, 0.5, 0.1, 0.1, m0)
    g = oracle.CSR.from_edges(n, s, d)
    del s, d
    n_ins_pairs, n_del_pairs = max(1, int(batch * 0.9) // 2), max(1, int(batch * 0.1) // 2)
    ins = [oracle.build_keys(*oracle.rmat_edges(n, aspen(1000 + t) & 0xFFFFFFFF, 0.5, 0.1, 0.1, n_ins_pairs))
           for t in range(nb)]
    # delete pairs (u, v), u < v, existing, every endpoint used once
    need = nb * n_del_pairs
    e = rng.integers(0, g.m, 4 * need + 1000).astype(np.uint64)
    u = (np.searchsorted(g.off, e, side="right") - 1).astype(np.int64)
    v = g.nbr[e].astype(np.int64)
    seen, pairs = set(), []
    for a, b in zip(u.tolist(), v.tolist()):
        if a < b and a not in seen and b not in seen:
            seen.update((a, b))
            pairs.append((a, b))
            if len(pairs) == need:
                break
    pairs = np.array(pairs, np.uint64)
    dels = [pairs[t * n_del_pairs:(t + 1) * n_del_pairs] for t in range(nb)]
    dels = [oracle.sort_unique(np.concatenate([(p[:, 0] << np.uint64(32)) | p[:, 1],
                                              (p[:, 1] << np.uint64(32)) | p[:, 0]])) for p in dels]
    sample = (f"rMAT scale {scale}, {m0} generated edges ({g.m} directed); {nb} mixed batches of {batch} "
              f"(90% inserts / 10% deletes of existing edges); add_edge_batch + parallel remove_edge, "
              f"sort excluded as in graph_bm.cc")
    if os.path.exists(oracle.REF_HARNESS):
        with tempfile.TemporaryDirectory() as td:
            def write(path, keys):
                keys = np.asarray(keys, np.uint64)
                pairs_ = np.empty((len(keys), 2), np.uint32)
                pairs_[:, 0] = (keys >> np.uint64(32)).astype(np.uint32)
                pairs_[:, 1] = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
                with open(path, "wb") as f:
                    np.array([n, len(keys)], np.uint64).tofile(f)
                    pairs_.tofile(f)

            gpath = os.path.join(td, "g.bin")
            with open(gpath, "wb") as f:
                np.array([n, g.m], np.uint64).tofile(f)
                pairs_ = np.empty((g.m, 2), np.uint32)
                pairs_[:, 0] = np.repeat(np.arange(n, dtype=np.uint32), g.degrees())
                pairs_[:, 1] = g.nbr[:g.m]
                pairs_.tofile(f)
                del pairs_
            files = []
            for t in range(nb):
                files += [os.path.join(td, f"i{t}.bin"), os.path.join(td, f"d{t}.bin")]
                write(files[-2], ins[t])
                write(files[-1], dels[t])
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            log(f"cpu baseline: reference build, {threads} threads, {sample}")
            r = subprocess.run([oracle.REF_HARNESS, "mixed", gpath] + files, capture_output=True, text=True,
                               env=env, timeout=900)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                return {"value": res["edges_per_s"], "unit": "edges/s", "cores": res["threads"],
                        "kind": "reference", "sample": sample, "insert_s": res["insert_s"],
                        "delete_s": res["delete_s"], "static_build_s": res["static_s"]}
            log("reference harness failed:", r.stderr[-500:])
    # port: the oracle's set-model update (union, then difference), one thread
    keys = oracle.sort_unique((np.repeat(np.arange(n, dtype=np.uint64), g.degrees()) << np.uint64(32))
                              | g.nbr[:g.m].astype(np.uint64))
    t0 = time.perf_counter()
    total = 0
    for t in range(nb):
        keys = oracle.set_difference(oracle.set_union(keys, ins[t]), dels[t])
        total += len(ins[t]) + len(dels[t])
    dt = time.perf_counter() - t0
    return {"value": total / max(dt, 1e-9), "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "oracle set-model update, " + sample}


SERVE_PROBS = (0.57, 0.19, 0.19)
SERVE_SUMMARY_REQUESTS = 60  # per rate in the default run's serving summary
SUSTAINED_BATCHES = 60  # the default run's sustained update figure


def serve_requests(tg, n, G, R, rng):
    """R requests of the serving mix on graph G's current key set: 30%
    update batches (10..1e6 edges, log-uniform, 90% inserts of fresh rMAT
    edges / 10% deletes of existing edges, both directions), 70% queries
    (60% BFS, 20% CC, 10% PageRank 10 iterations FP64, 10% local triangle
    count).  Arrival times are set per stream (serve_stream)."""
    off, nbr = G.export_adjacency()
    m = len(nbr)
    deg = off[1:] - off[:-1]
    active = np.nonzero(deg)[0].astype(np.uint32)
    reqs = []
    t_gen = 0
    for _ in range(R):
        u = rng.random()
        if u < 0.3:
            size = int(round(10 ** rng.uniform(1, 6)))
            npairs = max(1, int(size * 0.9) // 2)
            s, d, w = tg.rmat_device(n, aspen(7000 + t_gen) & 0xFFFFFFFF, *SERVE_PROBS, npairs, nominal=True,
                                     symmetric=True, drop_loops=True)
            t_gen += 1
            ins = (s.download(np.uint32, w), d.download(np.uint32, w))
            s.close(), d.close()
            nd = max(1, int(size * 0.1) // 2)
            idx = rng.integers(0, m, nd)
            ds = (np.searchsorted(off, idx, side="right") - 1).astype(np.uint32)
            dd = nbr[idx]
            dels = (np.concatenate([ds, dd]), np.concatenate([dd, ds]))
            reqs.append(("U", ins, dels))
        else:
            q = rng.random()
            v = int(active[rng.integers(0, len(active))])
            if q < 0.6:
                reqs.append(("BFS", v))
            elif q < 0.8:
                reqs.append(("CC",))
            elif q < 0.9:
                reqs.append(("PR",))
            else:
                reqs.append(("LTC", v))
    return reqs


def serve_stream(tg, build, reqs, arrivals, budget, verify=True):
    """The stream (requests with their arrival times) through the plain
    server (one request at a time, in arrival order) and the batched one
    (continuous batching, snapshot at admission), each on a fresh copy of
    the graph; then, when verify is set, every batched result against the
    same request run alone on the version it reports (the plain server
    replaying the batched server's order on a third copy)."""
    R = len(reqs)

    def submit(S, r, arrival):
        if r[0] == "U":
            return S.submit(tg.REQ_UPDATE, arrival, ins=r[1], dels=r[2])
        if r[0] == "BFS":
            return S.submit(tg.REQ_BFS, arrival, src=r[1])
        if r[0] == "CC":
            return S.submit(tg.REQ_CC, arrival)
        if r[0] == "PR":
            return S.submit(tg.REQ_PR, arrival, iters=10)
        return S.submit(tg.REQ_LTC, arrival, src=r[1])

    def replay(mode, stream):
        G = build()
        t_c = time.perf_counter()
        S = tg.Server(G, mode=mode, mem_budget=budget << 30)  # reserves the lanes' working memory
        startup = time.perf_counter() - t_c
        ids = [submit(S, r, a) for r, a in stream]
        t0 = time.perf_counter()
        S.run()
        wall = time.perf_counter() - t0
        res = [S.result(i) for i in ids]
        summ = S.summary()
        summ["wall_s"] = wall
        summ["startup_s"] = startup
        S.close()
        G.close()
        return summ, res

    sp, rp = replay("plain", list(zip(reqs, arrivals)))
    sb, rb = replay("admission", list(zip(reqs, arrivals)))
    mism = verified = 0
    if verify:
        order = sorted(range(R), key=lambda i: (rb[i]["version"], 0 if reqs[i][0] == "U" else 1, i))
        _, rv = replay("plain", [(reqs[i], 0.0) for i in order])
        for pos, i in enumerate(order):
            r, a, b = reqs[i], rv[pos], rb[i]
            verified += 1
            if a["status"] != 0 or b["status"] != 0 or a["version"] != b["version"]:
                mism += 1
            elif r[0] in ("BFS", "CC") and a["digest"] != b["digest"]:
                mism += 1
            elif r[0] == "LTC" and a["value"] != b["value"]:
                mism += 1
            elif r[0] == "PR" and abs(a["fvalue"] - b["fvalue"]) > 1e-9 * max(abs(a["fvalue"]), 1e-300):
                mism += 1
    ahead = sum(1 for i in range(R) if reqs[i][0] != "U" and rb[i]["version"] != rp[i]["version"])
    keep = ("requests_per_s", "p50_s", "p99_s", "p50_query_s", "p99_query_s", "makespan_s", "groups", "rejected",
            "wall_s", "startup_s")
    return {"plain": {k: sp[k] for k in keep if k in sp}, "batched": {k: sb[k] for k in keep if k in sb},
            "throughput_x": sb["requests_per_s"] / max(sp["requests_per_s"], 1e-9),
            "p99_ratio": sb["p99_s"] / max(sp["p99_s"], 1e-12),
            "verified": verified, "result_mismatches": mism, "queries_on_later_version": ahead}


def serve_graph_builder(tg, scale, edges):
    n = 1 << scale
    seed = aspen(100) & 0xFFFFFFFF

    def build():
        G = tg.Graph(n)
        s, d, w = tg.rmat_device(n, seed, *SERVE_PROBS, edges, nominal=True, symmetric=True, drop_loops=True)
        G.add_edge_batch_dev(s, d, w)
        s.close(), d.close()
        tg.lib().tg_device_sync()
        return G
    return build


def serve_summary(tg, scale, edges, R, budget, rng):
    """Bounded serving comparison for the default run's extra: R requests
    of the serving mix at overload (200/s), then at 80% of the plain
    server's capacity measured in that run (its requests/s), each verified
    against the alone-on-version replay."""
    build = serve_graph_builder(tg, scale, edges)
    G = build()
    m_directed = G.get_num_edges()
    reqs = serve_requests(tg, 1 << scale, G, R, rng)
    G.close()
    gaps = rng.exponential(1.0, R)
    out = {"workload": f"rmat scale {scale} skewed (.57/.19/.19), {edges} generated edges "
                       f"({m_directed} directed); {R} requests: 30% updates of 10..1e6 edges, "
                       "70% queries (BFS/CC/PR/LTC)",
           "kinds": {k: sum(1 for r in reqs if r[0] == k) for k in ("U", "BFS", "CC", "PR", "LTC")}}
    log(f"serve summary: {R} requests at 200/s")
    over = serve_stream(tg, build, reqs, np.cumsum(gaps / 200.0), budget)
    out["overload_200_per_s"] = over
    rate = 0.8 * over["plain"]["requests_per_s"]
    log(f"serve summary: {R} requests at {rate:.1f}/s (80% of the plain capacity)")
    near = serve_stream(tg, build, reqs, np.cumsum(gaps / rate), budget)
    near["rate_per_s"] = rate
    out["near_capacity"] = near
    log(f"serve summary: overload x{over['throughput_x']:.2f} p99 ratio {over['p99_ratio']:.3f}; "
        f"near capacity x{near['throughput_x']:.2f} p99 ratio {near['p99_ratio']:.3f}; mismatches "
        f"{over['result_mismatches'] + near['result_mismatches']}")
    return out


def serve_bench(args):
    """Mixed request stream through the plain and the batched server.

    rMAT (skewed a=.57 b=.19 c=.19) graph built on the device; R requests of
    the serving mix (serve_requests) arrive as a Poisson process at the
    given rate.  The stream is replayed one request at a time (plain) and
    with continuous batching (snapshot at admission); every batched result
    is then checked against the query run alone on the version it reports
    (serve_stream)."""
    from terrace_amd import graph as tg

    os.environ.setdefault("TG_SERVE_PROGRESS", "1")
    n = 1 << args.scale
    rng = np.random.default_rng(2024)
    build = serve_graph_builder(tg, args.scale, args.edges)
    log(f"serve: building scale {args.scale} skewed graph, {args.edges} generated edges")
    t0 = time.perf_counter()
    G1 = build()
    m_directed = G1.get_num_edges()
    log(f"serve: graph built in {time.perf_counter() - t0:.2f} s, {m_directed} edges, "
        f"{G1.get_size() / 2**30:.1f} GiB, tiers {G1.tier_counts()}")
    R = args.requests
    reqs = serve_requests(tg, n, G1, R, rng)
    G1.close()
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, R))
    log(f"serve: {R} requests at {args.rate}/s")
    res = serve_stream(tg, build, reqs, arrivals, args.budget, verify=not args.no_verify)
    log(f"serve: plain {res['plain']}, batched {res['batched']}, {res['result_mismatches']} mismatches")
    kinds = {k: sum(1 for r in reqs if r[0] == k) for k in ("U", "BFS", "CC", "PR", "LTC")}
    out = {
        "metric": "mixed-stream requests/s (batched server)",
        "value": res["batched"]["requests_per_s"],
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": R,
        "warmup": 0,
        "ms_per_step": res["batched"]["makespan_s"] / R * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/f64",
        "data": "synthetic rMAT generated on the device (seeded); Poisson arrivals",
        "config": {"workload": f"serving: rmat scale {args.scale} skewed (.57/.19/.19), {args.edges} generated "
                               f"edges; {R} requests at {args.rate}/s", "vertices": n,
                   "directed_edges": m_directed, "requests": kinds, "parallelism": "single GPU"},
        "serve": dict(res, mode="snapshot at admission"),
    }
    print(json.dumps(out))


def batching_bench(args):
    """Batching on one resident graph (rMAT scale 24, a=.5 b=.1 c=.1):
    1024 BFS and 256 BC queries from seeded random sources, one traversal per
    query versus 64 sources per lane mask; 4096 small update batches (1..1000
    edges, log-uniform, 90% inserts / 10% deletes of existing edges) applied
    one batch at a time versus coalesced per launch.  The autotuner runs
    first.  Multi-source levels are compared with the per-source levels for
    a sample of lanes, multi-source BC with the per-source BC within 1e-9,
    and the two graphs' adjacency after the updates must be identical."""
    from terrace_amd import graph as tg
    from terrace_amd._lib import DeviceBuffer, lib, check, ptr

    n = 1 << args.scale
    probs = (0.5, 0.1, 0.1)
    seed = aspen(100) & 0xFFFFFFFF
    rng = np.random.default_rng(4242)

    def build():
        G = tg.Graph(n)
        s, d, w = tg.rmat_device(n, seed, *probs, args.edges, nominal=True, symmetric=True, drop_loops=True)
        G.add_edge_batch_dev(s, d, w)
        del s, d
        tg.lib().tg_device_sync()
        return G

    log(f"batching: building scale {args.scale}, {args.edges} generated edges")
    G = build()
    m_directed = G.get_num_edges()
    log(f"batching: {m_directed} directed edges, tiers {G.tier_counts()}")
    t0 = time.perf_counter()
    tune = G.autotune()
    tune["autotune_s"] = time.perf_counter() - t0
    tune["chosen"] = G.get_tuning()
    log(f"batching: autotune {tune}")
    deg = G.degrees()
    active = np.nonzero(deg)[0].astype(np.uint32)
    bfs_src = active[rng.integers(0, len(active), 1024)].astype(np.uint32)
    bc_src = active[rng.integers(0, len(active), 256)].astype(np.uint32)
    res = {"autotune": tune}

    # ---- BFS: per source vs 64 per lane mask (device outputs)
    par, lev = DeviceBuffer(4 * n), DeviceBuffer(4 * n)
    depth = ctypes.c_uint32()
    t0 = time.perf_counter()
    for v in bfs_src:
        check(lib().tg_bfs_dev(G.handle, int(v), par.ptr, lev.ptr, ctypes.byref(depth)))
    t_single = time.perf_counter() - t0
    mlev = DeviceBuffer(64 * 4 * n)
    t0 = time.perf_counter()
    for c in range(0, len(bfs_src), 64):
        chunk = np.ascontiguousarray(bfs_src[c:c + 64])
        check(lib().tg_bfs_multi_dev(G.handle, ptr(chunk), len(chunk), mlev.ptr))
    t_multi = time.perf_counter() - t0
    # the last chunk's lanes against per-source levels (sampled lanes)
    last = np.ascontiguousarray(bfs_src[-64:])
    bfs_bad = 0
    for j in range(0, 64, 8):
        check(lib().tg_bfs_dev(G.handle, int(last[j]), par.ptr, lev.ptr, ctypes.byref(depth)))
        a = lev.download(np.uint32, n)
        b = np.empty(n, np.uint32)
        check(lib().tg_memcpy_d2h(ptr(b), mlev.offset(4 * n * j), 4 * n))
        bfs_bad += int(not np.array_equal(a, b))
    res["bfs"] = {"queries": len(bfs_src), "single_s": t_single, "multi_s": t_multi,
                  "single_per_s": len(bfs_src) / t_single, "multi_per_s": len(bfs_src) / t_multi,
                  "speedup": t_single / t_multi, "lanes_checked": 8, "lanes_mismatched": bfs_bad}
    log(f"batching: BFS {res['bfs']}")
    mlev.close(), par.close(), lev.close()

    # ---- BC: per source vs 64 per lane mask
    bco = DeviceBuffer(8 * n)
    t0 = time.perf_counter()
    for v in bc_src:
        check(lib().tg_bc_dev(G.handle, int(v), bco.ptr))
    t_single = time.perf_counter() - t0
    mbc = DeviceBuffer(64 * 8 * n)
    t0 = time.perf_counter()
    for c in range(0, len(bc_src), 64):
        chunk = np.ascontiguousarray(bc_src[c:c + 64])
        check(lib().tg_bc_multi_dev(G.handle, ptr(chunk), len(chunk), mbc.ptr))
    t_multi = time.perf_counter() - t0
    last = np.ascontiguousarray(bc_src[-64:])
    worst = 0.0
    bc_bad = 0
    for j in range(0, 64, 16):
        check(lib().tg_bc_dev(G.handle, int(last[j]), bco.ptr))
        a = bco.download(np.float64, n)
        b = np.empty(n, np.float64)
        check(lib().tg_memcpy_d2h(ptr(b), mbc.offset(8 * n * j), 8 * n))
        fa, fb = np.isfinite(a), np.isfinite(b)
        if not np.array_equal(fa, fb):
            bc_bad += 1
            continue
        err = float(np.abs(a[fa] - b[fa]).sum() / max(np.abs(a[fa]).sum(), 1e-300))
        worst = max(worst, err)
        bc_bad += int(err > 1e-9)
    res["bc"] = {"queries": len(bc_src), "single_s": t_single, "multi_s": t_multi,
                 "single_per_s": len(bc_src) / t_single, "multi_per_s": len(bc_src) / t_multi,
                 "speedup": t_single / t_multi, "lanes_checked": 4, "lanes_out_of_tolerance": bc_bad,
                 "worst_rel_l1": worst}
    log(f"batching: BC {res['bc']}")
    mbc.close(), bco.close()

    # ---- 4096 small update batches: one at a time vs coalesced
    off, nbr = G.export_adjacency()
    sizes = np.round(10 ** rng.uniform(0, 3, 4096)).astype(np.int64)
    n_ins = np.maximum(1, (sizes * 9) // 10)
    n_del = np.maximum(0, sizes - n_ins)
    tot_ins = int(n_ins.sum())
    s_d, d_d, w = tg.rmat_device(n, aspen(9000) & 0xFFFFFFFF, *probs, tot_ins, nominal=True, symmetric=False,
                                 drop_loops=True)
    ins_s, ins_d = s_d.download(np.uint32, w), d_d.download(np.uint32, w)
    del s_d, d_d
    idx = rng.integers(0, len(nbr), int(n_del.sum()))
    del_s = (np.searchsorted(off, idx, side="right") - 1).astype(np.uint32)
    del_d = nbr[idx]
    del off, nbr
    batches, io, do = [], 0, 0
    for b in range(4096):
        ni = min(int(n_ins[b]), len(ins_s) - io)
        batches.append((ins_s[io:io + ni], ins_d[io:io + ni], del_s[do:do + n_del[b]], del_d[do:do + n_del[b]]))
        io += ni
        do += int(n_del[b])
    G2 = build()
    t0 = time.perf_counter()
    for bs, bd, xs, xd in batches:
        if len(bs):
            G2.add_edge_batch(bs, bd)
        if len(xs):
            G2.remove_edge_batch(xs, xd)
    t_single = time.perf_counter() - t0
    per_launch = 256
    chunks = []
    for c in range(0, 4096, per_launch):
        part = batches[c:c + per_launch]
        cs = np.concatenate([np.concatenate([p[0], p[2]]) for p in part])
        cd = np.concatenate([np.concatenate([p[1], p[3]]) for p in part])
        co = np.concatenate([np.concatenate([np.zeros(len(p[0]), np.uint32), np.ones(len(p[2]), np.uint32)])
                             for p in part])
        chunks.append((cs, cd, co))
    t0 = time.perf_counter()
    for cs, cd, co in chunks:
        G.apply_ops(cs, cd, co)
    t_coal = time.perf_counter() - t0
    o1, n1 = G.export_adjacency()
    o2, n2 = G2.export_adjacency()
    same = bool(np.array_equal(o1, o2) and np.array_equal(n1, n2))
    del o1, n1, o2, n2
    edges = int(sizes.sum())
    res["updates"] = {"batches": 4096, "edges": edges, "batches_per_launch": per_launch,
                      "single_s": t_single, "coalesced_s": t_coal, "speedup": t_single / t_coal,
                      "single_edges_per_s": edges / t_single, "coalesced_edges_per_s": edges / t_coal,
                      "adjacency_identical": same}
    log(f"batching: updates {res['updates']}")
    G.close(), G2.close()
    out = {
        "metric": "multi-source BFS queries/s (64 per lane mask)",
        "value": res["bfs"]["multi_per_s"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(bfs_src),
        "warmup": 0,
        "ms_per_step": res["bfs"]["multi_s"] / len(bfs_src) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/f64",
        "data": "synthetic rMAT generated on the device (seeded)",
        "config": {"workload": f"batching: rmat scale {args.scale} (.5/.1/.1), {args.edges} generated edges; "
                               "1024 BFS + 256 BC sources, 4096 update batches of 1..1000 edges",
                   "vertices": n, "directed_edges": m_directed, "parallelism": "single GPU"},
        "batching": res,
    }
    print(json.dumps(out))


def skew_bench(args):
    """Skew stress across all three tiers: rMAT scale 22 with a=.57 b=.19
    c=.19 (max degree above 1e5; inline, PMA and B-tree tiers populated).
    BFS, PageRank (10 iterations), CC, BC, SSSP_BF and TC from the
    highest-degree vertex (results left in HBM), then mixed update batches of 1e6 edges whose sources are the
    top-1% degree vertices (90% inserts to random active vertices, 10%
    deletes of existing edges of those sources; both directions).  The
    adjacency after the first batch is checked against a host set model."""
    from terrace_amd import graph as tg

    n = 1 << args.scale
    probs = (0.57, 0.19, 0.19)
    seed = aspen(100) & 0xFFFFFFFF
    rng = np.random.default_rng(99)
    log(f"skew: building scale {args.scale}, {args.edges} generated edges")
    G = tg.Graph(n)
    s, d, w = tg.rmat_device(n, seed, *probs, args.edges, nominal=True, symmetric=True, drop_loops=True)
    tg.lib().tg_device_sync()
    t0 = time.perf_counter()
    G.add_edge_batch_dev(s, d, w)
    tg.lib().tg_device_sync()
    t_build = time.perf_counter() - t0
    del s, d
    m = G.get_num_edges()
    deg = G.degrees()
    tiers = G.tier_counts()
    src = int(np.argmax(deg))
    log(f"skew: {m} directed edges in {t_build:.3f} s, max degree {int(deg.max())}, tiers {tiers}")
    res = {"static_build_s": t_build, "directed_edges": m, "max_degree": int(deg.max()),
           "active_vertices": int((deg > 0).sum()), "tiers": tiers, "source": src}
    from terrace_amd._lib import DeviceBuffer as _DB, check as _check
    par_d, lev_d, f64_d, i32_d = _DB(4 * n), _DB(4 * n), _DB(8 * n), _DB(4 * n)
    depth = ctypes.c_uint32()
    L = tg.lib()
    L.tg_device_sync()
    # the first traversal also allocates the frontier buffers and builds the
    # tier lists of this graph version; then a warm-up CC, untimed
    t0 = time.perf_counter()
    _check(L.tg_bfs_dev(G.handle, src, par_d.ptr, lev_d.ptr, ctypes.byref(depth)))
    res["bfs_first_call_ms"] = (time.perf_counter() - t0) * 1e3
    _check(L.tg_cc_dev(G.handle, i32_d.ptr, ctypes.byref(depth)))
    L.tg_device_sync()
    t0 = time.perf_counter()
    _check(L.tg_bfs_dev(G.handle, src, par_d.ptr, lev_d.ptr, ctypes.byref(depth)))
    t = time.perf_counter() - t0
    levels = lev_d.download(np.uint32, n)
    reached = levels != np.iinfo(np.uint32).max
    res["bfs_ms"], res["bfs_gteps"] = t * 1e3, float(deg[reached].astype(np.int64).sum()) / t / 1e9
    # PageRank: the first call builds the sliced adjacency copy, the second
    # reuses it; then the tiered pull on the same graph
    t0 = time.perf_counter()
    _check(L.tg_pagerank_dev(G.handle, 10, f64_d.ptr))
    t = time.perf_counter() - t0
    res["pr_first_call_ms"] = t * 1e3
    t0 = time.perf_counter()
    _check(L.tg_pagerank_dev(G.handle, 10, f64_d.ptr))
    t = time.perf_counter() - t0
    res["pr_ms"], res["pr_gteps"] = t * 1e3, 10 * m / t / 1e9
    G.set_tuning(pr_sliced=0)
    t0 = time.perf_counter()
    _check(L.tg_pagerank_dev(G.handle, 10, f64_d.ptr))
    res["pr_tiered_pull_ms"] = (time.perf_counter() - t0) * 1e3
    G.set_tuning(pr_sliced=1)
    rounds = ctypes.c_uint32()
    t0 = time.perf_counter()
    _check(L.tg_cc_dev(G.handle, i32_d.ptr, ctypes.byref(rounds)))
    res["cc_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    _check(L.tg_bc_dev(G.handle, src, f64_d.ptr))
    res["bc_ms"] = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    _check(L.tg_sssp_bf_dev(G.handle, src, i32_d.ptr, ctypes.byref(rounds)))
    res["sssp_bf_ms"] = (time.perf_counter() - t0) * 1e3
    for b_ in (par_d, lev_d, f64_d, i32_d):
        b_.close()
    log(f"skew: BFS {res['bfs_ms']:.1f} ms, PR {res['pr_ms']:.1f} ms, CC {res['cc_ms']:.1f} ms, "
        f"BC {res['bc_ms']:.1f} ms, SSSP_BF {res['sssp_bf_ms']:.1f} ms")
    t0 = time.perf_counter()
    res["tc"] = tg.TC(G)
    res["tc_ms"] = (time.perf_counter() - t0) * 1e3
    log(f"skew: TC {res['tc']} in {res['tc_ms']:.1f} ms")
    # hub-targeted 1e6 mixed batches
    off, nbr = G.export_adjacency()
    active = np.nonzero(deg)[0].astype(np.uint32)
    top = np.argsort(deg)[-max(1, int((deg > 0).sum()) // 100):].astype(np.uint32)
    batches = []
    for b in range(args.skew_batches + 1):
        pairs = 450_000
        hs = top[rng.integers(0, len(top), pairs)]
        hd = active[rng.integers(0, len(active), pairs)]
        ins_s, ins_d = np.concatenate([hs, hd]), np.concatenate([hd, hs])
        vs = top[rng.integers(0, len(top), 50_000)]
        lo, hi = off[vs].astype(np.int64), off[vs + 1].astype(np.int64)
        pos = lo + (rng.random(len(vs)) * (hi - lo)).astype(np.int64)
        dd = nbr[pos]
        del_s, del_d = np.concatenate([vs, dd]), np.concatenate([dd, vs])
        dev = [tg.DeviceBuffer(4 * len(x)) for x in (ins_s, ins_d, del_s, del_d)]
        for buf, x in zip(dev, (ins_s, ins_d, del_s, del_d)):
            buf.upload(np.ascontiguousarray(x, np.uint32))
        batches.append((dev, len(ins_s), len(del_s), (ins_s, ins_d, del_s, del_d)))
    # batch 0 (warm-up) is checked against the host set model
    dev, ni, nd, host = batches[0]
    G.update_batch_dev(dev[0], dev[1], ni, dev[2], dev[3], nd)
    keys = (np.repeat(np.arange(n, dtype=np.uint64), np.diff(off).astype(np.int64)) << np.uint64(32)) | nbr
    del nbr
    ins = np.unique((host[0].astype(np.uint64) << np.uint64(32)) | host[1])
    dl = np.unique((host[2].astype(np.uint64) << np.uint64(32)) | host[3])
    p = np.searchsorted(keys, ins).clip(max=len(keys) - 1)
    want = np.concatenate([keys, ins[keys[p] != ins]])
    del keys
    want.sort()
    p = np.searchsorted(want, dl).clip(max=len(want) - 1)
    keep = np.ones(len(want), bool)
    keep[p[want[p] == dl]] = False
    want = want[keep]
    o2, n2 = G.export_adjacency()
    got = (np.repeat(np.arange(n, dtype=np.uint64), np.diff(o2).astype(np.int64)) << np.uint64(32)) | n2
    res["update_check_equal"] = bool(np.array_equal(got, want))
    del got, want, o2, n2
    log(f"skew: batch 0 adjacency equal to the set model: {res['update_check_equal']}")
    tg.lib().tg_device_sync()
    t0 = time.perf_counter()
    tot = 0
    for dev, ni, nd, _ in batches[1:]:
        G.update_batch_dev(dev[0], dev[1], ni, dev[2], dev[3], nd)
        tot += ni + nd
    tg.lib().tg_device_sync()
    t = time.perf_counter() - t0
    res["update_batches"] = args.skew_batches
    res["update_edges_per_s"] = tot / t
    res["ms_per_batch"] = t / args.skew_batches * 1e3
    res["tiers_after"] = G.tier_counts()
    log(f"skew: {args.skew_batches} hub-targeted batches of {tot // args.skew_batches} edges: "
        f"{res['update_edges_per_s']:.3e} edges/s")
    G.close()
    out = {
        "metric": "updated edges/s (hub-targeted 1e6 mixed batches, skewed graph)",
        "value": res["update_edges_per_s"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.skew_batches,
        "warmup": 1,
        "ms_per_step": res["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic rMAT generated on the device (seeded)",
        "config": {"workload": f"skew stress: rmat scale {args.scale} (.57/.19/.19), {args.edges} generated edges; "
                               "1e6-edge mixed batches from the top-1% degree vertices",
                   "vertices": n, "directed_edges": m, "parallelism": "single GPU"},
        "skew": res,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--scale", type=int, default=None, help="default 22 (26 with --serve)")
    ap.add_argument("--edges", type=int, default=None, help="generated edges, default 69M (775M with --serve)")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-extra-batches", action="store_true", help="skip the 1e3 / 1e7 batch sizes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller graph for smoke runs")
    ap.add_argument("--copy-probe", action="store_true",
                    help="finish with three 1 GiB device copies (HBM counter calibration)")
    ap.add_argument("--serve", action="store_true", help="mixed-stream serving comparison instead")
    ap.add_argument("--skew", action="store_true", help="skew stress across all three tiers instead")
    ap.add_argument("--skew-batches", type=int, default=5)
    ap.add_argument("--batching", action="store_true",
                    help="multi-source BFS/BC and coalesced small updates instead")
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=200.0, help="request arrivals per second")
    ap.add_argument("--budget", type=int, default=160, help="server memory budget, GiB")
    ap.add_argument("--no-verify", action="store_true", help="serve: skip the verification replay")
    ap.add_argument("--no-serve", action="store_true",
                    help="default run: skip the bounded serving summary (scale 26) in extra")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.scale is None:
        args.scale = 26 if args.serve else 24 if args.batching else 22
    if args.edges is None:
        args.edges = (775_000_000 if args.serve else 136_000_000 if args.batching
                      else 125_000_000 if args.skew else 69_000_000)
    if args.skew:
        skew_bench(args)
        return
    if args.serve:
        serve_bench(args)
        return
    if args.batching:
        batching_bench(args)
        return

    from terrace_amd import graph as tg

    scale = args.scale
    m0 = args.edges
    if args.quick:
        scale, m0 = 18, 4_000_000
    n = 1 << scale
    default_workload = (scale, m0) == (22, 69_000_000)
    probs = (0.5, 0.1, 0.1)
    seed = aspen(100) & 0xFFFFFFFF
    rng = np.random.default_rng(7)

    log(f"static build: scale {scale}, {m0} generated edges")
    G = tg.Graph(n)
    s, d, w = tg.rmat_device(n, seed, *probs, m0, nominal=True, symmetric=True, drop_loops=True)
    tg.lib().tg_device_sync()
    t0 = time.perf_counter()
    G.add_edge_batch_dev(s, d, w)
    tg.lib().tg_device_sync()
    t_static = time.perf_counter() - t0
    del s, d
    m_static = G.get_num_edges()
    log(f"static: {m_static} directed edges in {t_static:.3f} s; stats {G.stats()}")
    off, nbr = G.export_adjacency()
    if not args.no_extra:
        # untimed: first use of the traversal's frontier and scratch buffers
        # (allocation), so the timed queries below measure the steady state
        # of "update batch, then query" (per-version tier lists rebuilt)
        tg.BFS_with_edge_map(G, 9, with_levels=True)
        tg.CC(G)

    batches = make_batches(tg, G, n, probs, args.batch, args.warmup + args.steps, 1000, rng, off, nbr)
    for i in range(args.warmup):
        apply_batch(G, batches[i])
    tg.lib().tg_device_sync()
    # the timed steps also carry the roofline's HIP events: one before the
    # first step on the graph's stream and one after the last (nothing is
    # recorded between steps)
    st0 = G.stats()
    G.profile(True)
    G.profile_reset()
    t0 = time.perf_counter()
    for i in range(args.steps):
        apply_batch(G, batches[args.warmup + i])
    tg.lib().tg_device_sync()
    t_steps = time.perf_counter() - t0
    step_ms, step_n = G.profile_get("update_step")
    G.profile(False)
    st1 = G.stats()
    upd = sum(b[2] + b[5] for b in batches[args.warmup:])
    ins = sum(b[2] for b in batches[args.warmup:])
    value = upd / t_steps
    log(f"updates: {upd} in {t_steps:.4f} s -> {value:.3e} edges/s")

    extra = {"static_build_edges_per_s": m_static / t_static, "static_edges": m_static,
             "insert_share_edges_per_s": ins / t_steps, "delete_share_edges_per_s": (upd - ins) / t_steps}
    # Roofline of the headline work: the mixed update step over the timed
    # steps, its time from the HIP events above, its algorithmic bytes from
    # the sources the device counted for the steps.
    step_avg_ms = step_ms / max(step_n, 1)
    alg_bytes = step_bytes(st1, st0, ins, upd - ins) / max(step_n, 1)
    achieved = alg_bytes / (step_avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": UPDATE_STEP_TRAFFIC_DEFAULT if default_workload else None,
                "kernel": "mixed update step (tg_update_batch_dev: sort, existence, tier merges, PMA and "
                          "B-tree updates; every kernel of the step)",
                "avg_ms": step_avg_ms, "launches": step_n, "algorithmic_bytes": alg_bytes,
                "bytes_formula": "12 b_ins + 16 u_ins + 8 b_del + 16 u_del"}
    log(f"update step: {step_avg_ms:.4f} ms (events), {roofline['algorithmic_bytes']:.3e} B, frac {roofline['frac']:.2e}")
    if not args.no_extra:
        src = 9
        deg = G.degrees()
        # BFS / PR / CC through the device-output entry points: results stay in
        # HBM (as the reference's stay in host memory); the copy back is not timed
        from terrace_amd._lib import DeviceBuffer as _DB, check as _check
        par_d, lev_d, pr_d, cc_d = _DB(4 * n), _DB(4 * n), _DB(8 * n), _DB(4 * n)
        depth = ctypes.c_uint32()
        tg.lib().tg_device_sync()
        t0 = time.perf_counter()
        _check(tg.lib().tg_bfs_dev(G.handle, src, par_d.ptr, lev_d.ptr, ctypes.byref(depth)))
        t_bfs = time.perf_counter() - t0
        levels = lev_d.download(np.uint32, n)
        reached = levels != np.iinfo(np.uint32).max
        bfs_edges = int(deg[reached].astype(np.int64).sum())
        extra["bfs_ms"] = t_bfs * 1e3
        extra["bfs_gteps"] = bfs_edges / t_bfs / 1e9
        extra["bfs_levels"] = int(depth.value)
        # PR, profiled (per-kernel HIP events): the first call also allocates
        # and builds the sliced adjacency copy (cold)
        G.profile(True)
        G.profile_reset()
        t0 = time.perf_counter()
        _check(tg.lib().tg_pagerank_dev(G.handle, 10, pr_d.ptr))
        extra["pr_first_call_ms"] = (time.perf_counter() - t0) * 1e3
        m_now = G.get_num_edges()
        # Roofline of the dominant kernel work, one PageRank iteration, timed
        # with HIP events on the traversal's stream.  Graphs up to 6M vertices
        # run the XCD-sliced pull (k_prs_pull + k_prs_fix + k_prs_reduce per
        # iteration):
        # algorithmic bytes = the PageRank per-iteration figure 4m + 24n
        # (4 B per edge id, 8 B metadata and 16 B p read / p_next write per
        # vertex).  Larger graphs run the tiered pull, whose dominant kernel is
        # k_pull_pma<PullSum>: 4 B per PMA edge beyond the inline slots +
        # 24 B per PMA-tier vertex.
        it_ms, it_n = G.profile_get("pr_iteration")
        if it_n:
            it_avg = it_ms / it_n
            it_bytes = 4 * m_now + 24 * n
            achieved = it_bytes / (it_avg * 1e-3) / 1e9
            pull_ms_, pull_n_ = G.profile_get("prs_pull")
            red_ms_, red_n_ = G.profile_get("prs_reduce")
            fix_ms_, fix_n_ = G.profile_get("prs_fix")
            extra["pr_roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": PR_ITER_TRAFFIC_DEFAULT if default_workload else None,
                        "kernel": "PageRank iteration: k_prs_pull + k_prs_fix + k_prs_reduce (XCD-sliced pull)",
                        "avg_ms": it_avg, "launches": it_n, "algorithmic_bytes": it_bytes,
                        "k_prs_pull_avg_ms": pull_ms_ / max(pull_n_, 1),
                        "k_prs_fix_avg_ms": fix_ms_ / max(fix_n_, 1),
                        "k_prs_reduce_avg_ms": red_ms_ / max(red_n_, 1)}
        else:
            tc_ = G.tier_counts()
            pma_ms, pma_n = G.profile_get("pull_pma_sum")
            pma_bytes = 4 * tc_["pma_edges"] + 24 * tc_["pma_vertices"]
            pma_avg = pma_ms / max(pma_n, 1)
            achieved = pma_bytes / (pma_avg * 1e-3) / 1e9
            extra["pr_roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                        "kernel": "k_pull_pma<PullSum> (PageRank dense pull over PMA-tier vertices)",
                        "avg_ms": pma_avg, "launches": pma_n, "algorithmic_bytes": pma_bytes, "units": tc_}
        # the tiered pull on the same graph, for comparison
        G.set_tuning(pr_sliced=0)
        G.profile_reset()
        t0 = time.perf_counter()
        _check(tg.lib().tg_pagerank_dev(G.handle, 10, pr_d.ptr))
        extra["pr_tiered_pull_ms"] = (time.perf_counter() - t0) * 1e3
        for kn in ("pull_inline_sum", "pull_pma_sum", "pull_tree_sum"):
            ms_, n_ = G.profile_get(kn)
            extra[kn + "_avg_ms"] = ms_ / max(n_, 1)
        G.set_tuning(pr_sliced=1)
        G.profile(False)
        # PR right after an update (one edge removed and put back: same graph,
        # new version, so the slices are rebuilt with warm buffers), as when
        # every batch is followed by a PageRank
        one_s, one_d = np.array([src], np.uint32), np.array([int(G.neighbors(src)[0])], np.uint32)
        G.remove_edge_batch(one_s, one_d)
        G.add_edge_batch(one_s, one_d)
        tg.lib().tg_device_sync()
        t0 = time.perf_counter()
        _check(tg.lib().tg_pagerank_dev(G.handle, 10, pr_d.ptr))
        t_pr = time.perf_counter() - t0
        extra["pr_ms"] = t_pr * 1e3
        extra["pr_gteps"] = 10 * m_now / t_pr / 1e9
        # CC
        rounds = ctypes.c_uint32()
        t0 = time.perf_counter()
        _check(tg.lib().tg_cc_dev(G.handle, cc_d.ptr, ctypes.byref(rounds)))
        extra["cc_ms"] = (time.perf_counter() - t0) * 1e3
        # PageRank with its iterations replayed from a captured HIP graph
        # (the profiled run above launches them one by one; the graph is
        # captured on the second call on this graph version, replayed after)
        for _ in range(2):
            t0 = time.perf_counter()
            _check(tg.lib().tg_pagerank_dev(G.handle, 10, pr_d.ptr))
            extra["pr_graph_replay_ms"] = (time.perf_counter() - t0) * 1e3
        # BFS and CC again on the same version: levels replayed from their
        # captured two-level graphs (the first calls above launch them one
        # by one); best of three
        for name, fn in (("bfs_replay_ms", lambda: tg.lib().tg_bfs_dev(G.handle, src, par_d.ptr, lev_d.ptr,
                                                                        ctypes.byref(depth))),
                         ("cc_replay_ms", lambda: tg.lib().tg_cc_dev(G.handle, cc_d.ptr, ctypes.byref(rounds)))):
            ts = []
            for _ in range(4):
                tg.lib().tg_device_sync()
                t0 = time.perf_counter()
                _check(fn())
                ts.append((time.perf_counter() - t0) * 1e3)
            extra[name] = min(ts[1:])
        extra["bfs_replay_gteps"] = bfs_edges / (extra["bfs_replay_ms"] * 1e-3) / 1e9
        for b_ in (par_d, lev_d, pr_d, cc_d):
            b_.close()
        for bsz in (() if args.no_extra_batches else (1000, 10_000_000)):
            bt = make_batches(tg, G, n, probs, bsz, 2, 5000 + bsz % 97, rng, off, nbr)
            apply_batch(G, bt[0])
            tg.lib().tg_device_sync()
            t0 = time.perf_counter()
            apply_batch(G, bt[1])
            tg.lib().tg_device_sync()
            dt = time.perf_counter() - t0
            extra[f"update_edges_per_s_b{bsz}"] = (bt[1][2] + bt[1][5]) / dt
        # insert-only and delete-only batches, timed apart as the reference's
        # graph_bm does (src/graph_bm.cc:325-388): b requested edges (rMAT,
        # symmetrized) inserted, then the same b edges deleted; median of
        # up to 5 batches per size
        ins_rate, del_rate = {}, {}
        for bsz in (() if args.no_extra_batches else (1000, 100_000, 10_000_000)):
            reps = 5 if bsz <= 100_000 else 1
            ti, td, cnt_ = [], [], 0
            for r_ in range(reps + 1):
                s_, d_, w_ = tg.rmat_device(n, aspen(20000 + 7 * r_ + bsz % 89) & 0xFFFFFFFF, *probs, bsz // 2,
                                            nominal=True, symmetric=True, drop_loops=True)
                tg.lib().tg_device_sync()
                t0 = time.perf_counter()
                G.add_edge_batch_dev(s_, d_, w_)
                tg.lib().tg_device_sync()
                t1 = time.perf_counter()
                G.remove_edge_batch_dev(s_, d_, w_)
                tg.lib().tg_device_sync()
                t2 = time.perf_counter()
                if r_:  # the first of each size warms its buffers
                    ti.append(t1 - t0)
                    td.append(t2 - t1)
                    cnt_ = w_
            ins_rate[bsz] = cnt_ / float(np.median(ti))
            del_rate[bsz] = cnt_ / float(np.median(td))
        extra["insert_only_edges_per_s"] = ins_rate
        extra["delete_only_edges_per_s"] = del_rate
        # the same 1e5 mixed batches handed over as host arrays (PCIe-inclusive)
        bt = make_batches(tg, G, n, probs, args.batch, 3, 8000, rng, off, nbr)
        hb = [(s_.download(np.uint32, w_), d_.download(np.uint32, w_), bs_.download(np.uint32, nd_),
               bd_.download(np.uint32, nd_)) for s_, d_, w_, bs_, bd_, nd_ in bt]
        tg.lib().tg_device_sync()
        t0 = time.perf_counter()
        for hs, hd, xs, xd in hb:
            G.update_batch(hs, hd, xs, xd)
        tg.lib().tg_device_sync()
        dt = time.perf_counter() - t0
        extra["update_edges_per_s_host_arrays_b1e5"] = sum(len(h[0]) + len(h[2]) for h in hb) / dt
        # sustained: 60 more mixed batches of the headline size as one timed
        # region, including the batches that promote a vertex to the B-tree
        # tier (about one in twelve here), which take the general passes
        bt = make_batches(tg, G, n, probs, args.batch, SUSTAINED_BATCHES, 9000, rng, off, nbr)
        sa = G.stats()
        tg.lib().tg_device_sync()
        t0 = time.perf_counter()
        for b_ in bt:
            apply_batch(G, b_)
        tg.lib().tg_device_sync()
        dt = time.perf_counter() - t0
        sb = G.stats()
        extra["update_sustained"] = {
            "batches": SUSTAINED_BATCHES, "edges_per_s": sum(b_[2] + b_[5] for b_ in bt) / dt,
            "ms_per_batch": dt / SUSTAINED_BATCHES * 1e3, "promotions": sb["promotions"] - sa["promotions"]}
        log(f"sustained: {SUSTAINED_BATCHES} batches, {extra['update_sustained']['ms_per_batch']:.4f} ms per batch, "
            f"{extra['update_sustained']['promotions']} promotions")
        extra["memory_bytes"] = G.get_size()
        extra["stats"] = G.stats()

    if not args.no_extra and not args.no_serve and default_workload:
        # the serving workload (scale 26 skewed) at a bounded size, beside
        # the headline: plain versus continuously batched, near capacity and
        # at overload, every result verified
        G.close()
        extra["serving"] = serve_summary(tg, 26, 775_000_000, SERVE_SUMMARY_REQUESTS, args.budget,
                                         np.random.default_rng(2025))
    if args.copy_probe:
        copy_probe()
    cpu = None if args.no_cpu else cpu_baseline(args.quick, args.batch)
    out = {
        "metric": f"updated edges/s (mixed batches of {args.batch}: 90% inserts, 10% deletes)",
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_steps / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic rMAT generated on the device (seeded, closed form)",
        "config": {"workload": f"rmat scale {scale}, {m0} generated edges (a=.5 b=.1 c=.1), "
                               f"symmetrized; mixed update batches of {args.batch} "
                               f"(90% insert / 10% delete)",
                   "vertices": n, "directed_edges": m_static, "batch": args.batch,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
