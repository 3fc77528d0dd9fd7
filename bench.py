"""Benchmark: batched edge updates on the scale-22 rMAT graph, one MI355X.

Workload: rMAT scale 22 (4,194,304 vertex ids), 69M generated edges with
a=0.5 b=0.1 c=0.1 (divisor 2^32), self-loops dropped, symmetrized and
deduplicated (135.9M directed edges), built on the device.  A step is one
mixed update batch of --batch directed edge updates: 90% inserts (fresh
rMAT edges, both directions) and 10% deletes of edges of the graph (both
directions), applied as one insert call then one delete call.  All batch
inputs are generated into HBM before the timed region.

After the timed steps the run also times BFS, PageRank (10 iterations,
FP64), CC, insert batches of 10^3 and 10^7 edges, and reports them under
"extra".  Prints one JSON line.
"""
from __future__ import annotations

import argparse
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
    s, d, w, bs, bd, nd = bt
    G.add_edge_batch_dev(s, d, w)
    G.remove_edge_batch_dev(bs, bd, nd)


def cpu_baseline(quick: bool):
    """The reference's own add_edge_batch (oracle/_ref, built from the
    reference sources) timed on this host on a bounded sample: scale-20
    rMAT (8M generated edges), 5 insert batches of 10^5 edges.  Falls back
    to the oracle's set-union restatement when the reference build is
    absent."""
    import oracle

    n = 1 << 20
    m0 = 2_000_000 if quick else 8_000_000
    keys = oracle.build_keys(*oracle.rmat_edges(n, aspen(100) & 0xFFFFFFFF, 0.5, 0.1, 0.1, m0))
    threads = min(16, os.cpu_count() or 1)
    if os.path.exists(oracle.REF_HARNESS):
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "g.bin")
            pairs = np.empty((len(keys), 2), np.uint32)
            pairs[:, 0] = (keys >> np.uint64(32)).astype(np.uint32)
            pairs[:, 1] = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
            with open(path, "wb") as f:
                np.array([n, len(keys)], np.uint64).tofile(f)
                pairs.tofile(f)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            r = subprocess.run([oracle.REF_HARNESS, "bench", path, "100000", "5", "1000"],
                               capture_output=True, text=True, env=env, timeout=600)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                return {"value": res["edges_per_s"], "unit": "edges/s", "cores": res["threads"],
                        "kind": "reference",
                        "sample": f"reference add_edge_batch (OpenMP build), rMAT scale 20 with {m0} "
                                  f"generated edges ({res['static_edges']} directed), 5 insert batches "
                                  f"of 1e5 directed edges; sort excluded as in graph_bm.cc"}
            log("reference harness failed:", r.stderr[-500:])
    # port: set-union insert of the oracle
    t0 = time.perf_counter()
    total = 0
    for t in range(5):
        bk = oracle.build_keys(*oracle.rmat_edges(n, aspen(1000 + t) & 0xFFFFFFFF, 0.5, 0.1, 0.1, 50000))
        t1 = time.perf_counter()
        keys = oracle.set_union(keys, bk)
        total += len(bk)
        t0 += 0
        elapsed = time.perf_counter() - t1
        t0 -= elapsed
    dt = time.perf_counter() - t0
    return {"value": total / max(dt, 1e-9), "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "oracle set-union insert, rMAT scale 20, 5 batches of 1e5"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edges", type=int, default=69_000_000)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller graph for smoke runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    from terrace_amd import graph as tg

    scale = args.scale
    m0 = args.edges
    if args.quick:
        scale, m0 = 18, 4_000_000
    n = 1 << scale
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

    batches = make_batches(tg, G, n, probs, args.batch, args.warmup + args.steps, 1000, rng, off, nbr)
    for i in range(args.warmup):
        apply_batch(G, batches[i])
    tg.lib().tg_device_sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        apply_batch(G, batches[args.warmup + i])
    tg.lib().tg_device_sync()
    t_steps = time.perf_counter() - t0
    upd = sum(b[2] + b[5] for b in batches[args.warmup:])
    ins = sum(b[2] for b in batches[args.warmup:])
    value = upd / t_steps
    log(f"updates: {upd} in {t_steps:.4f} s -> {value:.3e} edges/s")

    extra = {"static_build_edges_per_s": m_static / t_static, "static_edges": m_static,
             "insert_share_edges_per_s": ins / t_steps}
    roofline = None
    if not args.no_extra:
        src = 9
        deg = G.degrees()
        # BFS
        G.profile(True)
        G.profile_reset()
        t0 = time.perf_counter()
        parents = tg.BFS_with_edge_map(G, src)
        t_bfs = time.perf_counter() - t0
        reached = parents >= 0
        bfs_edges = int(deg[reached].astype(np.int64).sum())
        extra["bfs_ms"] = t_bfs * 1e3
        extra["bfs_gteps"] = bfs_edges / t_bfs / 1e9
        # PR
        G.profile_reset()
        t0 = time.perf_counter()
        tg.PR_S(G, 10)
        t_pr = time.perf_counter() - t0
        m_now = G.get_num_edges()
        extra["pr_ms"] = t_pr * 1e3
        extra["pr_gteps"] = 10 * m_now / t_pr / 1e9
        pull_ms, pull_n = G.profile_get("edge_map_pull")
        # PR pull per iteration: algorithmic bytes 4m + 8n (metadata) + 16n
        per_iter_bytes = 4 * m_now + 8 * n + 16 * n
        avg_ms = pull_ms / max(pull_n, 1)
        achieved = per_iter_bytes / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "kernel": "PageRank dense pull (k_pull_inline + k_pull_pma + k_pull_tree), one iteration",
                    "avg_ms": avg_ms}
        # CC
        t0 = time.perf_counter()
        tg.CC(G)
        extra["cc_ms"] = (time.perf_counter() - t0) * 1e3
        G.profile(False)
        for bsz in (1000, 10_000_000):
            bt = make_batches(tg, G, n, probs, bsz, 2, 5000 + bsz % 97, rng, off, nbr)
            apply_batch(G, bt[0])
            tg.lib().tg_device_sync()
            t0 = time.perf_counter()
            apply_batch(G, bt[1])
            tg.lib().tg_device_sync()
            dt = time.perf_counter() - t0
            extra[f"update_edges_per_s_b{bsz}"] = (bt[1][2] + bt[1][5]) / dt
        extra["memory_bytes"] = G.get_size()
        extra["stats"] = G.stats()

    cpu = None if args.no_cpu else cpu_baseline(args.quick)
    out = {
        "metric": "inserted edges/s",
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
