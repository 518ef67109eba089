#!/usr/bin/env python3
"""Benchmark: 4-hop walks from 4096 sources on an R-MAT scale-22 graph, one MI355X.

A step is one complete walk: X_0 (4096 one-hot rows) -> X_1 .. X_4, each hop a
symbolic pass and a numeric pass over the device-resident A^T bit tiles; every X_k stays
on the device. Graph generation and the tile build are setup (untimed).

Prints ONE JSON line (rank 0). Extra keys: tile_products_per_s, per_hop, roofline
(dominant kernel: the numeric walk kernel, HIP events on its launch stream over the
timed region), cpu_baseline (the same tile algorithm on the host cores, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "walk queries/s + tile-products/s, 4-hop R-MAT-22, learned+batched vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def two_hop(M, ctx, scale):
    """Two-hop pair enumeration (the CSR-tiled walk's result) on the device, and the oracle's
    bitset product A*A on the host cores for the same graph (checked equal in count)."""
    import oracle as O

    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 2)
    n, m = el.info()
    el.two_hop().close()
    ctx.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        p = el.two_hop()
        n_pairs = p.info()[1]
        p.close()
    gpu_s = (time.perf_counter() - t0) / reps
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, 2)
    rows = np.zeros((n, (n + 63) // 64), np.uint64)
    np.bitwise_or.at(rows, (s, d // 64), np.left_shift(np.uint64(1), (d % 64).astype(np.uint64)))
    t0 = time.perf_counter()
    c = O.bitset_matmul(rows, rows, n, n, n, threads=O.max_threads())
    cpu_s = time.perf_counter() - t0
    idx = np.arange(n)
    diag = (c[idx, idx // 64] >> (idx % 64).astype(np.uint64)) & np.uint64(1)
    cpu_pairs = int(np.unpackbits(c.view(np.uint8)).sum()) - int(diag.sum())
    print(json.dumps({"two_hop": {
        "graph": f"R-MAT scale {scale}, edge factor 16, seed 2, {m} edges", "pairs": n_pairs,
        "device_ms": round(gpu_s * 1e3, 3), "pairs_per_s": round(n_pairs / gpu_s, 1),
        "cpu_baseline": {"value": round(cpu_pairs / cpu_s, 1), "unit": "pairs/s", "cores": O.max_threads(),
                         "kind": "port", "ms": round(cpu_s * 1e3, 3),
                         "sample": "the whole graph: oracle bitset product A*A (OpenMP), diagonal excluded",
                         "pairs_equal": cpu_pairs == n_pairs}}}), flush=True)


def cpu_baseline(M, O, ctx, w, at, n_edges, hops, n_src, stride, threads, reps=1):
    """Time the oracle's tile algorithm (OpenMP) on every `stride`-th strip of each hop
    (stride 1: the whole walk), using the GPU's X_{k-1} as input, `reps` times, and scale each
    hop by total / sampled A^T bits. Also checks the sampled strips against the GPU frontier
    (bit-exact)."""
    ath = at.download()
    ns = ath.n_strips
    strips = np.arange(0, ns, stride)
    est_total = 0.0
    sample_s = 0.0
    for k in range(1, hops + 1):
        x_in = w.frontier(k - 1)
        x_ref = w.frontier(k)
        x_out = np.zeros_like(x_in)
        dt = 0.0
        for _ in range(reps):
            x_out[:] = 0
            t0 = time.perf_counter()
            _, visited = O.walk_tile_hop(ns, ath.offset_row, ath.col_idx, ath.data, x_in, 0, ns, threads,
                                         x_out=x_out, stride=stride)
            dt += time.perf_counter() - t0
        sample_s += dt
        dt /= reps
        rows = (strips[:, None] * 64 + np.arange(64)[None, :]).ravel()
        if not np.array_equal(x_out[rows], x_ref[rows]):
            raise SystemExit(f"cpu_baseline: CPU tile algorithm disagrees with the GPU at hop {k}")
        est_total += dt * (n_edges / max(visited, 1))
        log(f"cpu hop {k}: {dt:.2f}s on {len(strips)} strips, {visited} A^T bits -> est {dt * n_edges / max(visited, 1):.2f}s")
    return {
        "value": n_src / est_total,
        "unit": "source-walks/s",
        "cores": threads,
        "host_cpus": os.cpu_count(),
        "nproc": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": (f"oracle tile algorithm (OpenMP, {threads} threads) on every output strip of each of the {hops} "
                   f"hops (the whole walk), {reps} times, {sample_s:.1f}s measured; every output row checked bit-exact "
                   f"against the GPU") if stride == 1 else
                  (f"oracle tile algorithm (OpenMP, {threads} threads) on every {stride}-th output strip of each of "
                   f"the {hops} hops ({len(strips)} strips/hop, {reps} times, {sample_s:.1f}s measured), scaled per "
                   f"hop by total/sampled A^T bits; sampled strips checked bit-exact against the GPU"),
        "est_seconds_per_walk": est_total,
    }


def plain_walks(M, ctx, graphs, queries, reps):
    """Every query run alone (its own walk, words per vertex sized to the query): seconds per
    pass over all queries, and each query's X_hops digest."""
    walks = {}

    def walk_for(q):
        W = 1
        while W * 64 < len(q.sources):
            W *= 2
        key = (q.graph, W)
        if key not in walks:
            walks[key] = M.Walk(ctx, graphs[q.graph], 64 * W, 6)
        return walks[key]

    def run_all():
        for q in queries:
            walk_for(q).run(q.sources, q.hops)

    run_all()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run_all()
    ctx.synchronize()
    secs = (time.perf_counter() - t0) / reps
    digests = []
    for q in queries:  # a walk object is shared by queries: rerun each for its own digest
        w = walk_for(q)
        w.run(q.sources, q.hops)
        digests.append(w.frontier_digest(q.hops))
    for w in walks.values():
        w.close()
    return secs, digests


def batch_section(M, ctx, seed0: int, query_seed: int, n_queries: int, reps: int):
    """Uneven query batch: 512 queries over 8 graphs (R-MAT scales 16..20, three ER
    n=2^18 p=2^-13), sizes log-uniform in [1, 4096], hops uniform in [1, 6].
    Plain = every query run alone; batched = one captured HIP graph over panels, launch
    options from the autotuner. Per-query X_hops digests must agree between the two forms."""
    from matrixgraph_amd import autotune
    from matrixgraph_amd import workloads as WL

    specs = WL.batch_graph_specs(seed0)
    t0 = time.perf_counter()
    built = [WL.build_graph(ctx, s) for s in specs]
    graphs = [g for g, _ in built]
    queries = WL.batch_queries(specs, n_queries, query_seed)
    log(f"batch: {len(specs)} graphs built in {time.perf_counter() - t0:.1f}s, {len(queries)} queries")

    plain_s, plain_digests = plain_walks(M, ctx, graphs, queries, reps)
    log(f"batch plain: {plain_s * 1e3:.1f} ms per batch ({n_queries / plain_s:.1f} queries/s)")

    best, table = autotune.tune(ctx, graphs, queries, reps=2, log=log)
    b = M.WalkBatch(ctx, graphs, queries, packing=best["packing"], chunk_tiles=best["chunk_tiles"], use_graph=True)
    b.set_hop_kernel(best.get("segments", 0))
    b.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    ctx.synchronize()
    batched_s = (time.perf_counter() - t0) / reps
    exact = all(b.query_digest(i) == plain_digests[i] for i in range(len(queries)))
    tot = b.totals()
    n_panels, max_hops, n_chunks = b.info()
    b.close()
    log(f"batch batched: {batched_s * 1e3:.1f} ms per batch ({n_queries / batched_s:.1f} queries/s), "
        f"{n_panels} panels, bit-exact={exact}")
    if not exact:
        raise SystemExit("batch: batched digests differ from the per-query runs")
    result = {
        "workload": f"{n_queries} walk queries over {len(specs)} graphs (" + "; ".join(s.describe() for s in specs)
                    + f"), query seed {query_seed}, sizes log-uniform [1,4096], hops uniform [1,6]",
        "plain_queries_per_s": round(n_queries / plain_s, 2),
        "batched_queries_per_s": round(n_queries / batched_s, 2),
        "batched_vs_plain": round(plain_s / batched_s, 3),
        "plain_ms": round(plain_s * 1e3, 3),
        "batched_ms": round(batched_s * 1e3, 3),
        "batched_tile_products_per_s": round(tot["tile_products"] / batched_s, 1),
        "panels": n_panels,
        "chunks": n_chunks,
        "autotune": {"best": best, "table": table},
        "bit_exact": exact,
    }
    return result, specs, graphs, [m for _, m in built], best


def learned_section(M, ctx, train_specs, train_graphs, train_edges, chunk_tiles, n_queries, reps):
    """Cost model trained on logged runs of the uneven-query batch graphs (seeds 10-17,
    fresh query seeds), evaluated on a batch over new graphs (seeds 20-27): plain per-query
    walks vs the batch in its static launch order, the learned order (with the exact
    fallback), and the order by exact logged costs."""
    from matrixgraph_amd import learned as LN
    from matrixgraph_amd import workloads as WL

    t0 = time.perf_counter()
    gstats_tr = LN.graph_stats(train_graphs, train_edges)
    train_sets = [WL.batch_queries(train_specs, n_queries, s) for s in range(101, 109)]
    X, T = LN.collect(ctx, train_graphs, gstats_tr, train_sets, chunk_tiles=chunk_tiles, log=log)
    mlp, tinfo = LN.train(ctx, X, T, epochs=6, batch=4096, lr=2e-3, log=log)
    t_train = time.perf_counter() - t0

    specs = WL.batch_graph_specs(20)
    built = [WL.build_graph(ctx, s) for s in specs]
    graphs = [g for g, _ in built]
    gstats = LN.graph_stats(graphs, [m for _, m in built])
    queries = WL.batch_queries(specs, n_queries, 2)
    plain_s, plain_digests = plain_walks(M, ctx, graphs, queries, reps)

    def make():
        return M.WalkBatch(ctx, graphs, queries, chunk_tiles=chunk_tiles, use_graph=True)

    static = make()
    static_s = LN.time_batch(ctx, static, reps)
    ok_static = all(static.query_digest(i) == plain_digests[i] for i in range(len(queries)))
    static.close()
    learned = make()
    metrics = LN.schedule_learned(ctx, learned, mlp, gstats)
    learned_s = LN.time_batch(ctx, learned, reps)
    ok_learned = all(learned.query_digest(i) == plain_digests[i] for i in range(len(queries)))
    learned.close()
    exact_b = make()
    LN.schedule_exact(ctx, exact_b)
    exact_s = LN.time_batch(ctx, exact_b, reps)
    exact_b.close()
    for g in graphs:
        g.close()
    log(f"learned: plain {plain_s * 1e3:.1f} ms, static {static_s * 1e3:.1f} ms, learned {learned_s * 1e3:.1f} ms, "
        f"exact-order {exact_s * 1e3:.1f} ms, mispredict {metrics['jobs_misprediction_rate']:.3f}, "
        f"fallback {metrics['fallback_levels']}/{metrics['levels']}, bit-exact={ok_static and ok_learned}")
    if not (ok_static and ok_learned):
        raise SystemExit("learned: batched digests differ from the per-query runs")
    return {
        "workload": f"{n_queries} walk queries over 8 fresh graphs (" + "; ".join(s.describe() for s in specs)
                    + "), query seed 2",
        "model": "MLP 8->64->64->2, BF16 weights, FP32 accumulation, trained on the GPU by the library's kernels",
        "training": dict(tinfo, seconds=round(t_train, 1),
                         source="per-chunk outcomes logged from batch runs over the seeds 10-17 graphs, "
                                "query seeds 101-108"),
        "plain_queries_per_s": round(n_queries / plain_s, 2),
        "static_queries_per_s": round(n_queries / static_s, 2),
        "learned_queries_per_s": round(n_queries / learned_s, 2),
        "exact_order_queries_per_s": round(n_queries / exact_s, 2),
        "learned_batched_vs_plain": round(plain_s / learned_s, 3),
        "learned_vs_static": round(static_s / learned_s, 4),
        "exact_order_vs_static": round(static_s / exact_s, 4),
        "chunk_tiles": chunk_tiles,
        "metrics": metrics,
        "bit_exact": True,
    }


def square_section(M, ctx, scale=14, seed=1, reps=50):
    """C = A * A (boolean) of R-MAT scale 14 (edge factor 16, seed 1): the symbolic pass
    (structural tile set, allocation of C) and the numeric pass, inputs resident on the device.
    Wall time per product (host clock around the two calls and a synchronisation), device time
    on the context's stream (HIP events around the same calls), tile products per second, and
    the oracle's tile algorithm (OpenMP) on the host for the same product, checked equal."""
    import oracle as O

    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    _, m = el.info()
    el.close()

    def once():
        c, st = M.multiply_symbolic(ctx, a, at)
        M.multiply_numeric(ctx, a, at, c)
        ctx.synchronize()
        return c, st

    for _ in range(5):
        c, st = once()
        c.close()
    wall, dev = [], []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.timer_start()
        c, st = once()
        dev.append(ctx.timer_stop())
        wall.append((time.perf_counter() - t0) * 1e3)
        if len(wall) < reps:
            c.close()
    got = c.download()
    c.close()
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    ah = O.edges_to_tiles(n, s, d)
    ath = O.edges_to_tiles(n, s, d, transpose=True)
    threads = O.max_threads()
    t0 = time.perf_counter()
    cpu_reps = 3
    for _ in range(cpu_reps):
        c_off, c_row, c_col, tp = O.tile_symbolic(ah[0], ah[1], ah[3], ath[1], ath[3])
        c_data = O.tile_numeric(ah[0], ah[1:], ath[1:], c_row, c_col, threads=threads)
    cpu_ms = (time.perf_counter() - t0) / cpu_reps * 1e3
    equal = (np.array_equal(got.row_idx, c_row) and np.array_equal(got.col_idx, c_col)
             and np.array_equal(got.data, c_data) and st["tile_products"] == tp)
    if not equal:
        raise SystemExit("square: C = A*A differs from the oracle")
    wall_ms = float(np.median(wall))
    dev_ms = float(np.median(dev))
    return {
        "workload": f"C = A*A, A = R-MAT scale {scale} (edge factor 16, seed {seed}, {m} edges, {a.n_nz_tile} tiles)",
        "tile_products": st["tile_products"],
        "out_tiles": st["out_tiles"],
        "wall_ms": round(wall_ms, 4),
        "device_ms": round(dev_ms, 4),
        "wall_over_device": round(wall_ms / dev_ms, 3),
        "tile_products_per_s": round(st["tile_products"] / (wall_ms * 1e-3), 1),
        "alg_bytes": st["alg_bytes"],
        "dense_equivalent_lane_ops_per_s": round(st["tile_products"] * 8192 / (dev_ms * 1e-3), 1),
        "dense_equivalent_note": "8192 32-bit lane-ops per tile product (64x64 AND-and-test of 64-bit rows); the "
                                 "kernels multiply sparse tiles from their entry lists and do far less",
        "reps": reps,
        "bit_exact_vs_oracle": equal,
        "cpu_baseline": {"value": round(st["tile_products"] / (cpu_ms * 1e-3), 1), "unit": "tile-products/s",
                         "ms": round(cpu_ms, 3), "cores": threads, "kind": "port",
                         "sample": f"the whole product: oracle tile_symbolic + tile_numeric (OpenMP), {cpu_reps} runs"},
    }


def prealloc_section(M, ctx, train_specs, train_graphs, n_train, n_eval):
    """Learned pre-allocation of pruned products: an MLP (ROW_FEATURES -> 64 -> 64 -> 2, the
    library's kernels) trained on per-row outcomes of pruned product walks over the seeds 10-17
    graphs predicts each output row's non-empty tiles from its symbolic-pass features; the
    predicted count sizes mg_multiply_pruned's single numeric pass (exact retry on overflow).
    Evaluated on the first n_eval queries of the batch over fresh graphs (seeds 20-27, query
    seed 2), each run as a product walk X_{k+1} = X_k * A with pruned frontiers, against the two
    model-free ways to size a pruned product (exact: flag pass + write pass; structural: one
    pass into scratch for every structural tile) and against the reference's structural
    product (symbolic allocation of every structural tile, zero tiles kept)."""
    from matrixgraph_amd import learned as LN
    from matrixgraph_amd import workloads as WL

    t0 = time.perf_counter()
    train_q = WL.batch_queries(train_specs, n_train, 201)
    groups = []
    X, T, pairs = LN.collect_products(ctx, train_graphs, train_q, log=log, groups=groups)
    mlp = M.MLP(ctx, M.ROW_FEATURES, LN.ROW_TARGETS, seed=7)
    rng = np.random.default_rng(7)
    perm = rng.permutation(len(X))
    l0, _ = mlp.loss_grad(X[perm[:65536]], T[perm[:65536]])
    l1 = mlp.train(X[perm], T[perm], epochs=40, batch=512, lr=2e-3)
    margin = LN.calibrate_margin(mlp, X, T, groups)
    t_train = time.perf_counter() - t0
    log(f"prealloc: trained on {len(X)} rows ({pairs} tile pairs), loss {l0:.5f} -> {l1:.5f}, margin {margin:.3f}, "
        f"{t_train:.1f}s")
    specs = WL.batch_graph_specs(20)
    built = [WL.build_graph(ctx, sp) for sp in specs]
    graphs = [g for g, _ in built]
    queries = WL.batch_queries(specs, 512, 2)[:n_eval]
    modes = {}
    digests = {}
    for mode in ("learned", "exact", "structural"):
        tot = None
        ds = []
        for q in queries:
            d, st = LN.pruned_walk(ctx, graphs[q.graph], q.sources, q.hops, mode, mlp, margin)
            ds.append(d)
            if tot is None:
                tot = dict(st)
            else:
                for k, v in st.items():
                    tot[k] = max(tot[k], v) if k in ("scratch_bytes", "output_bytes") else tot[k] + v
        digests[mode] = ds
        modes[mode] = tot
        log(f"prealloc {mode}: {tot['seconds'] * 1e3:.1f} ms, passes {tot['passes']}, fallbacks {tot['fallbacks']}, "
            f"peak scratch {tot['scratch_bytes'] / 2**20:.1f} MiB")
    exact = digests["learned"] == digests["exact"] == digests["structural"]
    # The reference's sizing: every structural tile allocated (zero tiles kept in X_k).
    ref_s = 0.0
    ref_bytes = 0
    for q in queries:
        xs, stats = WL.product_walk(ctx, graphs[q.graph], graphs[q.graph].n_strips * 64, q.sources, q.hops, None)
        ref_s += sum(st["symbolic_ms"] + st["numeric_ms"] for st in stats) * 1e-3
        ref_bytes = max(ref_bytes, max(x.n_nz_tile for x in xs[1:]) * 512)
        for x in xs:
            x.close()
    for g in graphs:
        g.close()
    mlp.close()
    if not exact:
        raise SystemExit("prealloc: pruned walks differ between sizing modes")
    L, E, S = modes["learned"], modes["exact"], modes["structural"]
    return {
        "workload": f"first {n_eval} of the 512 queries over 8 fresh graphs (seeds 20-27, query seed 2), each as a "
                    "product walk X_{k+1} = X_k * A with pruned frontiers",
        "model": f"MLP {M.ROW_FEATURES}->64->64->{LN.ROW_TARGETS}, BF16 weights, FP32 accumulation, trained on the GPU",
        "training": {"rows": int(len(X)), "tile_pairs": int(pairs), "queries": n_train,
                     "source": "per-row outcomes of pruned product walks over the seeds 10-17 graphs, query seed 201",
                     "loss_initial": round(float(l0), 6), "loss_final": round(float(l1), 6),
                     "seconds": round(t_train, 1),
                     "capacity_margin": round(margin, 4),
                     "margin_rule": "95th percentile over the training products of actual/predicted - 1"},
        "products": L["products"],
        "learned_ms": round(L["seconds"] * 1e3, 2),
        "exact_two_pass_ms": round(E["seconds"] * 1e3, 2),
        "structural_scratch_ms": round(S["seconds"] * 1e3, 2),
        "reference_structural_ms": round(ref_s * 1e3, 2),
        "learned_vs_exact": round(E["seconds"] / L["seconds"], 3),
        "learned_vs_structural_scratch": round(S["seconds"] / L["seconds"], 3),
        "numeric_passes": {"learned": L["passes"], "exact": E["passes"], "structural": S["passes"]},
        "fallbacks": L["fallbacks"],
        "fallback_frequency": round(L["fallbacks"] / max(L["products"], 1), 4),
        "rows_predicted": L["rows"],
        "row_misprediction_rate": round(L["rows_mispredicted"] / max(L["rows"], 1), 4),
        "misprediction_definition": "predicted/actual non-empty tiles of a row outside [1/2, 2] (+1 smoothing)",
        "predicted_tiles": L["predicted_tiles"], "numeric_tiles": L["numeric_tiles"],
        "capacity_tiles": L["capacity_tiles"],
        "allocated_tile_slots": {"reference_structural": L["structural_tiles"], "learned": L["allocated_tiles"],
                                 "exact": E["allocated_tiles"], "structural_scratch": S["allocated_tiles"]},
        "peak_scratch_bytes": {"learned": L["scratch_bytes"], "exact": E["scratch_bytes"],
                               "structural": S["scratch_bytes"]},
        "peak_output_bytes_reference_structural": ref_bytes,
        "peak_output_bytes_pruned": L["output_bytes"],
        "bit_exact": exact,
    }


def mfma_section(M, ctx, densities=(0.01, 0.03, 0.1, 0.3), reps=3):
    """Dense tile pairs on the matrix cores against the vector ALUs: C = A * A of 64 planted
    communities of 4096 vertices (community c = vertices [16384c, 16384c + 4096) of 2^20) at
    edge probability p, so every C tile of a community sums 64 pairs of tiles holding about
    4096 p bits each (16.8 M pairs). Per p: the numeric pass with every pair on the matrix cores
    (min_bits 0) and with none (min_bits 4097), bit-exact against each other; the crossover
    density sets where the dispatch threshold pays."""
    from matrixgraph_amd import workloads as WL

    n = 1 << 20
    out = []
    for p in densities:
        num = int(round(p * 2**32))
        el = M.EdgeList.communities(ctx, n, 64, 4096, 16384, num, 32, 3)
        a = el.to_tiled_matrix(transpose=False)
        at = el.to_tiled_matrix(transpose=True)
        _, m = el.info()
        el.close()
        row = {"p": p, "edges": m, "tiles": a.n_nz_tile}
        digests = {}
        for mb in (4097, 0):
            times = []
            for r in range(reps + 1):
                c, st = M.multiply_symbolic(ctx, a, at)
                ctx.synchronize()
                ctx.timer_start()
                dense = M.multiply_numeric_dense(ctx, a, at, c, mb)
                ms = ctx.timer_stop()
                if r:
                    times.append(ms)
                if r == reps:
                    digests[mb] = c.digest()
                c.close()
            key = "vector" if mb > 4096 else "matrix_cores"
            row[key + "_ms"] = round(float(np.median(times)), 3)
            row[key + "_products"] = dense
            row["tile_products"] = st["tile_products"]
        row["matrix_cores_vs_vector"] = round(row["vector_ms"] / row["matrix_cores_ms"], 3)
        row["int8_ops_per_s_matrix_cores"] = round(row["matrix_cores_products"] * 2 * 64**3
                                                   / (row["matrix_cores_ms"] * 1e-3), 1)
        row["bit_exact"] = digests[4097] == digests[0]
        if not row["bit_exact"]:
            raise SystemExit("mfma: matrix-core product differs from the vector product")
        log(f"mfma p={p}: vector {row['vector_ms']} ms, matrix cores {row['matrix_cores_ms']} ms "
            f"({row['matrix_cores_vs_vector']}x), {row['tile_products']} pairs")
        out.append(row)
        a.close()
        at.close()
    wins = [r["p"] for r in out if r["matrix_cores_vs_vector"] > 1.0]
    return {
        "workload": "C = A*A, A = 64 communities of 4096 vertices (stride 16384, 2^20 vertices) at edge "
                    "probability p, seed 3; numeric pass only (HIP events), symbolic pass excluded",
        "per_density": out,
        "matrix_cores_win_from_p": min(wins) if wins else None,
        "int8_peak_ops_per_s": 5.0e15,
    }


def dense_section(M, ctx, thresholds=(4097, 2048, 1024, 512, 0, None)):
    """Density-mixed 3-hop walk: R-MAT scale 20 (edge factor 16, seed 3) union 64 planted
    communities of 4096 vertices (community c = vertices [16384c, 16384c + 4096), p = 0.3),
    4096 sources (seed 3). The gather walk is timed, then the same walk by tile products
    X_{k+1} = X_k * A with the matrix-core path for tile pairs whose two tiles both have at
    least `min_bits` set bits (4097 = vector path only; None = mg_multiply_numeric, which
    multiplies compact tiles from their entry lists); every variant's X_3 is compared with the
    gather walk's export."""
    from matrixgraph_amd import workloads as WL

    scale, hops, n_src = 20, 3, 4096
    n = 1 << scale
    t0 = time.perf_counter()
    el = WL.union_graph(ctx, scale, 3, 64, 4096, 16384)
    _, m = el.info()
    at = el.to_tiled_matrix(transpose=True)
    el.close()
    sources = M.sample_sources(n, n_src, 3)
    log(f"dense: union graph {m} edges, {at.n_nz_tile} tiles, built in {time.perf_counter() - t0:.1f}s")
    w = M.Walk(ctx, at, n_src, hops)
    w.run(sources, hops)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        w.run(sources, hops)
    ctx.synchronize()
    gather_ms = (time.perf_counter() - t0) / 3 * 1e3
    ref = w.export(hops)
    ref_digest = ref.digest()
    tp_total = sum(w.hop_stats(k)["tile_products"] for k in range(1, hops + 1))
    ref.close()
    w.close()
    variants = []
    for mb in thresholds:
        xs, stats = WL.product_walk(ctx, at, n, sources, hops, mb)
        ok = xs[-1].digest() == ref_digest
        num_ms = sum(st["numeric_ms"] for st in stats)
        sym_ms = sum(st["symbolic_ms"] for st in stats)
        dense = sum(st["dense_products"] for st in stats)
        tp = sum(st["tile_products"] for st in stats)
        for x in xs:
            x.close()
        log(f"dense: min_bits {mb}: numeric {num_ms:.1f} ms, symbolic {sym_ms:.1f} ms, matrix-core products "
            f"{dense}/{tp}, bit-exact={ok}")
        if not ok:
            raise SystemExit("dense: tile-product walk differs from the gather walk")
        variants.append({"min_bits": mb, "engine": "vector, compact tiles from entry lists" if mb is None else
                         ("vector" if mb > 4096 else "matrix cores for pairs of tiles with >= min_bits bits"),
                         "numeric_ms": round(num_ms, 3), "symbolic_ms": round(sym_ms, 3),
                         "matrix_core_products": dense, "tile_products": tp,
                         "tile_products_per_s_numeric": round(tp / (num_ms * 1e-3), 1)})
    at.close()
    best = min((v for v in variants if v["min_bits"] is not None), key=lambda v: v["numeric_ms"])
    payload = [v for v in variants if v["min_bits"] is None]
    return {
        "workload": "3-hop walk from 4096 sources on R-MAT scale 20 (edge factor 16, seed 3) union 64 communities "
                    "of 4096 vertices at stride 16384, p = 1288490189/2^32, seed 3",
        "edges": m,
        "gather_walk_ms": round(gather_ms, 3),
        "gather_walk_source_walks_per_s": round(n_src / (gather_ms * 1e-3), 1),
        "tile_products": tp_total,
        "product_walk_variants": variants,
        "best_min_bits": best["min_bits"],
        "payload_numeric_ms": payload[0]["numeric_ms"] if payload else None,
        "bit_exact": True,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sources", type=int, default=4096)
    ap.add_argument("--hops", type=int, default=4)
    ap.add_argument("--cpu-stride", type=int, default=1, help="CPU baseline samples every n-th strip (1: all)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="CPU baseline repetitions (timings averaged)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-query batch section")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-scheduling section")
    ap.add_argument("--no-dense", action="store_true", help="skip the density-mixed (matrix-core) section")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the uneven-query batch section as its own JSON line")
    ap.add_argument("--mfma-only", action="store_true",
                    help="only the matrix-core community A*A section as its own JSON line")
    ap.add_argument("--no-prealloc", action="store_true", help="skip the learned pre-allocation section")
    ap.add_argument("--prealloc-train", type=int, default=384, help="training queries of the pre-allocation model")
    ap.add_argument("--prealloc-eval", type=int, default=128, help="evaluated queries of the pre-allocation section")
    ap.add_argument("--batch-queries", type=int, default=512)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--prune", action="store_true", help="structural tile sets from numerically non-empty tiles")
    ap.add_argument("--sweep-first-hop", action="store_true",
                    help="sweep every strip at the first hop too (no push from the source strips)")
    ap.add_argument("--square-only", action="store_true",
                    help="only the C = A*A section (R-MAT 14, seed 1) as its own JSON line")
    ap.add_argument("--no-square", action="store_true", help="skip the C = A*A section")
    ap.add_argument("--hop-kernel", type=int, default=None, metavar="SEGMENTS",
                    help="hop kernel after the first hop: 0 sweep, 1/2/4/8 job list (default: the library's)")
    ap.add_argument("--vertex-order", action="store_true",
                    help="keep frontier strips in vertex order (no internal strip permutation)")
    ap.add_argument("--two-hop", type=int, default=0, metavar="SCALE",
                    help="only time the two-hop pair enumeration (mg_edges_two_hop) on an R-MAT graph of this "
                         "scale (edge factor 16, seed 2), beside the oracle's bitset product on the host")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import matrixgraph_amd as M

    ctx = M.Context(0)
    if args.two_hop:
        two_hop(M, ctx, args.two_hop)
        return
    if args.square_only:
        print(json.dumps({"square": square_section(M, ctx)}), flush=True)
        return
    if args.mfma_only:
        print(json.dumps({"mfma": mfma_section(M, ctx)}), flush=True)
        return
    if args.batch_only:
        res = batch_section(M, ctx, 10, 1, args.batch_queries, args.batch_reps)[0]
        print(json.dumps({"batch": res}), flush=True)
        return
    ctx.synchronize()
    t0 = time.perf_counter()
    el = M.EdgeList.rmat(ctx, args.scale, args.edge_factor, 0.57, 0.19, 0.19, args.seed)
    n_vertices, n_edges = el.info()
    ctx.synchronize()
    t1 = time.perf_counter()
    at = el.to_tiled_matrix(transpose=True)
    ctx.synchronize()
    t2 = time.perf_counter()
    el.close()
    ns, n_tiles, _ = at.metadata()
    log(f"graph: 2^{args.scale} vertices, {n_edges} edges, {n_tiles} A^T tiles, setup {t2 - t0:.1f}s")
    sources = M.sample_sources(n_vertices, args.sources, args.seed)
    flags = (M.WALK_PRUNE if args.prune else 0) | (M.WALK_SWEEP_FIRST_HOP if args.sweep_first_hop else 0)
    flags |= M.WALK_VERTEX_ORDER if args.vertex_order else 0
    t3 = time.perf_counter()
    w = M.Walk(ctx, at, args.sources, args.hops, flags)
    if args.hop_kernel is not None:
        w.set_hop_kernel(args.hop_kernel)
    ctx.synchronize()
    t4 = time.perf_counter()
    setup = {
        "note": "per graph, before the timed walks (not in value)",
        "edge_list_ms": round((t1 - t0) * 1e3, 1),
        "tile_build_ms": round((t2 - t1) * 1e3, 1),
        "walk_create_ms": round((t4 - t3) * 1e3, 1),
        "walk_create_covers": "A^T index relabelled to the engine's strip order, row masks, compact payloads, "
                              "dense-tile rows, scheduler chunks and costs, column index for the pushed first hop, "
                              "job list and column counts of the hop kernel",
    }

    for _ in range(args.warmup):
        w.run(sources, args.hops)
    ctx.synchronize()
    w.set_timing(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.run(sources, args.hops)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    w.set_timing(False)

    ms_per_step = 1e3 * elapsed / args.steps
    per_hop = []
    total_tp = 0
    total_gathers = 0
    launches = []  # the roofline kernel's launches (hops that sweep every strip)
    kernel_names = {M.HOP_SWEEP: "sweep", M.HOP_PUSH: "push", M.HOP_JOBS: "jobs"}
    roof_kernel = None  # device kernel name of the hops after the first
    for k in range(1, args.hops + 1):
        st = w.hop_stats(k)
        sym_ms, num_ms, n_runs = w.kernel_times(k)
        avg_num = num_ms / max(n_runs, 1)
        avg_sym = sym_ms / max(n_runs, 1)
        total_tp += st["tile_products"]
        total_gathers += st["edges_visited"]
        kname = kernel_names[st["kernel"]]
        if st["kernel"] in (M.HOP_SWEEP, M.HOP_JOBS):
            roof_kernel = ("walk_numeric64_kernel" if st["kernel"] == M.HOP_SWEEP else
                           "walk_lean_kernel" if st["segments"] == 1 else f"walk_seg_kernel<{st['segments']}>")
            launches.append({"hop": k, "alg_bytes": st["alg_bytes_numeric"], "ms": round(avg_num, 4),
                             "frac_alg": round(st["alg_bytes_numeric"] / (avg_num * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)})
        per_hop.append({
            "hop": k, "kernel": kname, "segments": st["segments"], "tile_products": st["tile_products"],
            "alg_bytes": st["alg_bytes"], "alg_bytes_numeric": st["alg_bytes_numeric"],
            "left_tiles": st["left_tiles"], "left_tiles_numeric": st["left_tiles_numeric"],
            "right_tiles_used": st["right_tiles_used"], "right_tiles_used_numeric": st["right_tiles_used_numeric"],
            "out_tiles": st["out_tiles"], "out_tiles_numeric": st["out_tiles_numeric"],
            "popcount": st["popcount"], "gathers": st["edges_visited"],
            "numeric_kernel_ms": round(avg_num, 4), "symbolic_ms": round(avg_sym, 4),
            "GBps_alg_numeric": round(st["alg_bytes_numeric"] / (avg_num * 1e-3) / 1e9, 1) if avg_num > 0 else None,
        })
        log(f"hop {k} ({kname}): TP={st['tile_products']} B_alg={st['alg_bytes'] / 1e9:.2f}GB "
            f"B_alg(numeric)={st['alg_bytes_numeric'] / 1e9:.2f}GB kernels={avg_num:.3f}ms popcount={st['popcount']}")
    # The roofline is stated over the steady hops (3 and later), where X_k is numerically full.
    # At hop 2 the byte model charges every A^T tile in a used column at 520 B although the
    # kernel reads one 4-byte job per set bit, and most of its row gathers hit in the caches
    # (X_1 is sparse), so the model's bytes exceed what the launch could move at peak: that
    # launch is listed with its model figure (frac_alg above 1, not a claim) and its measured
    # HBM-side traffic, and is not averaged in.
    steady = [x for x in launches if x["hop"] >= 3] or launches
    traffic = None
    hop_traffic = {}
    tname = {"walk_numeric64_kernel": "walk_numeric", "walk_lean_kernel": "walk_lean"}.get(roof_kernel)
    tfile = ROOT / "profiles" / f"{tname}_rmat22_traffic.json"
    if (tname and tfile.exists() and args.scale == 22 and args.sources == 4096 and args.hops == 4 and not args.prune
            and not args.vertex_order):
        tj = json.loads(tfile.read_text())
        traffic = tj.get("bytes_per_steady_launch")
        hop_traffic = {int(h): v["read_bytes"] + v["write_bytes"] for h, v in tj.get("per_hop", {}).items()}
    for x in launches:
        x["in_roofline"] = x in steady
        if x["hop"] in hop_traffic:
            x["traffic_bytes"] = hop_traffic[x["hop"]]
            x["frac_traffic"] = round(hop_traffic[x["hop"]] / (x["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if x["frac_alg"] > 1:
            x["note"] = ("model bytes above what the launch can move: A^T is read as 4-byte jobs, not 520-byte "
                         "tiles, and the sparse X_k rows are gathered from the caches; frac_traffic is the "
                         "HBM-side fraction")
    n_launch = max(len(steady), 1)
    sum_bytes = sum(x["alg_bytes"] for x in steady)
    sum_ms = sum(x["ms"] for x in steady)
    achieved = (sum_bytes / n_launch) / ((sum_ms / n_launch) * 1e-3) / 1e9 if sum_ms > 0 else 0.0
    roofline = {
        "kernel": roof_kernel,
        "launches": f"hops {', '.join(str(x['hop']) for x in steady)} of each walk (steady: X_k numerically full)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "frac_traffic": (round(traffic / ((sum_ms / n_launch) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                         if traffic and sum_ms > 0 else None),
        "traffic_source": ("stored: HBM-side bytes per steady launch (TCC_EA0_RDREQ_32B/64B/128B, WRREQ/WRREQ_64B) "
                           f"from rocprofv3 --pmc runs of this bench, profiles/{tname}_rmat22_traffic.md; "
                           "not measured in this run") if traffic else None,
        "alg_bytes_per_launch": sum_bytes // n_launch,
        "alg_bytes_definition": "520*(T_L + T_R,used + T_C) + 4*(2*rows + 2) over numerically non-empty tiles "
                                "(T_L: X_k tiles, T_R,used: A^T tiles in their rows, T_C: X_{k+1} tiles)",
        "avg_launch_ms": round(sum_ms / n_launch, 4),
        "per_launch": launches,
    }

    result = {
        "metric": METRIC,
        "value": round(args.sources / (ms_per_step * 1e-3), 1),
        "unit": "source-walks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_walk{args.hops}hop_{args.sources}src",
            "graph": f"R-MAT scale {args.scale}, edge factor {args.edge_factor}, (a,b,c,d)=(0.57,0.19,0.19,0.05), "
                     f"seed {args.seed}, duplicates/self-loops removed, {n_edges} edges, {n_tiles} tiles",
            "sources": args.sources,
            "hops": args.hops,
            "tile_size": 64,
            "structural_mode": "prune" if args.prune else "canonical",
            "strip_order": "vertex" if args.vertex_order else "engine (fixed permutation)",
            "parallelism": "single-gpu",
        },
        "walk_queries_per_s": round(1.0 / (ms_per_step * 1e-3), 3),
        "tile_products_per_s": round(total_tp / (ms_per_step * 1e-3), 1),
        "tile_products_per_walk": total_tp,
        "tile_products_note": "structural (r,k,c) triples counted by the symbolic pass; gathers = A^T bits whose "
                              "source row is non-empty (one 512-B frontier-row gather each in the sweep kernel; "
                              "the one-segment job-list kernel gathers a source row once per target strip and "
                              "ORs it into all of that strip's rows it reaches)",
        "gathers_per_walk": total_gathers,
        "gathers_per_s": round(total_gathers / (ms_per_step * 1e-3), 1),
        "setup": setup,
        "per_hop": per_hop,
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        import oracle as O

        threads = O.max_threads()
        result["cpu_baseline"] = cpu_baseline(M, O, ctx, w, at, n_edges, args.hops, args.sources, args.cpu_stride,
                                              threads, reps=max(1, args.cpu_reps))
    if not args.no_square:
        result["square"] = square_section(M, ctx)
    if not args.no_dense:
        result["dense"] = dense_section(M, ctx)
        result["mfma"] = mfma_section(M, ctx)
    if not args.no_batch:
        w.close()
        result["batch"], specs, graphs, edges, best = batch_section(M, ctx, 10, 1, args.batch_queries, args.batch_reps)
        if not args.no_learned:
            result["learned"] = learned_section(M, ctx, specs, graphs, edges, best["chunk_tiles"], args.batch_queries,
                                                args.batch_reps)
            result["learned_batched_vs_plain"] = result["learned"]["learned_batched_vs_plain"]
        if not args.no_prealloc:
            result["learned_prealloc"] = prealloc_section(M, ctx, specs, graphs, args.prealloc_train,
                                                          args.prealloc_eval)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
