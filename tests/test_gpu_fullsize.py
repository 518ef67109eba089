"""The density-mixed, batched and learned-schedule walk workloads at their full sizes, against
the CPU oracle:

- density-mixed 3-hop walk: R-MAT scale 20 (edge factor 16, seed 3) union 64 planted
  communities of 4096 vertices (community c = vertices [16384c, 16384c + 4096), p = 0.3),
  4096 sources (seed 3); the gather walk's per-hop digests equal the oracle's vertex-major
  frontiers (O3), and the walk by tile products (matrix-core threshold 512 and the
  entry-list product) equals the gather walk at every hop;
- 512 uneven queries over 8 graphs (R-MAT scales 16-20, seeds 10-14; three ER n = 2^18,
  p = 2^-13, seeds 15-17), query seed 1, one captured HIP graph: every query equals its
  single walk (digest) and every 16th query equals O3 (arrays);
- the learned launch order on 8 fresh graphs (seeds 20-27, query seed 2): every query equals
  the static order's result and its single walk, every 32nd query equals O3.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def oracle_in_csr(spec_or_edges, n):
    s, d = spec_or_edges
    return O.in_csr(n, s, d)


def oracle_query_frontier(n, csr, sources, hops):
    off, isrc = csr
    x = O.one_hot_frontier(n, sources, (len(sources) + 63) // 64)
    for _ in range(hops):
        x = O.frontier_step(n, off, isrc, x)
    return x


def test_density_mixed_walk_full_size(ctx):
    M = mg()
    from matrixgraph_amd import workloads as WL

    scale, seed, n_comm, comm_size, stride, n_src, hops = 20, 3, 64, 4096, 16384, 4096, 3
    n = 1 << scale
    print("density-mixed: union graph R-MAT 20 + 64 x 4096 communities, seed 3", flush=True)
    el = WL.union_graph(ctx, scale, seed, n_comm, comm_size, stride)
    s1, d1 = O.rmat_edges(scale, 16, *WL.RMAT_ABC, seed)
    s2, d2 = O.community_edges(n, n_comm, comm_size, stride, WL.COMMUNITY_P_NUM, WL.COMMUNITY_P_LOG2, seed)
    s, d = O.sort_unique_edges(np.concatenate([s1, s2]), np.concatenate([d1, d2]))
    del s1, d1, s2, d2
    assert el.info() == (n, len(s))
    assert el.digest() == O.edges_digest(s, d)
    at = el.to_tiled_matrix(transpose=True)
    el.close()
    csr = O.in_csr(n, s, d)
    del s, d
    sources = M.sample_sources(n, n_src, seed)
    assert np.array_equal(sources, O.sample_sources(n, n_src, seed))
    w = M.Walk(ctx, at, n_src, hops)
    w.run(sources, hops)
    x = O.one_hot_frontier(n, sources, 64)
    exports = []
    for k in range(1, hops + 1):
        x = O.frontier_step(n, csr[0], csr[1], x)
        assert w.frontier_digest(k) == O.frontier_digest(x, n_src), f"gather walk hop {k}"
        e = w.export(k)
        exports.append(e.digest())
        e.close()
    w.close()
    del x, csr
    for min_bits in (512, None):
        print(f"density-mixed: walk by tile products, min_bits {min_bits}", flush=True)
        xs, stats = WL.product_walk(ctx, at, n, sources, hops, min_bits)
        for k in range(1, hops + 1):
            assert xs[k].digest() == exports[k - 1], f"product walk min_bits {min_bits} hop {k}"
        if min_bits == 512:
            assert sum(st["dense_products"] for st in stats) > 0
        for m in xs:
            m.close()
    at.close()


def _plain_digests(M, ctx, graphs, queries):
    """Each query run alone (one walk object per graph and frontier width)."""
    walks = {}
    out = []
    for q in queries:
        W = 1
        while W * 64 < len(q.sources):
            W *= 2
        key = (q.graph, W)
        if key not in walks:
            walks[key] = M.Walk(ctx, graphs[q.graph], 64 * W, 6)
        w = walks[key]
        w.run(q.sources, q.hops)
        out.append(w.frontier_digest(q.hops))
    for w in walks.values():
        w.close()
    return out


def _check_oracle_subset(b, specs, queries, every):
    from matrixgraph_amd import workloads as WL

    by_graph = {}
    for i in range(0, len(queries), every):
        by_graph.setdefault(queries[i].graph, []).append(i)
    for g, idx in sorted(by_graph.items()):
        n = specs[g].n_vertices
        csr = oracle_in_csr(WL.oracle_edges(specs[g]), n)
        for i in idx:
            q = queries[i]
            want = oracle_query_frontier(n, csr, q.sources, q.hops)
            assert np.array_equal(b.query_frontier(i), want), f"query {i} ({specs[g].describe()})"


def test_uneven_query_batch_full_size(ctx):
    M = mg()
    from matrixgraph_amd import workloads as WL

    specs = WL.batch_graph_specs(10)
    print("batch: 8 graphs, seeds 10-17, 512 queries, query seed 1", flush=True)
    graphs = [WL.build_graph(ctx, s)[0] for s in specs]
    queries = WL.batch_queries(specs, 512, 1)
    b = M.WalkBatch(ctx, graphs, queries, use_graph=True)
    b.run()
    b.run()  # graph replay
    ctx.synchronize()
    plain = _plain_digests(M, ctx, graphs, queries)
    got = [b.query_digest(i) for i in range(len(queries))]
    bad = [i for i in range(len(queries)) if got[i] != plain[i]]
    assert not bad, f"queries differing from their single walks: {bad[:10]}"
    _check_oracle_subset(b, specs, queries, 16)
    b.close()
    for g in graphs:
        g.close()


def test_learned_schedule_full_size(ctx):
    M = mg()
    from matrixgraph_amd import learned as LN
    from matrixgraph_amd import workloads as WL

    print("learned: training graphs seeds 10-17, two logged query sets", flush=True)
    tr_specs = WL.batch_graph_specs(10)
    tr_built = [WL.build_graph(ctx, s) for s in tr_specs]
    tr_graphs = [g for g, _ in tr_built]
    gstats_tr = LN.graph_stats(tr_graphs, [m for _, m in tr_built])
    train_sets = [WL.batch_queries(tr_specs, 512, s) for s in (101, 102)]
    X, T = LN.collect(ctx, tr_graphs, gstats_tr, train_sets)
    for g in tr_graphs:
        g.close()
    mlp, info = LN.train(ctx, X, T, epochs=3, batch=4096, lr=2e-3)
    assert info["loss_final"] < info["loss_initial"]
    print("learned: evaluation graphs seeds 20-27, 512 queries, query seed 2", flush=True)
    specs = WL.batch_graph_specs(20)
    built = [WL.build_graph(ctx, s) for s in specs]
    graphs = [g for g, _ in built]
    gstats = LN.graph_stats(graphs, [m for _, m in built])
    queries = WL.batch_queries(specs, 512, 2)
    static = M.WalkBatch(ctx, graphs, queries, use_graph=True)
    static.run()
    ref = [static.query_digest(i) for i in range(len(queries))]
    static.close()
    learned = M.WalkBatch(ctx, graphs, queries, use_graph=True)
    metrics = LN.schedule_learned(ctx, learned, mlp, gstats)
    assert 0.0 <= metrics["jobs_misprediction_rate"] <= 1.0
    learned.run()
    learned.run()
    got = [learned.query_digest(i) for i in range(len(queries))]
    assert got == ref, "learned launch order changed a query's result"
    plain = _plain_digests(M, ctx, graphs, queries)
    bad = [i for i in range(len(queries)) if got[i] != plain[i]]
    assert not bad, f"queries differing from their single walks: {bad[:10]}"
    _check_oracle_subset(learned, specs, queries, 32)
    learned.close()
    mlp.close()
    for g in graphs:
        g.close()
