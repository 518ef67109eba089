"""Learned scheduling of batched walks: per-chunk outcome logging, the MLP cost model trained
on the GPU from logged runs, and launch orders chosen from its predictions. Results must not
depend on the order (bit-exact against the static order and single walks)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def setup(ctx, seeds=(41, 42, 43)):
    import matrixgraph_amd as M
    from matrixgraph_amd import workloads as WL

    specs = [WL.GraphSpec("rmat", 12, seeds[0]), WL.GraphSpec("rmat", 13, seeds[1]), WL.GraphSpec("er", 12, seeds[2])]
    built = [WL.build_graph(ctx, s) for s in specs]
    return M, WL, specs, [g for g, _ in built], [m for _, m in built]


def test_logging_records_match_counters(ctx):
    M, WL, specs, graphs, edges = setup(ctx)
    qs = WL.batch_queries(specs, 24, 7, max_sources=600, max_hops=4)
    # Every level-0 chunk swept (and logged): pushed first hops have no chunk records.
    b = M.WalkBatch(ctx, graphs, qs, chunk_tiles=64, use_graph=True, sweep_first_hop=True)
    assert b.pushed_panels() == 0
    b.set_logging(True)
    b.run()
    ctx.synchronize()
    total_jobs = 0
    for k in range(b.info()[1]):
        rec = b.level_chunks(k)
        assert np.all(rec["jobs"] <= rec["bits"])
        assert np.all(rec["tile_end"] >= rec["tile_begin"])
        total_jobs += int(rec["jobs"].sum())
    assert total_jobs == b.totals()["edges_visited"]
    # bits of a level's chunks of one panel cover the whole graph once
    rec0 = b.level_chunks(0)
    p0 = rec0["panel"][0]
    g = b.panel_info(int(p0))["graph"]
    assert int(rec0["bits"][rec0["panel"] == p0].sum()) == edges[g]


def test_order_does_not_change_results_and_rejects_bad_orders(ctx):
    M, WL, specs, graphs, edges = setup(ctx)
    qs = WL.batch_queries(specs, 30, 9, max_sources=700, max_hops=5)
    b = M.WalkBatch(ctx, graphs, qs, chunk_tiles=64, use_graph=True)
    b.run()
    ref = [b.query_digest(i) for i in range(len(qs))]
    rng = np.random.default_rng(3)
    for k in range(b.info()[1]):
        n = len(b.level_chunks(k))
        b.set_level_order(k, rng.permutation(n))
    b.run()
    b.run()
    assert [b.query_digest(i) for i in range(len(qs))] == ref
    n = len(b.level_chunks(0))
    with pytest.raises(M.MGError):
        b.set_level_order(0, np.zeros(n, np.uint32))
    with pytest.raises(M.MGError):
        b.set_level_order(0, np.arange(n - 1))


def test_learned_schedule_trains_and_stays_bit_exact(ctx):
    M, WL, specs, graphs, edges = setup(ctx)
    from matrixgraph_amd import learned as LN

    gstats = LN.graph_stats(graphs, edges)
    train_sets = [WL.batch_queries(specs, 40, s, max_sources=800, max_hops=5) for s in (11, 12, 13)]
    X, T = LN.collect(ctx, graphs, gstats, train_sets, chunk_tiles=64)
    assert X.shape[1] == LN.N_FEATURES and len(X) == len(T) and len(X) > 1000
    assert np.isfinite(X).all() and np.isfinite(T).all()
    mlp, info = LN.train(ctx, X, T, epochs=6, batch=1024, lr=3e-3)
    assert info["loss_final"] < info["loss_initial"]
    qs = WL.batch_queries(specs, 40, 99, max_sources=800, max_hops=5)
    static = M.WalkBatch(ctx, graphs, qs, chunk_tiles=64, use_graph=True)
    static.run()
    ref = [static.query_digest(i) for i in range(len(qs))]
    learned = M.WalkBatch(ctx, graphs, qs, chunk_tiles=64, use_graph=True)
    metrics = LN.schedule_learned(ctx, learned, mlp, gstats)
    assert 0.0 <= metrics["jobs_misprediction_rate"] <= 1.0
    assert 0 <= metrics["fallback_levels"] <= metrics["levels"]
    learned.run()
    assert [learned.query_digest(i) for i in range(len(qs))] == ref
    # single-walk parity of a few queries
    for i in range(0, len(qs), 9):
        q = qs[i]
        w = M.Walk(ctx, graphs[q.graph], len(q.sources), q.hops)
        w.run(q.sources, q.hops)
        assert w.frontier_digest(q.hops) == ref[i]
        w.close()
