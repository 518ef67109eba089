"""Batched walks: many queries of uneven size and hop count packed into panels, run as one
captured HIP graph; every query's result equals the same query run alone (digest and
arrays) and the CPU oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def small_workload(ctx, n_queries=40, seed=5):
    from matrixgraph_amd import workloads as WL

    specs = [WL.GraphSpec("rmat", 12, 31), WL.GraphSpec("rmat", 13, 32), WL.GraphSpec("er", 12, 33)]
    graphs = [WL.build_graph(ctx, s)[0] for s in specs]
    qs = WL.batch_queries(specs, n_queries, seed, max_sources=700, max_hops=5)
    return specs, graphs, qs


@pytest.mark.parametrize("packing,use_graph", [(0, True), (0, False), (1, True)])
def test_batch_matches_single_walks_and_oracle(ctx, packing, use_graph):
    M = mg()
    from matrixgraph_amd import workloads as WL

    specs, graphs, qs = small_workload(ctx)
    b = M.WalkBatch(ctx, graphs, qs, packing=packing, use_graph=use_graph)
    b.run()
    b.run()  # replay: results must not depend on the previous run
    ctx.synchronize()
    n_panels, max_hops, _ = b.info()
    if packing == 1:
        assert n_panels == len(qs)
    else:
        assert n_panels < len(qs)
    edges = [WL.oracle_edges(s) for s in specs]
    for i, q in enumerate(qs):
        g = q.graph
        n = specs[g].n_vertices
        w = M.Walk(ctx, graphs[g], len(q.sources), q.hops)
        w.run(q.sources, q.hops)
        assert b.query_digest(i) == w.frontier_digest(q.hops), i
        if i % 7 == 0:
            xs = O.walk_frontiers(n, edges[g][0], edges[g][1], q.sources, q.hops, W=(len(q.sources) + 63) // 64)
            assert np.array_equal(b.query_frontier(i), xs[q.hops]), i
        w.close()


def test_batch_export_and_errors(ctx):
    M = mg()
    specs, graphs, qs = small_workload(ctx, 12, 9)
    b = M.WalkBatch(ctx, graphs, qs)
    b.run()
    q = qs[3]
    w = M.Walk(ctx, graphs[q.graph], len(q.sources), q.hops)
    w.run(q.sources, q.hops)
    e1 = b.query_export(3).download()
    e2 = w.export(q.hops).download()
    assert np.array_equal(e1.row_idx, e2.row_idx) and np.array_equal(e1.data, e2.data)
    bad = [M.workloads.QuerySpec(0, np.array([1, 1], np.uint32), 2)]
    with pytest.raises(M.MGError):
        M.WalkBatch(ctx, graphs, bad)
    bad = [M.workloads.QuerySpec(7, np.array([1], np.uint32), 2)]
    with pytest.raises(M.MGError):
        M.WalkBatch(ctx, graphs, bad)


@pytest.mark.parametrize("packing,n_queries,max_sources", [(0, 12, 60), (1, 30, 300)])
def test_batch_first_hop_push_equals_sweep(ctx, packing, n_queries, max_sources):
    """Panels whose first hop is pushed (few source strips on the larger graph) give the same
    per-query frontiers and batch totals as sweeping every panel's first hop; panels of the
    small graph stay swept in the same batch."""
    M = mg()
    from matrixgraph_amd import workloads as WL

    specs = [WL.GraphSpec("rmat", 16, 51), WL.GraphSpec("er", 12, 52)]
    graphs = [WL.build_graph(ctx, s)[0] for s in specs]
    qs = WL.batch_queries(specs, n_queries, 11, max_sources=max_sources, max_hops=4)
    bp = M.WalkBatch(ctx, graphs, qs, packing=packing)
    bs = M.WalkBatch(ctx, graphs, qs, packing=packing, sweep_first_hop=True)
    assert bp.pushed_panels() > 0 and bs.pushed_panels() == 0
    for b in (bp, bs):
        b.run()
        b.run()
    ctx.synchronize()
    assert bp.totals() == bs.totals()
    for i in range(len(qs)):
        assert bp.query_digest(i) == bs.query_digest(i), i
    edges = [WL.oracle_edges(s) for s in specs]
    for i in range(0, len(qs), 5):
        q = qs[i]
        n = specs[q.graph].n_vertices
        xs = O.walk_frontiers(n, edges[q.graph][0], edges[q.graph][1], q.sources, q.hops,
                              W=(len(q.sources) + 63) // 64)
        assert np.array_equal(bp.query_frontier(i), xs[q.hops]), i


@pytest.mark.parametrize("segments", [2, 8])
def test_batch_job_list_kernel_equals_sweep(ctx, segments):
    """Batch levels after the first by the job-list kernel (queries packed in panels, finished
    queries' words masked) give the sweep kernel's per-query frontiers and batch totals, on a
    captured graph replayed twice, then again after switching kernels."""
    M = mg()
    specs, graphs, qs = small_workload(ctx, 30, 17)
    bj = M.WalkBatch(ctx, graphs, qs)
    bs = M.WalkBatch(ctx, graphs, qs)
    bj.set_hop_kernel(segments)
    bs.set_hop_kernel(0)
    for b in (bj, bs):
        b.run()
        b.run()
    ctx.synchronize()
    assert bj.totals() == bs.totals()
    for i in range(len(qs)):
        assert bj.query_digest(i) == bs.query_digest(i), i
    bj.set_hop_kernel(0)
    bs.set_hop_kernel(segments)
    for b in (bj, bs):
        b.run()
    ctx.synchronize()
    assert bj.totals() == bs.totals()
    for i in range(0, len(qs), 3):
        assert bj.query_digest(i) == bs.query_digest(i), i
