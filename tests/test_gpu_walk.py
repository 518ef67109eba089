"""GPU walks X_{k+1} = X_k * A (A supplied transposed) vs the CPU oracles: dense frontier
content (bit-exact), structural tile sets and tile-product counts, the reference-layout
export, pruning, edge cases and the full-size R-MAT-22 4-hop walk (digests)."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def setup_graph(ctx, scale, seed, ef=16):
    M = mg()
    el = M.EdgeList.rmat(ctx, scale, ef, 0.57, 0.19, 0.19, seed)
    at = el.to_tiled_matrix(transpose=True)
    s, d = O.rmat_edges(scale, ef, 0.57, 0.19, 0.19, seed)
    return el, at, s, d


def s_masks(n_strips, sources):
    s = np.zeros(n_strips, np.uint64)
    for i, v in enumerate(np.asarray(sources, np.int64)):
        s[v // 64] |= np.uint64(1) << np.uint64(i // 64)
    return s


def numeric_masks(x):
    """Per strip: bit w set iff some vertex of the strip has a bit in source word w."""
    ns = x.shape[0] // 64
    nz = (x.reshape(ns, 64, x.shape[1]) != 0).any(axis=1)
    return (nz.astype(np.uint64) << np.arange(x.shape[1], dtype=np.uint64)).sum(axis=1, dtype=np.uint64)


def popcount(a):
    return int(np.unpackbits(np.ascontiguousarray(a).view(np.uint8)).sum())


# 3000 sources fill 47 of a 64-word panel's words (the other 17 words stay zero); the batch
# tests cover panels whose finished queries' words are dead.
@pytest.mark.parametrize("scale,seed,n_src,hops", [(10, 1, 64, 4), (12, 4, 1, 3), (12, 5, 65, 4), (14, 2, 4096, 4),
                                                   (14, 6, 1000, 2), (13, 7, 3000, 4)])
def test_walk_matches_oracle(ctx, scale, seed, n_src, hops):
    M = mg()
    n = 1 << scale
    el, at, s, d = setup_graph(ctx, scale, seed)
    sources = O.sample_sources(n, n_src, seed)
    w = M.Walk(ctx, at, n_src, hops)
    w.run(sources, hops)
    xs = O.walk_frontiers(n, s, d, sources, hops, W=w.words_per_vertex)
    ns = n // 64
    ath = at.download()
    s_cur = s_masks(ns, sources)
    for k in range(1, hops + 1):
        got = w.frontier(k)
        assert np.array_equal(got, xs[k]), f"hop {k}"
        st = w.hop_stats(k)
        s_next, tp = O.walk_symbolic(ns, ath.offset_row, ath.col_idx, s_cur)
        assert st["tile_products"] == tp
        assert st["left_tiles"] == sum(bin(int(v)).count("1") for v in s_cur)
        assert st["out_tiles"] == sum(bin(int(v)).count("1") for v in s_next)
        assert st["right_tiles_used"] == int(np.count_nonzero(s_cur[ath.col_idx]))
        assert st["popcount"] == popcount(xs[k])
        assert st["out_tiles_numeric"] == sum(bin(int(v)).count("1") for v in numeric_masks(xs[k]))
        assert st["out_tiles_numeric"] <= st["out_tiles"]
        assert w.frontier_digest(k) == O.frontier_digest(xs[k], n_src)
        s_cur = s_next


def test_walk_export_matches_tile_oracle(ctx):
    """X_k exported in the reference layout equals the O4 tile product X_{k-1} * A."""
    M = mg()
    scale, seed, n_src = 11, 3, 130
    n = 1 << scale
    el, at, s, d = setup_graph(ctx, scale, seed)
    sources = O.sample_sources(n, n_src, 9)
    w = M.Walk(ctx, at, n_src, 3)
    w.run(sources, 3)
    a = O.edges_to_tiles(n, s, d)
    at_h = O.edges_to_tiles(n, s, d, transpose=True)
    ns = a[0]
    x_prev = w.export(0).download()
    # X_0 export: one-hot rows
    dense0 = x_prev.to_dense(n)[:n_src]
    assert np.array_equal(np.nonzero(dense0)[1], sources) and dense0.sum() == n_src
    for k in range(1, 4):
        xk = w.export(k).download()
        c_off, c_row, c_col, tp = O.tile_symbolic(ns, x_prev.offset_row, x_prev.col_idx, at_h[1], at_h[3])
        assert np.array_equal(xk.row_idx, c_row) and np.array_equal(xk.col_idx, c_col)
        assert np.array_equal(xk.offset_row, c_off)
        c_data = O.tile_numeric(ns, (x_prev.offset_row, x_prev.row_idx, x_prev.col_idx, x_prev.data), at_h[1:],
                                c_row, c_col)
        assert np.array_equal(xk.data, c_data)
        assert w.hop_stats(k)["tile_products"] == tp
        x_prev = xk


def test_walk_prune_same_content(ctx):
    M = mg()
    el, at, s, d = setup_graph(ctx, 13, 8)
    sources = O.sample_sources(1 << 13, 300, 1)
    w0 = M.Walk(ctx, at, 300, 4)
    w1 = M.Walk(ctx, at, 300, 4, flags=M.WALK_PRUNE)
    w0.run(sources, 4)
    w1.run(sources, 4)
    for k in range(1, 5):
        assert np.array_equal(w0.frontier(k), w1.frontier(k))
        s0, s1 = w0.hop_stats(k), w1.hop_stats(k)
        assert s1["out_tiles"] <= s0["out_tiles"] and s1["tile_products"] <= s0["tile_products"]
        assert s1["out_tiles_numeric"] == s0["out_tiles_numeric"]


def test_walk_cycle_closed_form_and_empty_graph(ctx):
    M = mg()
    n = 8192
    v = np.arange(n, dtype=np.uint32)
    cyc = M.EdgeList.upload(ctx, n, v, (v + 1) % n).to_tiled_matrix(transpose=True)
    sources = np.arange(0, n, 3, dtype=np.uint32)[:2000]
    w = M.Walk(ctx, cyc, 2000, 5)
    w.run(sources, 5)
    for k in range(1, 6):
        x = w.frontier(k)
        assert w.hop_stats(k)["popcount"] == 2000
        idx = (sources.astype(np.int64) + k) % n
        for i in range(0, 2000, 97):
            assert (int(x[idx[i], i // 64]) >> (i % 64)) & 1
    empty = M.EdgeList.upload(ctx, 4096, np.zeros(0, np.uint32), np.zeros(0, np.uint32)).to_tiled_matrix(True)
    we = M.Walk(ctx, empty, 10, 2)
    we.run(np.arange(10, dtype=np.uint32), 2)
    assert not we.frontier(1).any() and we.hop_stats(1)["tile_products"] == 0


def ragged_graph():
    """1000 vertices (the last strip is partial): three DAG layers 0-99 -> 100-199 -> 200-299 ->
    300-399 whose walks end after three hops, a complete 64 x 64 block on 640-703 (dense
    tiles), self-loops, and edges into the last vertex."""
    rng = np.random.default_rng(3)
    src, dst = [], []
    for layer in range(3):
        src += rng.integers(layer * 100, layer * 100 + 100, 400).tolist()
        dst += rng.integers(layer * 100 + 100, layer * 100 + 200, 400).tolist()
    blk = np.arange(640, 704)
    src += np.repeat(blk, 64).tolist()
    dst += np.tile(blk, 64).tolist()
    src += [999, 500, 640, 998, 997]
    dst += [999, 500, 999, 999, 999]
    return O.sort_unique_edges(np.array(src, np.uint32), np.array(dst, np.uint32))


def test_walk_ragged_dag_dense_block_and_last_vertex(ctx):
    """Every hop equals the oracle on the ragged graph; the DAG sources' walks are empty after
    three hops; padding rows stay zero; digests match the oracle's over the padded rows."""
    M = mg()
    n = 1000
    s, d = ragged_graph()
    at = M.EdgeList.upload(ctx, n, s, d).to_tiled_matrix(transpose=True)
    sources = np.array([0, 999, 5, 650, 320, 640, 998] + list(range(50, 90)), np.uint32)
    hops = 6
    w = M.Walk(ctx, at, len(sources), hops)
    w.run(sources, hops)
    W = w.words_per_vertex
    xs = O.walk_frontiers(n, s, d, sources, hops, W=W)
    for k in range(1, hops + 1):
        got = w.frontier(k)
        assert np.array_equal(got[:n], xs[k]), f"hop {k}"
        assert not got[n:].any()
        padded = np.zeros_like(got)
        padded[:n] = xs[k]
        assert w.frontier_digest(k) == O.frontier_digest(padded, len(sources))
        assert w.hop_stats(k)["popcount"] == popcount(xs[k])
    dag = [i for i, v in enumerate(sources.tolist()) if v < 400 and v not in (320,)]
    for k in range(4, hops + 1):
        x = xs[k]
        assert not any((int(x[v, i // 64]) >> (i % 64)) & 1 for i in dag for v in range(n))
    assert xs[hops][999, 1 // 64] >> np.uint64(1) & np.uint64(1)  # the last vertex keeps its self-loop walk


def test_walk_rerun_smaller_and_errors(ctx):
    M = mg()
    n = 1 << 12
    el, at, s, d = setup_graph(ctx, 12, 11)
    w = M.Walk(ctx, at, 4096, 3)
    big = O.sample_sources(n, 4096, 1)
    w.run(big, 3)
    small = O.sample_sources(n, 10, 2)
    w.run(small, 2)  # stale bits of the larger run must not leak
    xs = O.walk_frontiers(n, s, d, small, 2, W=w.words_per_vertex)
    assert np.array_equal(w.frontier(2), xs[2])
    with pytest.raises(M.MGError):
        w.run(np.array([1, 1], np.uint32), 1)
    with pytest.raises(M.MGError):
        w.run(small, 4)
    with pytest.raises(M.MGError):
        w.run(np.array([n], np.uint32), 1)
    with pytest.raises(M.MGError):
        w.hop_stats(3)


_STAT_KEYS = ("tile_products", "left_tiles", "right_tiles_used", "out_tiles", "out_tiles_numeric", "popcount",
              "alg_bytes", "edges_visited")


@pytest.mark.parametrize("scale,seed,sizes,flags", [(12, 3, (1, 50, 700), 0), (12, 3, (700, 1, 50), 1),
                                                    (14, 7, (100, 2000, 100), 0), (14, 5, (20, 130, 20), 1)])
def test_walk_first_hop_push_equals_sweep(ctx, scale, seed, sizes, flags):
    """The first hop pushed from the source strips' A^T columns gives the same frontiers,
    masks (seen through the later hops and re-runs) and statistics as sweeping every strip,
    over a sequence of runs whose source sets grow, shrink and repeat; the last run is also
    checked against the oracle."""
    M = mg()
    n = 1 << scale
    el, at, s, d = setup_graph(ctx, scale, seed)
    hops = 3
    wp = M.Walk(ctx, at, max(sizes), hops, flags=flags)
    ws = M.Walk(ctx, at, max(sizes), hops, flags=flags | M.WALK_SWEEP_FIRST_HOP)
    pushed = 0
    for r, size in enumerate(sizes):
        sources = O.sample_sources(n, size, 100 + r)
        wp.run(sources, hops)
        ws.run(sources, hops)
        for k in range(1, hops + 1):
            assert np.array_equal(wp.frontier(k), ws.frontier(k)), f"run {r} hop {k}"
            a, b = wp.hop_stats(k), ws.hop_stats(k)
            assert {f: a[f] for f in _STAT_KEYS} == {f: b[f] for f in _STAT_KEYS}, f"run {r} hop {k}"
            assert b["kernel"] == (M.HOP_SWEEP if k == 1 else a["kernel"])
            if k > 1:
                assert a["kernel"] in (M.HOP_SWEEP, M.HOP_JOBS)
        pushed += wp.hop_stats(1)["kernel"] == M.HOP_PUSH
    assert pushed >= 1
    xs = O.walk_frontiers(n, s, d, sources, hops, W=wp.words_per_vertex)
    for k in range(1, hops + 1):
        assert np.array_equal(wp.frontier(k), xs[k])


@pytest.mark.skipif(os.environ.get("MG_SKIP_FULL") == "1", reason="full-size check disabled by MG_SKIP_FULL=1")
def test_walk_rmat22_full_scale_digests(ctx):
    """Configuration: R-MAT scale 22 (seed 2), 4096 sources (seed 2), 4 hops; every hop's
    frontier digest equals the CPU oracle's (O3 vertex-major frontier)."""
    M = mg()
    scale, seed, n_src, hops = 22, 2, 4096, 4
    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    assert el.digest() == O.edges_digest(s, d)
    at = el.to_tiled_matrix(transpose=True)
    sources = O.sample_sources(n, n_src, seed)
    w = M.Walk(ctx, at, n_src, hops)
    w.run(sources, hops)
    off, isrc = O.in_csr(n, s, d)
    del s, d
    x = O.one_hot_frontier(n, sources, 64)
    assert w.hop_stats(1)["kernel"] == M.HOP_PUSH
    for k in range(1, hops + 1):
        x = O.frontier_step(n, off, isrc, x)
        assert w.frontier_digest(k) == O.frontier_digest(x, n_src), f"hop {k}"
        assert w.hop_stats(k)["popcount"] == popcount(x)



def _stats(w, k):
    st = w.hop_stats(k)
    return {f: st[f] for f in _STAT_KEYS}


@pytest.mark.parametrize("segments", [1, 2, 4, 8])
def test_job_list_kernel_equals_sweep_and_oracle(ctx, segments):
    """R-MAT 18 (hub strips are split across chunks), 4096 sources: the job-list hop kernel with
    1, 2, 4 or 8 row segments gives the sweep kernel's frontiers and statistics at every hop,
    over runs whose source sets shrink and grow again (stores skipped outside the previous
    masks), and the oracle's frontiers on the last run."""
    M = mg()
    scale, seed, hops = 18, 9, 4
    n = 1 << scale
    el, at, s, d = setup_graph(ctx, scale, seed)
    wj = M.Walk(ctx, at, 4096, hops)
    ws = M.Walk(ctx, at, 4096, hops)
    assert wj.set_hop_kernel(segments) == segments
    assert ws.set_hop_kernel(0) == 0
    for r, size in enumerate((4096, 700, 4096)):
        sources = O.sample_sources(n, size, 40 + r)
        wj.run(sources, hops)
        ws.run(sources, hops)
        for k in range(1, hops + 1):
            assert np.array_equal(wj.frontier(k), ws.frontier(k)), f"run {r} hop {k}"
            assert _stats(wj, k) == _stats(ws, k), f"run {r} hop {k}"
            if k > 1:
                assert wj.hop_stats(k)["kernel"] == M.HOP_JOBS and wj.hop_stats(k)["segments"] == segments
                assert ws.hop_stats(k)["kernel"] == M.HOP_SWEEP
    xs = O.walk_frontiers(n, s, d, sources, hops, W=64)
    for k in range(1, hops + 1):
        assert np.array_equal(wj.frontier(k), xs[k]), f"oracle hop {k}"


@pytest.mark.parametrize("segments", [1, 4])
def test_job_list_kernel_ragged_graph(ctx, segments):
    """The job-list kernel (one segment: whole rows, eight waves; four segments) on a
    1000-vertex graph (partial last strip, dense tiles, walks that die out, self-loops) with a
    4096-source walk: every hop equals the oracle."""
    M = mg()
    n = 1000
    s, d = ragged_graph()
    at = M.EdgeList.upload(ctx, n, s, d).to_tiled_matrix(transpose=True)
    sources = np.array([0, 999, 5, 650, 320, 640, 998] + list(range(50, 90)), np.uint32)
    hops = 6
    w = M.Walk(ctx, at, 4096, hops)
    assert w.set_hop_kernel(segments) == segments
    w.run(sources, hops)
    xs = O.walk_frontiers(n, s, d, sources, hops, W=64)
    for k in range(1, hops + 1):
        got = w.frontier(k)
        assert np.array_equal(got[:n], xs[k]), f"hop {k}"
        assert not got[n:].any()
        assert w.hop_stats(k)["popcount"] == popcount(xs[k])


def test_hop_kernel_choice_errors(ctx):
    M = mg()
    el, at, s, d = setup_graph(ctx, 10, 1)
    w = M.Walk(ctx, at, 4096, 2)
    with pytest.raises(M.MGError):
        w.set_hop_kernel(3)
    narrow = M.Walk(ctx, at, 100, 2)
    assert narrow.set_hop_kernel(4) == 0  # fewer than 2049 sources keep the sweep


@pytest.mark.parametrize("scale,seed,n_src,hops,flags", [(14, 3, 4096, 4, 0), (13, 8, 300, 3, 0),
                                                         (14, 5, 4096, 3, 2), (10, 2, 1000, 4, 0)])
def test_engine_strip_order_equals_vertex_order(ctx, scale, seed, n_src, hops, flags):
    """The engine's permuted strip order (default) and the vertex order give the same
    frontiers (downloads), digests, exports and statistics at every hop, for full and narrow
    frontiers, pushed and swept first hops; the permuted walk also equals the oracle."""
    M = mg()
    n = 1 << scale
    el, at, s, d = setup_graph(ctx, scale, seed)
    sources = O.sample_sources(n, n_src, seed)
    wp = M.Walk(ctx, at, n_src, hops, flags)
    wv = M.Walk(ctx, at, n_src, hops, flags | M.WALK_VERTEX_ORDER)
    for w in (wp, wv):
        w.run(sources, hops)
    xs = O.walk_frontiers(n, s, d, sources, hops, W=wp.words_per_vertex)
    for k in range(1, hops + 1):
        fp = wp.frontier(k)
        assert np.array_equal(fp, wv.frontier(k)), f"hop {k}"
        assert np.array_equal(fp, xs[k]), f"hop {k} vs oracle"
        assert wp.frontier_digest(k) == wv.frontier_digest(k) == O.frontier_digest(xs[k], n_src)
        assert _stats(wp, k) == _stats(wv, k), f"hop {k}"
        assert wp.hop_stats(k)["kernel"] == wv.hop_stats(k)["kernel"]
        ep, ev = wp.export(k).download(), wv.export(k).download()
        assert np.array_equal(ep.offset_row, ev.offset_row) and np.array_equal(ep.col_idx, ev.col_idx)
        assert np.array_equal(ep.data, ev.data)


def test_batch_engine_strip_order_equals_vertex_order(ctx):
    M = mg()
    from matrixgraph_amd import workloads as WL

    specs = [WL.GraphSpec("rmat", 13, 61), WL.GraphSpec("er", 12, 62), WL.GraphSpec("rmat", 12, 63)]
    graphs = [WL.build_graph(ctx, sp)[0] for sp in specs]
    qs = WL.batch_queries(specs, 40, 21, max_sources=2000, max_hops=5)
    bp = M.WalkBatch(ctx, graphs, qs)
    bv = M.WalkBatch(ctx, graphs, qs, vertex_order=True)
    for b in (bp, bv):
        b.run()
        b.run()
    ctx.synchronize()
    assert bp.totals() == bv.totals()
    for i in range(len(qs)):
        assert bp.query_digest(i) == bv.query_digest(i), i
        if i % 4 == 0:
            assert np.array_equal(bp.query_frontier(i), bv.query_frontier(i)), i
            ep, ev = bp.query_export(i).download(), bv.query_export(i).download()
            assert np.array_equal(ep.col_idx, ev.col_idx) and np.array_equal(ep.data, ev.data), i
