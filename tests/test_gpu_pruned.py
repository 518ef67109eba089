"""Pruned products (mg_multiply_pruned): C = A*B holding only its numerically non-empty tiles.
Every sizing mode (exact two-pass, structural scratch, a capacity below the count that forces
the exact retry, the exact count, a larger capacity) gives the structural product of
mg_multiply_symbolic + mg_multiply_numeric with its zero tiles removed, and the oracle's tile
product the same way; the per-row features agree with the symbolic pass."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def drop_zero_tiles(h):
    """(offset_row, row_idx, col_idx, data) of a downloaded tile matrix without its zero tiles."""
    keep = np.any(h.data != 0, axis=1)
    row = h.row_idx[keep]
    off = np.zeros(h.n_strips + 1, np.uint32)
    np.add.at(off, row.astype(np.int64) + 1, 1)
    return np.cumsum(off).astype(np.uint32), row, h.col_idx[keep], h.data[keep]


def check_all_modes(ctx, a, bt):
    M = mg()
    c_full, st = M.multiply(ctx, a, bt)
    ctx.synchronize()
    want = drop_zero_tiles(c_full.download())
    n_num = len(want[1])
    caps = [M.PRUNED_EXACT, M.PRUNED_STRUCTURAL, max(1, n_num // 3), n_num, n_num + 1000]
    for cap in caps:
        if cap == 0 and n_num == 0:
            continue
        c, ps = M.multiply_pruned(ctx, a, bt, cap)
        got = c.download()
        assert np.array_equal(got.offset_row, want[0]), cap
        assert np.array_equal(got.row_idx, want[1]) and np.array_equal(got.col_idx, want[2]), cap
        assert np.array_equal(got.data, want[3]), cap
        assert ps["numeric_tiles"] == n_num and ps["structural_tiles"] == st["out_tiles"], cap
        assert ps["tile_products"] == st["tile_products"], cap
        if cap == M.PRUNED_EXACT:
            assert ps["passes"] == 2 and ps["fallback"] == 0
        elif cap == max(1, n_num // 3) and n_num > cap:
            assert ps["fallback"] == 1 and ps["passes"] == 2
        else:
            assert ps["fallback"] == 0 and ps["passes"] == 1
        c.close()
    c_full.close()
    return n_num, st


@pytest.mark.parametrize("scale,seed", [(10, 3), (12, 1)])
def test_pruned_square_matches_structural_product_and_oracle(ctx, scale, seed):
    M = mg()
    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    n_num, st = check_all_modes(ctx, a, at)
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    ah = O.edges_to_tiles(n, s, d)
    ath = O.edges_to_tiles(n, s, d, transpose=True)
    _, c_row, c_col, _ = O.tile_symbolic(ah[0], ah[1], ah[3], ath[1], ath[3])
    c_data = O.tile_numeric(ah[0], ah[1:], ath[1:], c_row, c_col)
    keep = np.any(c_data != 0, axis=1)
    c, _ = M.multiply_pruned(ctx, a, at, M.PRUNED_EXACT)
    got = c.download()
    assert np.array_equal(got.row_idx, c_row[keep]) and np.array_equal(got.col_idx, c_col[keep])
    assert np.array_equal(got.data, c_data[keep])


def test_pruned_walk_step_is_mostly_empty(ctx):
    """X_1 = X_0 * A from 2000 one-hot sources: part of the structural tiles are numerically empty
    (8104 structural, 4385 non-empty)."""
    M = mg()
    from matrixgraph_amd import workloads as WL

    scale = 14
    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 9)
    at = el.to_tiled_matrix(transpose=True)
    x0 = WL.sources_matrix(ctx, n, M.sample_sources(n, 2000, 9))
    n_num, st = check_all_modes(ctx, x0, at)
    assert 0 < n_num < st["out_tiles"]


def test_pruned_dense_blocks_and_empty_operand(ctx):
    M = mg()
    rng = np.random.default_rng(4)
    n = 512
    A = np.zeros((n, n), np.uint8)
    B = np.zeros((n, n), np.uint8)
    for i in range(8):
        for j in range(8):
            p = [0, 0.001, 0.05, 0.5][rng.integers(4)]
            A[i * 64:(i + 1) * 64, j * 64:(j + 1) * 64] = rng.random((64, 64)) < p
            q = [0, 0.001, 0.05, 0.5][rng.integers(4)]
            B[i * 64:(i + 1) * 64, j * 64:(j + 1) * 64] = rng.random((64, 64)) < q
    a = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(A))
    bt = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(B.T.copy()))
    check_all_modes(ctx, a, bt)
    z = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(np.zeros((n, n), np.uint8)))
    c, ps = M.multiply_pruned(ctx, a, z, M.PRUNED_EXACT)
    assert ps["numeric_tiles"] == 0 and ps["structural_tiles"] == 0 and c.n_nz_tile == 0


def test_row_features_follow_the_symbolic_pass(ctx):
    M = mg()
    el = M.EdgeList.rmat(ctx, 12, 16, 0.57, 0.19, 0.19, 6)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    feat, cnt, tp = M.multiply_row_features(ctx, a, at)
    c, st = M.multiply_symbolic(ctx, a, at)
    h = c.download()
    assert np.array_equal(cnt, np.diff(h.offset_row)) and tp == st["tile_products"]
    assert feat.shape == (a.n_strips, M.ROW_FEATURES) and np.isfinite(feat).all()
    assert np.allclose(feat[:, 0], np.log1p(cnt.astype(np.float64)) / 16, rtol=1e-5, atol=1e-6)
    ah = a.download()
    assert np.allclose(feat[:, 2], np.log1p(np.diff(ah.offset_row).astype(np.float64)) / 16, rtol=1e-5, atol=1e-6)
    assert np.allclose(feat[:, 7], cnt / a.n_strips, rtol=1e-5, atol=1e-7)


def test_learned_capacity_walk_equals_exact_and_structural(ctx):
    """A model trained on a few pruned product walks sizes the next walks' products; every mode
    (learned capacity with its fallback, exact, structural) gives the same final frontier, equal
    to the gather walk's export with its zero tiles removed."""
    M = mg()
    from matrixgraph_amd import learned as LN
    from matrixgraph_amd import workloads as WL

    specs = [WL.GraphSpec("rmat", 12, 71), WL.GraphSpec("er", 12, 72)]
    graphs = [WL.build_graph(ctx, sp)[0] for sp in specs]
    train_q = WL.batch_queries(specs, 12, 5, max_sources=600, max_hops=3)
    X, T, pairs = LN.collect_products(ctx, graphs, train_q)
    assert X.shape[1] == M.ROW_FEATURES and len(X) == len(T) and pairs > 0
    assert np.all((T[:, 0] >= 0) & (T[:, 0] <= 1))
    mlp = M.MLP(ctx, M.ROW_FEATURES, LN.ROW_TARGETS, seed=3)
    mlp.train(X, T, epochs=20, batch=256, lr=3e-3)
    for q in WL.batch_queries(specs, 6, 8, max_sources=900, max_hops=3):
        at = graphs[q.graph]
        d_l, st_l = LN.pruned_walk(ctx, at, q.sources, q.hops, "learned", mlp)
        d_e, st_e = LN.pruned_walk(ctx, at, q.sources, q.hops, "exact")
        d_s, st_s = LN.pruned_walk(ctx, at, q.sources, q.hops, "structural")
        assert d_l == d_e == d_s
        assert st_e["passes"] == 2 * q.hops and st_s["passes"] == q.hops
        assert st_l["passes"] == q.hops + st_l["fallbacks"]
        w = M.Walk(ctx, at, len(q.sources), q.hops)
        w.run(q.sources, q.hops)
        e = w.export(q.hops).download()
        keep = np.any(e.data != 0, axis=1)
        xw = M.TiledMatrix.upload(ctx, M.HostTiledMatrix(e.n_strips, drop_zero_tiles(e)[0], e.row_idx[keep],
                                                         e.col_idx[keep], e.data[keep]))
        assert xw.digest() == d_e
        w.close()


def test_prepared_operand_gives_the_same_products(ctx):
    """Operands whose payloads are kept (mg_tiled_matrix_prepare_operand) give the same plain,
    pruned and model-free products; a numeric pass into a prepared matrix drops its kept copy
    (prepared while its tiles were still zero, then filled: later products see the filled tiles)."""
    M = mg()
    el = M.EdgeList.rmat(ctx, 12, 16, 0.57, 0.19, 0.19, 5)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    c0, _ = M.multiply(ctx, a, at)
    p0, _ = M.multiply_pruned(ctx, a, at, M.PRUNED_STRUCTURAL)
    want_c, want_p = c0.digest(), p0.digest()
    a.prepare_operand()
    at.prepare_operand()
    at.prepare_operand()  # current already: nothing to do
    c1, _ = M.multiply(ctx, a, at)
    p1, _ = M.multiply_pruned(ctx, a, at, M.PRUNED_STRUCTURAL)
    assert c1.digest() == want_c and p1.digest() == want_p
    c2, _ = M.multiply_symbolic(ctx, a, at)
    c2.prepare_operand()  # all tiles zero at this point
    M.multiply_numeric(ctx, a, at, c2)
    ctx.synchronize()
    assert c2.digest() == want_c
    d_prep, _ = M.multiply(ctx, c2, at)
    d_ref, _ = M.multiply(ctx, c0, at)
    assert d_prep.digest() == d_ref.digest()
    for m in (c0, p0, c1, p1, c2, d_prep, d_ref):
        m.close()
