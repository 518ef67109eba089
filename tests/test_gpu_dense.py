"""Matrix-core (int8 MFMA) tile products for dense tile pairs: bit-exact against the vector
path and the dense oracle across densities, and the density-mixed walk computed by tile
products (per-pair engine choice) equal to the gather walk."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def block_matrix(n, rng, densities):
    """n x n 0/1 matrix whose 64 x 64 blocks have densities drawn from `densities`."""
    ns = n // 64
    out = np.zeros((n, n), np.uint8)
    for i in range(ns):
        for j in range(ns):
            p = densities[rng.integers(len(densities))]
            if p > 0:
                out[i * 64:(i + 1) * 64, j * 64:(j + 1) * 64] = rng.random((64, 64)) < p
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_mfma_products_match_vector_path_and_oracle(ctx, seed):
    M = mg()
    rng = np.random.default_rng(seed)
    n = 512
    dens = [0, 0.002, 0.05, 0.3, 0.5, 0.9, 1.0]
    A = block_matrix(n, rng, dens)
    B = block_matrix(n, rng, dens)
    ref = (A.astype(np.int32) @ B.astype(np.int32)) > 0
    a = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(A))
    bt = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(B.T.copy()))
    outs = {}
    for min_bits in (0, 700, 4097):
        c, st = M.multiply_symbolic(ctx, a, bt)
        dense = M.multiply_numeric_dense(ctx, a, bt, c, min_bits)
        if min_bits == 0:
            assert dense == st["tile_products"]
        if min_bits > 4096:
            assert dense == 0
        h = c.download()
        assert np.array_equal(h.to_dense(n).astype(bool), ref), min_bits
        outs[min_bits] = h
    for k in (0, 700):
        assert np.array_equal(outs[k].data, outs[4097].data)


def test_mfma_single_bit_layout(ctx):
    """Every (row, k, column) position through one tile pair: L = e_r e_k^T, R = e_k e_c^T
    gives exactly bit (r, c) -- fixes the operand and result lane layout of the instruction."""
    M = mg()
    rng = np.random.default_rng(5)
    n = 64 * 8
    for _ in range(3):
        A = np.zeros((n, n), np.uint8)
        B = np.zeros((n, n), np.uint8)
        trip = []
        for blk in range(8):  # diagonal blocks, independent products
            r, k, c = rng.integers(0, 64, 3)
            A[blk * 64 + r, blk * 64 + k] = 1
            B[blk * 64 + k, blk * 64 + c] = 1
            trip.append((blk, r, c))
        a = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(A))
        bt = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.from_dense(B.T.copy()))
        c_m, _ = M.multiply_symbolic(ctx, a, bt)
        M.multiply_numeric_dense(ctx, a, bt, c_m, 0)
        dense = c_m.download().to_dense(n)
        want = np.zeros((n, n), np.uint8)
        for blk, r, c in trip:
            want[blk * 64 + r, blk * 64 + c] = 1
        assert np.array_equal(dense, want)


def test_union_graph_matches_oracle(ctx):
    M = mg()
    from matrixgraph_amd import workloads as WL

    el = WL.union_graph(ctx, 11, 3, 4, 256, 512)
    s, d = el.download()
    so, do = WL.union_graph_oracle(11, 3, 4, 256, 512)
    assert np.array_equal(s, so) and np.array_equal(d, do)


@pytest.mark.parametrize("min_bits", [0, 512, 4097])
def test_product_walk_equals_gather_walk(ctx, min_bits):
    M = mg()
    from matrixgraph_amd import workloads as WL

    scale, n_src, hops = 12, 200, 3
    n = 1 << scale
    el = WL.union_graph(ctx, scale, 3, 4, 512, 1024)
    at = el.to_tiled_matrix(transpose=True)
    el.close()
    sources = M.sample_sources(n, n_src, 3)
    xs, stats = WL.product_walk(ctx, at, n, sources, hops, min_bits)
    w = M.Walk(ctx, at, n_src, hops)
    w.run(sources, hops)
    for k in range(1, hops + 1):
        e = w.export(k).download()
        p = xs[k].download()
        assert np.array_equal(e.row_idx, p.row_idx) and np.array_equal(e.col_idx, p.col_idx), k
        assert np.array_equal(e.data, p.data), k
        assert stats[k - 1]["tile_products"] == w.hop_stats(k)["tile_products"]
    if min_bits == 0:
        assert all(s["dense_products"] == s["tile_products"] for s in stats)
    if min_bits == 512:
        assert sum(s["dense_products"] for s in stats) > 0


@pytest.mark.parametrize("min_bits", [0, 64, 4097])
def test_community_square_on_matrix_cores_matches_oracle(ctx, min_bits):
    """A*A of planted dense communities (8 of 512 vertices at p = 0.3 on 2^13 vertices) plus a
    sparse R-MAT part: every C tile of a community sums 8 dense pairs on the matrix cores;
    the result equals the vector path and the oracle's bitset product."""
    M = mg()
    from matrixgraph_amd import workloads as WL

    scale, n_comm, size, stride = 13, 8, 512, 1024
    n = 1 << scale
    el = WL.union_graph(ctx, scale, 5, n_comm, size, stride)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    s, d = WL.union_graph_oracle(scale, 5, n_comm, size, stride)
    rows = np.zeros((n, n // 64), np.uint64)
    np.bitwise_or.at(rows, (s, d // 64), np.left_shift(np.uint64(1), (d % 64).astype(np.uint64)))
    want = O.bitset_matmul(rows, rows, n, n, n, threads=O.max_threads())
    c, st = M.multiply_symbolic(ctx, a, at)
    dense = M.multiply_numeric_dense(ctx, a, at, c, min_bits)
    got = c.download()
    dense_rows = np.zeros((n, n // 64), np.uint64)
    tiles = got.data.reshape(-1, 64)
    for t in range(len(got.row_idx)):
        x, y = int(got.row_idx[t]), int(got.col_idx[t])
        dense_rows[x * 64:(x + 1) * 64, y] = tiles[t]
    assert np.array_equal(dense_rows, want)
    if min_bits == 0:
        assert dense == st["tile_products"]
    elif min_bits == 64:
        assert 0 < dense < st["tile_products"]
    else:
        assert dense == 0
