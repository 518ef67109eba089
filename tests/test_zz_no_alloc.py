"""The allocation-free numeric pass (mg_multiply_numeric_noalloc, static per-device payload
scratch) against mg_multiply_numeric through the C-ABI: two calls in a row, the second on
smaller operands (the scratch keeps the first call's payloads), then the size limit."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def product_pair(ctx, scale, seed):
    import matrixgraph_amd as M

    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    c_ref, st = M.multiply_symbolic(ctx, a, at)
    c_new, _ = M.multiply_symbolic(ctx, a, at)
    return el, a, at, c_ref, c_new, st


def test_noalloc_numeric_equals_numeric_twice_then_smaller(ctx):
    import matrixgraph_amd as M

    lib = M.lib()
    for scale, seed in ((14, 1), (12, 7)):
        print(f"noalloc product R-MAT {scale} seed {seed}", flush=True)
        el, a, at, c_ref, c_new, st = product_pair(ctx, scale, seed)
        assert a.n_nz_tile + at.n_nz_tile <= M.NOALLOC_MAX_TILES
        assert lib.mg_multiply_numeric(ctx.h, a.h, at.h, c_ref.h) == 0, M.lib().mg_last_error()
        assert lib.mg_multiply_numeric_noalloc(ctx.h, a.h, at.h, c_new.h) == 0, M.lib().mg_last_error()
        ctx.synchronize()
        ref, got = c_ref.download(), c_new.download()
        assert np.array_equal(got.row_idx, ref.row_idx) and np.array_equal(got.col_idx, ref.col_idx)
        assert np.array_equal(got.data, ref.data), f"R-MAT {scale}: noalloc C differs from mg_multiply_numeric"
        assert int(np.unpackbits(got.data.view(np.uint8)).sum()) > 0
        if scale == 12:  # the smaller call is also checked against the oracle's tile product
            s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
            n = 1 << scale
            ah = O.edges_to_tiles(n, s, d)
            ath = O.edges_to_tiles(n, s, d, transpose=True)
            _, c_row, c_col, tp = O.tile_symbolic(ah[0], ah[1], ah[3], ath[1], ath[3])
            assert np.array_equal(got.col_idx, c_col) and st["tile_products"] == tp
            assert np.array_equal(got.data, O.tile_numeric(ah[0], ah[1:], ath[1:], c_row, c_col))
        for m in (c_ref, c_new, a, at, el):
            m.close()


def test_noalloc_numeric_rejects_operands_beyond_the_scratch(ctx):
    import matrixgraph_amd as M

    el = M.EdgeList.rmat(ctx, 18, 16, 0.57, 0.19, 0.19, 3)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    assert a.n_nz_tile + at.n_nz_tile > M.NOALLOC_MAX_TILES
    c, _ = M.multiply_symbolic(ctx, a, at)
    assert M.lib().mg_multiply_numeric_noalloc(ctx.h, a.h, at.h, c.h) == 1
    assert b"static" in M.lib().mg_last_error()
    for m in (c, a, at, el):
        m.close()
