"""Stream ordering: a walk created over a product C right after the asynchronous numeric pass
that fills C (no synchronisation in between) reads the finished C. Checked against the CPU
oracle: the dense bitset product for C, then the O3 vertex-major frontier walk."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_walk_created_right_after_async_numeric_pass(ctx):
    import matrixgraph_amd as M

    scale, seed, n_src, hops = 16, 21, 256, 2
    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    a = el.to_tiled_matrix(transpose=False)
    at = el.to_tiled_matrix(transpose=True)
    ctx.synchronize()
    c, _ = M.multiply_symbolic(ctx, a, at)
    M.multiply_numeric(ctx, a, at, c)  # asynchronous on the context's stream
    w = M.Walk(ctx, c, n_src, hops)  # its setup reads C's tiles
    sources = O.sample_sources(n, n_src, seed)
    w.run(sources, hops)
    # Oracle: C = A * A as dense bit rows; the walk's right operand is C in the A^T role, so
    # the in-neighbours of v in the walked graph are the set bits of row v of C.
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    rows = np.zeros((n, n // 64), np.uint64)
    np.bitwise_or.at(rows, (s, d // 64), np.left_shift(np.uint64(1), (d % 64).astype(np.uint64)))
    cd = O.bitset_matmul(rows, rows, n, n, n, threads=O.max_threads())
    del rows
    counts = np.zeros(n, np.uint64)
    parts = []
    step = 4096
    for r0 in range(0, n, step):
        bits = np.unpackbits(cd[r0:r0 + step].view(np.uint8), axis=1, bitorder="little")
        rr, cc = np.nonzero(bits)
        counts[r0:r0 + step] = np.bincount(rr, minlength=min(step, n - r0)).astype(np.uint64)
        parts.append(cc.astype(np.uint32))
    del cd
    in_off = np.zeros(n + 1, np.uint64)
    np.cumsum(counts, out=in_off[1:])
    in_src = np.concatenate(parts)
    x = O.one_hot_frontier(n, sources, w.words_per_vertex)
    for k in range(1, hops + 1):
        x = O.frontier_step(n, in_off, in_src, x)
        assert np.array_equal(w.frontier(k), x), f"hop {k}"
