"""GPU bit-tiled product C = A * B (B supplied transposed) vs the CPU oracle, including the
reference-wrapper entry points called with raw device pointers."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def host(ns, off, row, col, data):
    return mg().HostTiledMatrix(ns, off, row, col, data)


def gpu_square(ctx, n, s, d):
    """C = A * A on the GPU from an edge list (A and A^T built by the oracle converter)."""
    M = mg()
    a = O.edges_to_tiles(n, s, d)
    at = O.edges_to_tiles(n, s, d, transpose=True)
    ga = M.TiledMatrix.upload(ctx, host(*a))
    gat = M.TiledMatrix.upload(ctx, host(*at))
    c, st = M.multiply(ctx, ga, gat)
    return a, at, c.download(), st, c


def oracle_square(a, at):
    ns = a[0]
    c_off, c_row, c_col, tp = O.tile_symbolic(ns, a[1], a[3], at[1], at[3])
    c_data = O.tile_numeric(ns, a[1:], at[1:], c_row, c_col)
    return c_off, c_row, c_col, c_data, tp


def test_toy_graph_square(ctx):
    g = json.loads((GOLDEN / "toy_graph.json").read_text())
    e = np.array(g["edges"], np.uint32)
    s, d = O.sort_unique_edges(e[:, 0], e[:, 1])
    _, _, c, st, handle = gpu_square(ctx, g["n_vertices"], s, d)
    got = {(r, cc) for r in range(64) for cc in range(64) if (int(c.data[0, r]) >> cc) & 1}
    assert got == set(map(tuple, g["powers"]["2"]))
    assert handle.popcount() == g["total_count"]["2"]
    assert st["out_tiles"] == 1 and st["tile_products"] == 1


@pytest.mark.parametrize("scale,seed", [(12, 3), (14, 1)])
def test_rmat_square_matches_oracle(ctx, scale, seed):
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    a, at, c, st, handle = gpu_square(ctx, 1 << scale, s, d)
    c_off, c_row, c_col, c_data, tp = oracle_square(a, at)
    assert np.array_equal(c.offset_row, c_off)
    assert np.array_equal(c.row_idx, c_row) and np.array_equal(c.col_idx, c_col)
    assert np.array_equal(c.data, c_data)
    assert st["tile_products"] == tp and st["out_tiles"] == len(c_row)
    assert handle.popcount() == int(np.unpackbits(c_data.view(np.uint8)).sum())


def test_dense_and_closed_form_tiles(ctx):
    # complete bipartite blocks give all-ones tiles (the 64-iteration path), plus a cycle
    n = 200
    dense = np.zeros((n, n), np.uint8)
    dense[0:64, 64:128] = 1
    dense[64:128, 128:192] = 1
    dense[np.arange(n), (np.arange(n) + 1) % n] = 1
    dense[5, 7] = dense[150, 3] = 1
    s, d = np.nonzero(dense)
    a, at, c, st, _ = gpu_square(ctx, n, s.astype(np.uint32), d.astype(np.uint32))
    expect = O.bool_matmul_bruteforce(dense, dense)
    got = host(c.n_strips, c.offset_row, c.row_idx, c.col_idx, c.data).to_dense(n)
    assert np.array_equal(got, expect)


@pytest.mark.parametrize("seed", [7, 8])
def test_tiles_around_compact_limit(ctx, seed):
    """Tiles of 1..12 bits (the numeric pass multiplies tiles of up to 10 bits from their
    compact entry lists, one lane per pair, and denser ones as whole tiles through the inner
    indices both tiles use, so partial and empty overlaps of their column sets occur), 40-bit
    and full tiles, with empty tiles between them: C = A * A equals a brute-force Boolean
    product."""
    rng = np.random.default_rng(seed)
    ns = 8
    n = ns * 64
    counts = [1, 2, 3, 9, 10, 11, 12, 40, 4096]
    dense = np.zeros((n, n), np.uint8)
    for x in range(ns):
        for y in range(ns):
            if rng.random() < 0.25:
                continue
            c = counts[int(rng.integers(len(counts)))]
            cells = rng.choice(4096, size=c, replace=False)
            dense[x * 64 + cells // 64, y * 64 + cells % 64] = 1
    s, d = np.nonzero(dense)
    a, at, c, st, _ = gpu_square(ctx, n, s.astype(np.uint32), d.astype(np.uint32))
    got = host(c.n_strips, c.offset_row, c.row_idx, c.col_idx, c.data).to_dense(n)
    assert np.array_equal(got, O.bool_matmul_bruteforce(dense, dense))
    c_off, c_row, c_col, c_data, tp = oracle_square(a, at)
    assert np.array_equal(c.data, c_data) and st["tile_products"] == tp


def _hip_runtime():
    import ctypes

    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


def test_numeric_pass_on_a_caller_stream_then_back(ctx):
    """The numeric pass queued on a caller's stream, then the context switched back to its own
    stream (the switch waits for the caller's stream, whose pass reads the context's scratch)
    and a second, smaller product run there: both equal the oracle."""
    import ctypes

    M = mg()
    hip = _hip_runtime()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        results = []
        for k, (scale, seed) in enumerate([(12, 21), (11, 22)]):
            s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
            a = O.edges_to_tiles(1 << scale, s, d)
            at = O.edges_to_tiles(1 << scale, s, d, transpose=True)
            ctx.set_stream(stream.value if k == 0 else None)
            ga = M.TiledMatrix.upload(ctx, host(*a))
            gat = M.TiledMatrix.upload(ctx, host(*at))
            c, _ = M.multiply_symbolic(ctx, ga, gat)
            M.multiply_numeric(ctx, ga, gat, c)
            results.append((a, at, c, ga, gat))
        ctx.set_stream(None)
        ctx.synchronize()
        for a, at, c, _, _ in results:
            c_off, c_row, c_col, c_data, _ = oracle_square(a, at)
            got = c.download()
            assert np.array_equal(got.row_idx, c_row) and np.array_equal(got.col_idx, c_col)
            assert np.array_equal(got.data, c_data)
    finally:
        ctx.set_stream(None)
        ctx.synchronize()
        assert hip.hipStreamDestroy(stream) == 0


def test_empty_operand(ctx):
    M = mg()
    s, d = O.rmat_edges(8, 16, 0.57, 0.19, 0.19, 1)
    at = O.edges_to_tiles(256, s, d, transpose=True)
    ga = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.empty(4))
    gat = M.TiledMatrix.upload(ctx, host(*at))
    c, st = M.multiply(ctx, ga, gat)
    assert c.n_nz_tile == 0 and st["tile_products"] == 0


# ---------------------------------------------------------------------------
# Layer 1: reference-wrapper entry points on raw device buffers
# ---------------------------------------------------------------------------
_CTX = None


def _dev(arr):
    return mg().DeviceBuffer.from_array(_CTX, arr)


def _host(buf, dtype):
    _CTX.synchronize()
    return buf.to_array(dtype)


def _stream():
    return _CTX.stream


def _bits(words, nbits):
    return np.unpackbits(np.ascontiguousarray(words).view(np.uint8), bitorder="little")[:nbits]


def _pack(bits):
    nw = len(bits) // 64 + 1
    pad = np.zeros(nw * 64, np.uint8)
    pad[:len(bits)] = bits
    return np.packbits(pad, bitorder="little").view(np.uint64).copy()


def test_layer1_bit_and_and_bit_count(ctx):
    global _CTX
    _CTX = ctx
    lib = mg().lib()
    rng = np.random.default_rng(7)
    m, k, n = 70, 100, 90
    A = (rng.random((m, k)) < 0.03).astype(np.uint8)
    B = (rng.random((n, k)) < 0.03).astype(np.uint8)
    expect = (A.astype(np.int64) @ B.T.astype(np.int64) > 0).astype(np.uint8)
    da, db = _dev(_pack(A.ravel())), _dev(_pack(B.ravel()))
    prior = np.zeros(m * n, np.uint8)
    prior[rng.integers(0, m * n, 50)] = 1
    dc = _dev(_pack(prior))
    assert lib.mg_matrix_bit_and(_stream(), da.ptr, db.ptr, dc.ptr, m, k, n) == 0
    got = _bits(_host(dc, np.uint64), m * n).reshape(m, n)
    assert np.array_equal(got, expect | prior.reshape(m, n))  # OR-accumulates like atomicOr
    words = rng.integers(0, 2**63, 10, dtype=np.int64).astype(np.uint64)
    cnt = _dev(np.array([5], np.uint64))
    dw = _dev(words)  # keep the buffer alive until the launch has completed
    assert lib.mg_matrix_bit_count(_stream(), dw.ptr, cnt.ptr, 600) == 0
    assert int(_host(cnt, np.uint64)[0]) == 5 + int(_bits(words, 600).sum())


def test_layer1_init_metadata(ctx):
    global _CTX
    _CTX = ctx
    lib = mg().lib()
    rng = np.random.default_rng(8)
    ns = 100
    lay = (rng.random((ns, ns)) < 0.05).astype(np.uint8)
    xs, ys = np.nonzero(lay)
    dl = _dev(_pack(lay.ravel()))
    cnt = _dev(np.zeros(ns, np.uint32))
    dr = _dev(np.zeros(len(xs), np.uint32))
    dcol = _dev(np.zeros(len(xs), np.uint32))
    assert lib.mg_init_bit_tiled_matrix_metadata_by_layout_matrix(
        _stream(), dl.ptr, ns, cnt.ptr, dr.ptr, dcol.ptr) == 0
    assert np.array_equal(_host(cnt, np.uint32), lay.sum(axis=1).astype(np.uint32))
    assert np.array_equal(_host(dr, np.uint32), xs.astype(np.uint32))
    assert np.array_equal(_host(dcol, np.uint32), ys.astype(np.uint32))


def test_layer1_fill_tiles_and_count(ctx):
    global _CTX
    _CTX = ctx
    lib = mg().lib()
    s, d = O.rmat_edges(12, 16, 0.57, 0.19, 0.19, 9)
    a = O.edges_to_tiles(4096, s, d)
    at = O.edges_to_tiles(4096, s, d, transpose=True)
    c_off, c_row, c_col, c_data, _ = oracle_square(a, at)
    rng = np.random.default_rng(1)
    prior = np.zeros_like(c_data)
    idx = rng.integers(0, prior.size, 200)
    prior.flat[idx] = rng.integers(0, 2**63, 200, dtype=np.int64).astype(np.uint64)
    ns = a[0]
    bufs = [_dev(v) for v in (a[1], at[1], c_off, a[2], at[2], c_row, a[3], at[3], c_col, a[4], at[4], prior)]
    (a_off, b_off, co, a_row, b_row, cr, a_col, b_col, cc, a_dat, b_dat, dc) = bufs
    rc = lib.mg_fill_tiles(_stream(), 64, ns, len(a[2]), len(at[2]), len(c_row), None, a_off.ptr,
                           b_off.ptr, co.ptr, a_row.ptr, b_row.ptr, cr.ptr,
                           a_col.ptr, b_col.ptr, cc.ptr, a_dat.ptr, b_dat.ptr,
                           dc.ptr)
    assert rc == 0
    got = _host(dc, np.uint64).reshape(-1, 64)
    assert np.array_equal(got, c_data | prior)
    cnt = _dev(np.zeros(1, np.uint64))
    assert lib.mg_count(_stream(), 64, ns, len(c_row), dc.ptr, cnt.ptr) == 0
    assert int(_host(cnt, np.uint64)[0]) == int(np.unpackbits((c_data | prior).view(np.uint8)).sum())
    assert lib.mg_fill_tiles(_stream(), 32, ns, 0, 0, 1, None, *([None] * 12)) == 4
