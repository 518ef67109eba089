"""GPU input production vs the CPU oracle: generators, set operations, converter (bit-exact)."""
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


@pytest.mark.parametrize("scale,seed", [(10, 1), (14, 1), (16, 7)])
def test_rmat_matches_oracle(ctx, scale, seed):
    el = mg().EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    s, d = el.download()
    os_, od = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    assert el.digest() == O.edges_digest(os_, od)


def test_er_and_communities_match_oracle(ctx):
    el = mg().EdgeList.er(ctx, 12, 1, 6, 5)
    s, d = el.download()
    os_, od = O.er_edges(12, 1, 6, 5)
    assert np.array_equal(s, os_) and np.array_equal(d, od)
    cm = mg().EdgeList.communities(ctx, 16384, 4, 512, 4096, 1288490189, 32, 3)
    cs, cd = cm.download()
    ocs, ocd = O.community_edges(16384, 4, 512, 4096, 1288490189, 32, 3)
    assert np.array_equal(cs, ocs) and np.array_equal(cd, ocd)
    rm = mg().EdgeList.rmat(ctx, 14, 16, 0.57, 0.19, 0.19, 3)
    un = rm.union(cm)
    us, ud = un.download()
    rs, rd = O.rmat_edges(14, 16, 0.57, 0.19, 0.19, 3)
    es, ed = O.sort_unique_edges(np.concatenate([rs, ocs]), np.concatenate([rd, ocd]))
    assert np.array_equal(us, es) and np.array_equal(ud, ed)


def test_upload_keeps_self_loops_and_dedups(ctx):
    g = json.loads((GOLDEN / "toy_graph.json").read_text())
    e = np.array(g["edges"] + g["edges"][:5], np.uint32)
    el = mg().EdgeList.upload(ctx, g["n_vertices"], e[:, 0], e[:, 1])
    s, d = el.download()
    es, ed = O.sort_unique_edges(e[:, 0], e[:, 1])
    assert len(s) == 25 and np.array_equal(s, es) and np.array_equal(d, ed)
    assert np.any(s == d)


@pytest.mark.parametrize("scale", [10, 14])
def test_tiles_match_oracle(ctx, scale):
    el = mg().EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 1)
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, 1)
    n = 1 << scale
    for transpose in (False, True):
        m = el.to_tiled_matrix(transpose).download()
        ns, off, row, col, data = O.edges_to_tiles(n, s, d, transpose)
        assert m.n_strips == ns
        assert np.array_equal(m.offset_row, off) and np.array_equal(m.row_idx, row)
        assert np.array_equal(m.col_idx, col) and np.array_equal(m.data, data)


def test_tile_upload_roundtrip_and_validation(ctx):
    M = mg()
    s, d = O.rmat_edges(10, 16, 0.57, 0.19, 0.19, 4)
    ns, off, row, col, data = O.edges_to_tiles(1024, s, d)
    h = M.HostTiledMatrix(ns, off, row, col, data)
    t = M.TiledMatrix.upload(ctx, h)
    back = t.download()
    assert np.array_equal(back.data, data) and np.array_equal(back.offset_row, off)
    assert t.digest() == O.tiles_digest(ns, row, col, data)
    assert t.popcount() == len(s)
    bad = M.HostTiledMatrix(ns, off.copy(), row, col.copy(), data)
    bad.col_idx[0] = ns  # out of range
    with pytest.raises(M.MGError) as e:
        M.TiledMatrix.upload(ctx, bad)
    assert e.value.code == 1
    bad2 = M.HostTiledMatrix(ns, off, row, col, data, tile_size=32)
    with pytest.raises(M.MGError) as e:
        M.TiledMatrix.upload(ctx, bad2)
    assert e.value.code == 4
    # empty matrix
    empty = M.TiledMatrix.upload(ctx, M.HostTiledMatrix.empty(4))
    assert empty.n_nz_tile == 0 and empty.popcount() == 0


def test_edge_lists_beyond_2_28_entries(ctx):
    """40 communities of 4096 vertices at p = 1/2: about 335 M edges in one list, past the
    2^28 threads of a 2^20-block grid (element-wise kernels must stride over the grid)."""
    M = mg()
    n, n_comm, size = 40 * 4096, 40, 4096
    el = M.EdgeList.communities(ctx, n, n_comm, size, size, 1, 1, 5)
    _, m = el.info()
    assert m > (1 << 28)
    s, d = O.community_edges(n, n_comm, size, size, 1, 1, 5)
    assert m == len(s)
    assert el.digest() == O.edges_digest(s, d)
    t = el.to_tiled_matrix()
    assert t.popcount() == m


@pytest.mark.parametrize("scale,seed", [(12, 3), (14, 9)])
def test_tiles_to_edges_roundtrip(ctx, scale, seed):
    """Tile -> edge-list expansion: A's tiles give back the edge list, A^T's its transpose."""
    M = mg()
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, seed)
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    back = el.to_tiled_matrix().to_edges()
    assert back.info() == (1 << scale, len(s))
    assert back.digest() == el.digest() == O.edges_digest(s, d)
    bs, bd = el.to_tiled_matrix(transpose=True).to_edges().download()
    ts, td = O.sort_unique_edges(d, s)
    assert np.array_equal(bs, ts) and np.array_equal(bd, td)


def test_walk_export_to_edges_matches_oracle(ctx):
    """X_k exported and expanded: edge (s, v) iff source s reaches v in exactly k hops."""
    M = mg()
    scale, n_src, hops = 12, 150, 2
    n = 1 << scale
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 4)
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, 4)
    sources = O.sample_sources(n, n_src, 8)
    w = M.Walk(ctx, el.to_tiled_matrix(transpose=True), n_src, hops)
    w.run(sources, hops)
    xs = O.walk_frontiers(n, s, d, sources, hops, W=w.words_per_vertex)
    for k in range(1, hops + 1):
        es, ev = w.export(k).to_edges().download()
        bits = np.unpackbits(xs[k].view(np.uint8), bitorder="little").reshape(n, -1)
        v_idx, s_idx = np.nonzero(bits)
        keep = s_idx < n_src
        order = np.lexsort((v_idx[keep], s_idx[keep]))
        assert np.array_equal(es, s_idx[keep][order].astype(np.uint32))
        assert np.array_equal(ev, v_idx[keep][order].astype(np.uint32))


def _two_hop_rules(n, src, dst):
    """The CSR-tiled walk's pair rules (core/task/gpu_task/kernel/matrix_operations.cu:536-
    600): from s over an edge to m != s, then over an edge to v with v != m and v != s."""
    out = [set() for _ in range(n)]
    for u, v in zip(src.tolist(), dst.tolist()):
        out[u].add(v)
    pairs = set()
    for s_ in range(n):
        for m in out[s_]:
            if m == s_:
                continue
            for v in out[m]:
                if v != m and v != s_:
                    pairs.add((s_, v))
    return sorted(pairs)


def test_two_hop_pairs_match_rules_and_oracle(ctx):
    """Two-hop enumeration on the bit tiles equals the walk rules on the toy graph (with a
    self-loop added) and the oracle's A'*A' without its diagonal on an R-MAT graph with
    self-loops."""
    M = mg()
    toy = json.loads((Path(__file__).parent / "golden" / "toy_graph.json").read_text())
    n = toy["n_vertices"]
    e = np.array(toy["edges"], np.uint32).reshape(-1, 2)
    s = np.concatenate([e[:, 0], np.array([1, 2], np.uint32)])
    d = np.concatenate([e[:, 1], np.array([1, 2], np.uint32)])
    got_s, got_d = M.EdgeList.upload(ctx, n, s, d).two_hop().download()
    assert list(zip(got_s.tolist(), got_d.tolist())) == _two_hop_rules(n, s, d)

    scale = 10
    n = 1 << scale
    rs, rd = O.rmat_edges(scale, 8, 0.57, 0.19, 0.19, 7)
    loops = np.arange(0, n, 7, dtype=np.uint32)
    s = np.concatenate([rs, loops])
    d = np.concatenate([rd, loops])
    pairs = M.EdgeList.upload(ctx, n, s, d).two_hop()
    gs, gd = pairs.download()
    a = np.zeros((n, n), np.uint8)
    a[rs, rd] = 1  # A' (the R-MAT edges have no self-loops)
    c = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(a), n, n, n), n)
    np.fill_diagonal(c, 0)
    ws, wd = np.nonzero(c)
    assert pairs.info() == (n, len(ws))
    assert np.array_equal(gs, ws.astype(np.uint32)) and np.array_equal(gd, wd.astype(np.uint32))
    assert len(ws) == len(_two_hop_rules(n, s, d))


@pytest.mark.parametrize("n_chunks,tile_size", [(1, 64), (2, 64), (2, 256)])
def test_two_hop_from_csr_tiled_grid(ctx, tmp_path, n_chunks, tile_size):
    """A CSR-tiled grid (the enumerating walk's input format) read back to edges and walked
    gives the oracle's two-hop pairs of the graph it was written from."""
    M = mg()
    scale = 10
    n = 1 << scale
    s, d = O.rmat_edges(scale, 8, 0.57, 0.19, 0.19, 11)
    root = tmp_path / "csr_grid"
    M.write_csr_tiled_grid(root, n, s, d, n_chunks=n_chunks, tile_size=tile_size)
    rs, rd, meta = M.read_csr_tiled_grid(root)
    gs, gd = M.EdgeList.upload(ctx, max(meta["n_vertices"], meta["max_vid"] + 1), rs, rd).two_hop().download()
    a = np.zeros((n, n), np.uint8)
    a[s, d] = 1
    c = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(a), n, n, n), n)
    np.fill_diagonal(c, 0)
    ws, wd = np.nonzero(c)
    assert np.array_equal(gs, ws.astype(np.uint32)) and np.array_equal(gd, wd.astype(np.uint32))


def test_tile_to_edge_expansion_rejects_more_than_2_31_edges(ctx):
    """524,288 full tiles hold 2^31 set bits: the expansion's 32-bit scan would wrap, so the
    64-bit total is checked first and the call fails with an error instead of writing past
    its allocation."""
    M = mg()
    ns = 725  # 725^2 = 525,625 tile positions
    n = 1 << 19
    rows = (np.arange(n) // ns).astype(np.uint32)
    cols = (np.arange(n) % ns).astype(np.uint32)
    off = np.searchsorted(rows, np.arange(ns + 1)).astype(np.uint32)
    data = np.full((n, 64), np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64)
    m = M.TiledMatrix.upload(ctx, M.HostTiledMatrix(ns, off, rows, cols, data))
    del data
    with pytest.raises(M.MGError) as e:
        m.to_edges()
    assert "2^31" in str(e.value)
    m.close()


def test_symbolic_pass_rejects_more_than_u32_output_tiles(ctx):
    """A with tile (x, 0) in every row and Bt with tile (y, 0) in every row: C = A*B has every
    one of 65,537^2 > 2^32 - 1 tile positions, more than a u32 tile index holds; the symbolic
    pass (and the pruned product built on it) fail with an error."""
    M = mg()
    ns = 65537
    idx = np.arange(ns, dtype=np.uint32)
    off = np.arange(ns + 1, dtype=np.uint32)
    data = np.zeros((ns, 64), np.uint64)
    data[:, 0] = 1
    hm = M.HostTiledMatrix(ns, off, idx, np.zeros(ns, np.uint32), data)
    a = M.TiledMatrix.upload(ctx, hm)
    bt = M.TiledMatrix.upload(ctx, hm)
    with pytest.raises(M.MGError) as e:
        M.multiply_symbolic(ctx, a, bt)
    assert "u32" in str(e.value)
    with pytest.raises(M.MGError):
        M.multiply_pruned(ctx, a, bt, M.PRUNED_EXACT)
    a.close()
    bt.close()
