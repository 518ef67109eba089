"""The C++ drop-in path end to end: grids on disk in the reference format ->
ppr_query_exec (PPRQuery: symbolic -> numeric -> count) -> result grid on disk, compared
with the oracle's Boolean A * A, for one and for two chunks per dimension."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "matrixgraph_amd" / "ppr_query_exec"


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def grid_blocks(n, src, dst, n_chunks):
    """Grid-cut partition with block scope S = ceil(n / n_chunks) (tools/graph_partitioner/
    partitioner/grid_cut.cu:88-89,117-118): block (x, y) holds edges with src in block x and
    dst in block y, in local ids (u mod S, v mod S)."""
    S = -(-n // n_chunks)
    blocks = []
    for x in range(n_chunks):
        for y in range(n_chunks):
            m = (src // S == x) & (dst // S == y)
            blocks.append(mg().HostTiledMatrix(*O.edges_to_tiles(S, src[m] % S, dst[m] % S)))
    return blocks


def run_exec(a_dir, at_dir, c_dir):
    r = subprocess.run([str(EXE), "-i", f"{a_dir}/", "-it", f"{at_dir}/", "-o", f"{c_dir}/"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return int(re.search(r"total_count:(\d+)", r.stdout).group(1)), int(re.search(r"tile_products: (\d+)",
                                                                               r.stdout).group(1))


@pytest.mark.parametrize("n_chunks", [1, 2])
def test_ppr_query_exec_matches_oracle(tmp_path, n_chunks):
    M = mg()
    scale = 11
    n = 1 << scale
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, 21)
    M.write_grid(tmp_path / "A", grid_blocks(n, s, d, n_chunks), n, len(s))
    M.write_grid(tmp_path / "AT", grid_blocks(n, d, s, n_chunks), n, len(s))
    total, tp = run_exec(tmp_path / "A", tmp_path / "AT", tmp_path / "C")
    a = np.zeros((n, n), np.uint8)
    a[s, d] = 1
    expect = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(a), n, n, n), n)
    assert total == int(expect.sum())
    blocks, meta = M.read_grid(tmp_path / "C")
    assert meta["n_chunks"] == n_chunks
    S = n // n_chunks
    for x in range(n_chunks):
        for y in range(n_chunks):
            got = blocks[x * n_chunks + y].to_dense(S)
            assert np.array_equal(got, expect[x * S:(x + 1) * S, y * S:(y + 1) * S]), (x, y)
    if n_chunks == 1:
        ah = O.edges_to_tiles(n, s, d)
        ath = O.edges_to_tiles(n, s, d, transpose=True)
        c_off, c_row, c_col, otp = O.tile_symbolic(ah[0], ah[1], ah[3], ath[1], ath[3])
        assert tp == otp
        assert np.array_equal(blocks[0].row_idx, c_row) and np.array_equal(blocks[0].col_idx, c_col)


def test_ppr_query_exec_reports_bad_input(tmp_path):
    r = subprocess.run([str(EXE), "-i", f"{tmp_path}/missing/", "-it", f"{tmp_path}/missing/"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "meta.yaml" in r.stderr


@pytest.mark.parametrize("n_chunks", [2, 3])
def test_device_grid_cut_matches_host_and_runs(tmp_path, ctx, n_chunks):
    """The device grid cut (mg_edges_grid_block + conversion) writes the same A and A^T grids
    as the host partition, including a block scope that does not divide the vertex count
    (n_chunks = 3), and PPRQuery over them gives the oracle's A * A."""
    M = mg()
    scale = 11
    n = 1 << scale
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, 22)
    el = M.EdgeList.rmat(ctx, scale, 16, 0.57, 0.19, 0.19, 22)
    M.write_grid_from_edges(tmp_path / "A", el, n_chunks)
    M.write_grid_from_edges(tmp_path / "AT", el, n_chunks, transpose=True)
    for name, blocks in (("A", grid_blocks(n, s, d, n_chunks)), ("AT", grid_blocks(n, d, s, n_chunks))):
        got, meta = M.read_grid(tmp_path / name)
        assert meta == {"n_chunks": n_chunks, "n_vertices": n, "n_edges": len(s)}
        for g, (a, b) in enumerate(zip(got, blocks)):
            for f in ("offset_row", "row_idx", "col_idx", "data"):
                assert np.array_equal(getattr(a, f), getattr(b, f)), (name, g, f)
    total, _ = run_exec(tmp_path / "A", tmp_path / "AT", tmp_path / "C")
    a = np.zeros((n, n), np.uint8)
    a[s, d] = 1
    expect = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(a), n, n, n), n)
    assert total == int(expect.sum())
    blocks, meta = M.read_grid(tmp_path / "C")
    S = -(-n // n_chunks)
    for x in range(n_chunks):
        for y in range(n_chunks):
            got = blocks[x * n_chunks + y].to_dense(S)
            want = expect[x * S:(x + 1) * S, y * S:(y + 1) * S]
            assert np.array_equal(got[:want.shape[0], :want.shape[1]], want), (x, y)
            assert got.sum() == want.sum(), (x, y)
    with pytest.raises(M.MGError):
        el.grid_block(n_chunks, n_chunks, 0)
