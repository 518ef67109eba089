"""On-disk bit-tile format (host only, no GPU): byte layout of BitTiledMatrixIO /
GridBitTiledMatrixIO (core/io/bit_tiled_matrix_io.cu:30-136,
core/io/grid_bit_tiled_matrix_io.cu:22-67), round trips and malformed-input errors."""
import numpy as np
import pytest

import oracle as O


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def rmat_block(scale=10, seed=3, transpose=False):
    s, d = O.rmat_edges(scale, 16, 0.57, 0.19, 0.19, seed)
    return mg().HostTiledMatrix(*O.edges_to_tiles(1 << scale, s, d, transpose))


def test_block_bytes_match_reference_layout(tmp_path):
    M = mg()
    b = rmat_block()
    d = tmp_path / "blk"
    M.write_block(d, b)
    meta = (d / "meta.yaml").read_text()
    assert meta == (f"BitTiledMatrixMetadata:\n  n_strips: {b.n_strips}\n  n_nz_tile: {b.n_nz_tile}\n"
                    f"  tile_size: 64\n")
    mb = d / "meta_buf"
    assert (mb / "row_idx.bin").read_bytes() == b.row_idx.astype("<u4").tobytes()
    assert (mb / "col_idx.bin").read_bytes() == b.col_idx.astype("<u4").tobytes()
    assert (mb / "tile_offset_row.bin").read_bytes() == b.offset_row.astype("<u4").tobytes()
    assert (mb / "data.bin").read_bytes() == b.data.astype("<u8").tobytes()
    bm = np.frombuffer((mb / "nz_tile_bm.bin").read_bytes(), "<u8")
    assert len(bm) == b.n_strips * b.n_strips // 64 + 1
    for t in range(0, b.n_nz_tile, 7):
        pos = int(b.row_idx[t]) * b.n_strips + int(b.col_idx[t])
        assert (int(bm[pos >> 6]) >> (pos & 63)) & 1
    assert int(np.unpackbits(bm.view(np.uint8)).sum()) == b.n_nz_tile
    assert np.array_equal(bm, b.nz_tile_bm())
    back = M.read_block(d)
    for f in ("offset_row", "row_idx", "col_idx", "data"):
        assert np.array_equal(getattr(back, f), getattr(b, f))


def test_grid_round_trip_and_empty_blocks(tmp_path):
    M = mg()
    b = rmat_block(9, 4)
    empty = M.HostTiledMatrix.empty(b.n_strips)
    root = tmp_path / "grid"
    M.write_grid(root, [b, empty, empty, rmat_block(9, 5, True)], n_vertices=1024, n_edges=12345)
    assert (root / "meta.yaml").read_text() == "GridGraphMetadata:\n  n_vertices: 1024\n  n_edges: 12345\n  n_chunks: 2\n"
    blocks, meta = M.read_grid(root)
    assert meta == {"n_chunks": 2, "n_vertices": 1024, "n_edges": 12345}
    assert np.array_equal(blocks[0].data, b.data)
    assert blocks[1].n_nz_tile == 0 and blocks[1].n_strips == b.n_strips
    assert (root / "block1" / "meta.yaml").exists()


def test_yaml_variants_accepted(tmp_path):
    M = mg()
    b = rmat_block(8, 1)
    d = tmp_path / "v"
    M.write_block(d, b)
    (d / "meta.yaml").write_text(f"---\n# written by hand\nBitTiledMatrixMetadata:\n    tile_size: 64\n"
                                 f"    n_nz_tile: '{b.n_nz_tile}'\n\n    n_strips: {b.n_strips}  # strips\n")
    assert np.array_equal(M.read_block(d).data, b.data)


@pytest.mark.parametrize("damage", ["truncate_data", "missing_key", "not_a_number", "no_dir", "bad_shape"])
def test_malformed_inputs_raise_io_error(tmp_path, damage):
    M = mg()
    b = rmat_block(8, 2)
    d = tmp_path / "bad"
    M.write_block(d, b)
    if damage == "truncate_data":
        p = d / "meta_buf" / "data.bin"
        p.write_bytes(p.read_bytes()[:-8])
    elif damage == "missing_key":
        (d / "meta.yaml").write_text("BitTiledMatrixMetadata:\n  n_strips: 4\n  tile_size: 64\n")
    elif damage == "not_a_number":
        (d / "meta.yaml").write_text("BitTiledMatrixMetadata:\n  n_strips: four\n  n_nz_tile: 1\n  tile_size: 64\n")
    elif damage == "no_dir":
        d = tmp_path / "absent"
    elif damage == "bad_shape":
        (d / "meta.yaml").write_text("BitTiledMatrixMetadata: 3\n")
    with pytest.raises(M.MGError) as e:
        M.read_block(d)
    assert e.value.code == 5
