"""CSR-tiled on-disk format (host only, no GPU): CSRTiledMatrixIO blocks
(core/io/csr_tiled_matrix_io.cu:33-172) whose tiles are ImmutableCSR blobs
(core/data_structures/immutable_csr.cu:79-101) as Edgelist2CSRTiledMatrix lays them out
(core/util/format_converter.cuh:417-755), and the CSR-tiled grid root
(tools/graph_converter/converter/to_grid_csr_tiled_matrix.cuh:54-85).

The byte-level fixture below is derived by hand from those files for four edges; the
reference ships no CSR-tiled files or expected outputs."""
import numpy as np
import pytest

import oracle as O


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


# Four edges, block scope S = 8, tile extent T = 4 -> n_strips 2.
#   0->1: tile (0,0) x0 y1    0->3: tile (0,0) x0 y3    1->0: tile (0,0) x1 y0    5->6: tile (1,1) x1 y2
# Tile id 0 = (0,0): vertices x0 (global 0), x1 (global 1).
#   globalid [0,1] indegree [0,0] outdegree [2,1] in_offset [0,0,0] out_offset [0,2,3]
#   incoming [] outgoing [1,3,0] edges_globalid_by_localid (T+1 = 5) [0,1,0,3,0]  -> 20 words
# Tile id 3 = (1,1): vertex x1 (global 5).
#   globalid [5] indegree [0] outdegree [1] in_offset [0,0] out_offset [0,1]
#   incoming [] outgoing [2] edges_globalid_by_localid [0,0,6,0,0]                  -> 13 words
TOY_SRC = np.array([0, 0, 1, 5], np.uint32)
TOY_DST = np.array([1, 3, 0, 6], np.uint32)
TOY_BLOB = [0, 1, 0, 0, 2, 1, 0, 0, 0, 0, 2, 3, 1, 3, 0, 0, 1, 0, 3, 0,
            5, 0, 1, 0, 0, 0, 1, 2, 0, 0, 6, 0, 0]
TOY_META = ("CSRTiledMatrixMetadata:\n  n_strips: 2\n  n_nz_tile: 2\n  tile_size: 4\n  subgraphs:\n"
            "    - gid: 0\n      num_vertices: 2\n      num_incoming_edges: 0\n      num_outgoing_edges: 3\n"
            "      max_vid: 4\n      min_vid: 0\n"
            "    - gid: 3\n      num_vertices: 1\n      num_incoming_edges: 0\n      num_outgoing_edges: 1\n"
            "      max_vid: 4\n      min_vid: 0\n")


def edge_set(s, d):
    return set(zip(np.asarray(s).tolist(), np.asarray(d).tolist()))


def test_block_bytes_match_hand_derived_layout(tmp_path):
    M = mg()
    d = tmp_path / "blk"
    M.write_csr_tiled_block(d, TOY_SRC, TOY_DST, block_scope=8, tile_extent=4)
    assert (d / "meta.yaml").read_text() == TOY_META
    mb = d / "meta_buf"
    assert np.frombuffer((mb / "data.bin").read_bytes(), "<u4").tolist() == TOY_BLOB
    assert np.frombuffer((mb / "csr_offset.bin").read_bytes(), "<u8").tolist() == [0, 80, 132]
    assert np.frombuffer((mb / "row_idx.bin").read_bytes(), "<u4").tolist() == [0, 1]
    assert np.frombuffer((mb / "col_idx.bin").read_bytes(), "<u4").tolist() == [0, 1]
    assert np.frombuffer((mb / "tile_offset_row.bin").read_bytes(), "<u4").tolist() == [0, 1, 2]
    # n_strips^2 = 4 bits -> 4/64 + 1 = 1 word; bits 0*2+0 and 1*2+1
    assert np.frombuffer((mb / "nz_tile_bm.bin").read_bytes(), "<u8").tolist() == [0b1001]
    s, t, info = M.read_csr_tiled_block(d, with_tiles=True)
    assert info["n_strips"] == 2 and info["n_nz_tile"] == 2 and info["tile_size"] == 4 and info["n_edges"] == 4
    assert list(zip(s.tolist(), t.tolist())) == [(0, 1), (0, 3), (1, 0), (5, 6)]
    assert info["tile_row"].tolist() == [0, 0, 0, 1] and info["tile_col"].tolist() == [0, 0, 0, 1]


def test_reader_does_not_rely_on_tile_order(tmp_path):
    """With several host threads the reference converter lists tiles interleaved by thread;
    the reader follows the subgraphs list and csr_offset, so a permuted block reads the same."""
    M = mg()
    d = tmp_path / "perm"
    d.joinpath("meta_buf").mkdir(parents=True)
    blob0, blob3 = TOY_BLOB[:20], TOY_BLOB[20:]
    (d / "meta_buf" / "data.bin").write_bytes(np.array(blob3 + blob0, "<u4").tobytes())
    (d / "meta_buf" / "csr_offset.bin").write_bytes(np.array([0, 52, 132], "<u8").tobytes())
    (d / "meta_buf" / "row_idx.bin").write_bytes(np.array([1, 0], "<u4").tobytes())
    (d / "meta_buf" / "col_idx.bin").write_bytes(np.array([1, 0], "<u4").tobytes())
    (d / "meta_buf" / "tile_offset_row.bin").write_bytes(np.array([0, 1, 2], "<u4").tobytes())
    (d / "meta_buf" / "nz_tile_bm.bin").write_bytes(np.array([9], "<u8").tobytes())
    head, t0, t3 = TOY_META.split("    - ")
    (d / "meta.yaml").write_text(head + "    - " + t3 + "    - " + t0)
    s, t, info = M.read_csr_tiled_block(d, with_tiles=True)
    assert list(zip(s.tolist(), t.tolist())) == [(5, 6), (0, 1), (0, 3), (1, 0)]
    assert info["tile_row"].tolist() == [1, 0, 0, 0]


@pytest.mark.parametrize("style", ["dash_at_key_indent", "flow_items", "flow_list", "comments_quotes"])
def test_yaml_list_styles_accepted(tmp_path, style):
    M = mg()
    d = tmp_path / style
    M.write_csr_tiled_block(d, TOY_SRC, TOY_DST, 8, 4)
    items = [dict(gid=0, num_vertices=2, num_incoming_edges=0, num_outgoing_edges=3, max_vid=4, min_vid=0),
             dict(gid=3, num_vertices=1, num_incoming_edges=0, num_outgoing_edges=1, max_vid=4, min_vid=0)]
    head = "CSRTiledMatrixMetadata:\n  n_strips: 2\n  n_nz_tile: 2\n  tile_size: 4\n"
    flow = ["{" + ", ".join(f"{k}: {v}" for k, v in it.items()) + "}" for it in items]
    if style == "dash_at_key_indent":
        body = "  subgraphs:\n" + "".join(
            "  - " + "\n    ".join(f"{k}: {v}" for k, v in it.items()) + "\n" for it in items)
    elif style == "flow_items":
        body = "  subgraphs:\n" + "".join(f"    - {f}\n" for f in flow)
    elif style == "flow_list":
        body = "  subgraphs: [" + ", ".join(flow) + "]\n"
    else:
        body = "  subgraphs:   # tiles\n" + "".join(
            "    -\n" + "".join(f"      {k}: '{v}'\n" for k, v in it.items()) for it in items)
        head = "---\n# by hand\n" + head.replace("n_strips: 2", "n_strips: \"2\"")
    (d / "meta.yaml").write_text(head + body)
    s, t, _ = M.read_csr_tiled_block(d)
    assert edge_set(s, t) == edge_set(TOY_SRC, TOY_DST)


def test_empty_block_and_extent_rule(tmp_path):
    M = mg()
    d = tmp_path / "empty"
    M.write_csr_tiled_block(d, np.empty(0, np.uint32), np.empty(0, np.uint32), 1024, 16)
    assert (d / "meta.yaml").read_text() == (
        "CSRTiledMatrixMetadata:\n  n_strips: 64\n  n_nz_tile: 0\n  tile_size: 16\n  subgraphs: []\n")
    assert not (d / "meta_buf" / "data.bin").exists()
    s, t, info = M.read_csr_tiled_block(d)
    assert len(s) == 0 and info["n_nz_tile"] == 0
    # format_converter.cuh:424-426: an argument below 128 counts strips
    assert M.csr_tile_extent(16384, 64) == 256
    assert M.csr_tile_extent(1000, 64) == 16
    assert M.csr_tile_extent(16384, 128) == 128
    assert M.csr_tile_extent(100, 1) == 100


@pytest.mark.parametrize("n_chunks,tile_size", [(1, 64), (2, 64), (3, 256), (2, 5)])
def test_grid_round_trip_rmat(tmp_path, n_chunks, tile_size):
    M = mg()
    n = 1 << 10
    s, d = O.rmat_edges(10, 8, 0.57, 0.19, 0.19, 7)
    root = tmp_path / "grid"
    M.write_csr_tiled_grid(root, n, s, d, n_chunks=n_chunks, tile_size=tile_size)
    assert (root / "meta.yaml").read_text() == (
        f"GridGraphMetadata:\n  n_vertices: {n}\n  n_edges: {len(s)}\n  n_chunks: {n_chunks}\n"
        f"  max_vid: {int(max(s.max(), d.max()))}\n")
    assert len((root / "label.bin").read_bytes()) == 4 * n
    rs, rd, meta = M.read_csr_tiled_grid(root)
    assert meta["n_chunks"] == n_chunks and meta["n_edges"] == len(s)
    assert len(rs) == len(s) and edge_set(rs, rd) == edge_set(s, d)
    # every edge's tile agrees with the placement rule
    scope = -(-(meta["max_vid"] + 1) // n_chunks)
    ext = M.csr_tile_extent(scope, tile_size)
    for g in range(n_chunks * n_chunks):
        bs, bd, info = M.read_csr_tiled_block(root / f"block{g}", with_tiles=True)
        assert np.all(bs // scope == g // n_chunks) and np.all(bd // scope == g % n_chunks)
        assert np.array_equal(info["tile_row"], (bs % scope) // ext)
        assert np.array_equal(info["tile_col"], (bd % scope) // ext)


@pytest.mark.parametrize("damage", ["truncate_data", "count_mismatch", "target_beyond_max_vid", "offsets_decrease",
                                    "unknown_key", "no_label"])
def test_malformed_inputs_raise_io_error(tmp_path, damage):
    M = mg()
    d = tmp_path / "bad"
    M.write_csr_tiled_block(d, TOY_SRC, TOY_DST, 8, 4)
    mb = d / "meta_buf"
    if damage == "truncate_data":
        (mb / "data.bin").write_bytes((mb / "data.bin").read_bytes()[:-4])
    elif damage == "count_mismatch":
        (d / "meta.yaml").write_text(TOY_META.replace("n_nz_tile: 2", "n_nz_tile: 3"))
    elif damage == "target_beyond_max_vid":
        blob = list(TOY_BLOB)
        blob[12] = 9  # outgoing[0] of tile 0
        (mb / "data.bin").write_bytes(np.array(blob, "<u4").tobytes())
    elif damage == "offsets_decrease":
        (mb / "csr_offset.bin").write_bytes(np.array([0, 80, 60], "<u8").tobytes())
    elif damage == "unknown_key":
        (d / "meta.yaml").write_text(TOY_META.replace("min_vid: 0\n", "min_vertex: 0\n", 1))
    if damage == "no_label":
        root = tmp_path / "grid"
        M.write_csr_tiled_grid(root, 8, TOY_SRC, TOY_DST)
        (root / "label.bin").unlink()
        with pytest.raises(M.MGError) as e:
            M.read_csr_tiled_grid(root)
    else:
        with pytest.raises(M.MGError) as e:
            M.read_csr_tiled_block(d)
    assert e.value.code == 5


def test_writer_rejects_edges_from_two_grid_blocks(tmp_path):
    M = mg()
    with pytest.raises(M.MGError):
        M.write_csr_tiled_block(tmp_path / "x", np.array([1, 9], np.uint32), np.array([2, 2], np.uint32), 8, 4)
