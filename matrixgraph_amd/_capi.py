"""ctypes declarations of the C-ABI in include/matrixgraph_amd.h.

The shared library is built in-tree (matrixgraph_amd/libmatrixgraph_amd.so). Loading
fails loudly when it is missing: there is no Python or CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("MATRIXGRAPH_AMD_LIB", _HERE / "libmatrixgraph_amd.so"))

u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int, C.c_double
vp = C.c_void_p
P_u32 = C.POINTER(C.c_uint32)
P_u64 = C.POINTER(C.c_uint64)
PP_u32 = C.POINTER(P_u32)
PP_u64 = C.POINTER(P_u64)
sz = C.c_size_t


class ProductStats(C.Structure):
    _fields_ = [("tile_products", u64), ("out_tiles", u64), ("right_tiles_used", u64), ("alg_bytes", u64)]


class PrunedStats(C.Structure):
    _fields_ = [("structural_tiles", u64), ("numeric_tiles", u64), ("tile_products", u64), ("capacity", u64),
                ("scratch_bytes", u64), ("output_bytes", u64), ("passes", u32), ("fallback", u32)]


class HopStats(C.Structure):
    _fields_ = [
        ("tile_products", u64),
        ("left_tiles", u64),
        ("right_tiles_used", u64),
        ("out_tiles", u64),
        ("out_tiles_numeric", u64),
        ("popcount", u64),
        ("alg_bytes", u64),
        ("edges_visited", u64),
        ("symbolic_ms", f64),
        ("numeric_ms", f64),
        ("kernel", u32),
        ("segments", u32),
        ("left_tiles_numeric", u64),
        ("right_tiles_used_numeric", u64),
        ("alg_bytes_numeric", u64),
    ]


class ChunkRecord(C.Structure):
    _fields_ = [("panel", u32), ("strip_begin", u32), ("strip_end", u32), ("tile_begin", u32), ("tile_end", u32),
                ("bits", u32), ("jobs", u32), ("out_tiles", u32)]


class WalkQuery(C.Structure):
    _fields_ = [("graph", u32), ("n_sources", u32), ("sources", C.POINTER(u32)), ("hops", u32)]


class BatchOptions(C.Structure):
    _fields_ = [("packing", u32), ("chunk_tiles", u32), ("use_graph", u32), ("flags", u32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "mg_version": (C.c_char_p, []),
    "mg_last_error": (C.c_char_p, []),
    "mg_context_create": (i32, [i32, C.POINTER(vp)]),
    "mg_context_destroy": (i32, [vp]),
    "mg_context_set_stream": (i32, [vp, vp]),
    "mg_context_synchronize": (i32, [vp]),
    "mg_context_device": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "mg_context_stream": (i32, [vp, C.POINTER(vp)]),
    "mg_buffer_alloc": (i32, [vp, sz, C.POINTER(vp)]),
    "mg_buffer_free": (i32, [vp, vp]),
    "mg_buffer_upload": (i32, [vp, vp, vp, sz]),
    "mg_buffer_download": (i32, [vp, vp, vp, sz]),
    "mg_buffer_copy": (i32, [vp, vp, vp, sz]),
    # layer 1 (device pointers)
    "mg_matrix_bit_and": (i32, [vp, vp, vp, vp, u32, u32, u32]),
    "mg_matrix_bit_count": (i32, [vp, vp, vp, u64]),
    "mg_init_bit_tiled_matrix_metadata_by_layout_matrix": (i32, [vp, vp, u32, vp, vp, vp]),
    "mg_fill_tiles": (i32, [vp, sz, sz, sz, sz, sz, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "mg_count": (i32, [vp, sz, sz, sz, vp, vp]),
    # layer 2
    "mg_tiled_matrix_upload": (i32, [vp, u32, u32, u32, vp, vp, vp, vp, C.POINTER(vp)]),
    "mg_tiled_matrix_metadata": (i32, [vp, P_u32, P_u32, P_u32]),
    "mg_tiled_matrix_download": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "mg_tiled_matrix_device_arrays": (i32, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]),
    "mg_tiled_matrix_digest": (i32, [vp, vp, P_u64]),
    "mg_tiled_matrix_popcount": (i32, [vp, vp, P_u64]),
    "mg_tiled_matrix_prepare_operand": (i32, [vp, vp]),
    "mg_tiled_matrix_destroy": (i32, [vp]),
    "mg_io_read_block_metadata": (i32, [C.c_char_p, P_u32, P_u32, P_u32]),
    "mg_io_read_block": (i32, [C.c_char_p, vp, vp, vp, vp, vp]),
    "mg_io_write_block": (i32, [C.c_char_p, u32, u32, u32, vp, vp, vp, vp]),
    "mg_io_read_grid_metadata": (i32, [C.c_char_p, P_u32, P_u32, P_u32]),
    "mg_io_write_grid_metadata": (i32, [C.c_char_p, u32, u32, u32]),
    "mg_io_read_csr_tiled_block_info": (i32, [C.c_char_p, P_u32, P_u32, P_u32, P_u64]),
    "mg_io_read_csr_tiled_block": (i32, [C.c_char_p, vp, vp, u64, vp, vp]),
    "mg_io_write_csr_tiled_block": (i32, [C.c_char_p, u32, u32, vp, vp, u64]),
    "mg_io_read_csr_tiled_grid_metadata": (i32, [C.c_char_p, P_u32, P_u32, P_u32, P_u32]),
    "mg_io_write_csr_tiled_grid_metadata": (i32, [C.c_char_p, u32, u32, u32, u32]),
    "mg_multiply_symbolic": (i32, [vp, vp, vp, C.POINTER(vp), C.POINTER(ProductStats)]),
    "mg_multiply_numeric": (i32, [vp, vp, vp, vp]),
    "mg_multiply_numeric_dense": (i32, [vp, vp, vp, vp, u32, P_u64]),
    "mg_multiply_numeric_noalloc": (i32, [vp, vp, vp, vp]),
    "mg_context_timer_start": (i32, [vp]),
    "mg_context_timer_stop": (i32, [vp, C.POINTER(f64)]),
    "mg_multiply_pruned": (i32, [vp, vp, vp, u64, C.POINTER(vp), C.POINTER(PrunedStats)]),
    "mg_multiply_row_features": (i32, [vp, vp, vp, vp, vp, P_u64]),
    "mg_multiply_pruned_model": (i32, [vp, vp, vp, vp, C.c_float, vp, C.POINTER(vp), C.POINTER(PrunedStats)]),
    # edges
    "mg_edges_rmat": (i32, [vp, u32, u32, f64, f64, f64, u64, C.POINTER(vp)]),
    "mg_edges_er": (i32, [vp, u32, u64, u32, u64, C.POINTER(vp)]),
    "mg_edges_communities": (i32, [vp, u32, u32, u32, u32, u64, u32, u64, C.POINTER(vp)]),
    "mg_edges_union": (i32, [vp, vp, vp, C.POINTER(vp)]),
    "mg_edges_upload": (i32, [vp, u32, u64, vp, vp, C.POINTER(vp)]),
    "mg_edges_info": (i32, [vp, P_u32, P_u64]),
    "mg_edges_download": (i32, [vp, vp, vp, vp]),
    "mg_edges_digest": (i32, [vp, vp, P_u64]),
    "mg_edges_to_tiled_matrix": (i32, [vp, vp, i32, C.POINTER(vp)]),
    "mg_edges_destroy": (i32, [vp]),
    "mg_edges_grid_block": (i32, [vp, vp, u32, u32, u32, C.POINTER(vp)]),
    "mg_tiled_matrix_to_edges": (i32, [vp, vp, C.POINTER(vp)]),
    "mg_walk_batch_pushed_panels": (i32, [vp, C.POINTER(u32)]),
    "mg_edges_two_hop": (i32, [vp, vp, C.POINTER(vp)]),
    "mg_sample_sources": (i32, [u32, u32, u64, vp]),
    # walks
    "mg_walk_create": (i32, [vp, vp, u32, u32, u32, C.POINTER(vp)]),
    "mg_walk_run": (i32, [vp, vp, u32, vp, u32]),
    "mg_walk_set_timing": (i32, [vp, i32]),
    "mg_walk_set_hop_kernel": (i32, [vp, vp, u32]),
    "mg_walk_hop_kernel": (i32, [vp, C.POINTER(u32)]),
    "mg_walk_hop_stats": (i32, [vp, vp, u32, C.POINTER(HopStats)]),
    "mg_walk_kernel_times": (i32, [vp, vp, u32, C.POINTER(f64), C.POINTER(f64), P_u32]),
    "mg_walk_words_per_vertex": (i32, [vp, P_u32]),
    "mg_walk_frontier_download": (i32, [vp, vp, u32, vp]),
    "mg_walk_frontier_digest": (i32, [vp, vp, u32, P_u64]),
    "mg_walk_export": (i32, [vp, vp, u32, C.POINTER(vp)]),
    "mg_walk_destroy": (i32, [vp]),
    "matrixgraph_bit_matmult": (i32, [vp, vp, vp, i32, i32, i32]),
    "matrixgraph_bit_walk": (i32, [vp, vp, i32, i32, vp, i32, i32, vp, vp]),
    # batched walks
    "mg_walk_batch_create": (i32, [vp, vp, u32, vp, u32, vp, C.POINTER(vp)]),
    "mg_walk_batch_run": (i32, [vp, vp]),
    "mg_walk_batch_info": (i32, [vp, P_u32, P_u32, P_u64]),
    "mg_walk_batch_query_stats": (i32, [vp, vp, u32, u32, C.POINTER(HopStats)]),
    "mg_walk_batch_totals": (i32, [vp, vp, P_u64, P_u64, P_u64]),
    "mg_walk_batch_query_digest": (i32, [vp, vp, u32, P_u64]),
    "mg_walk_batch_query_download": (i32, [vp, vp, u32, vp]),
    "mg_walk_batch_query_export": (i32, [vp, vp, u32, C.POINTER(vp)]),
    "mg_walk_batch_destroy": (i32, [vp]),
    "mg_walk_batch_set_logging": (i32, [vp, i32]),
    "mg_walk_batch_set_hop_kernel": (i32, [vp, vp, u32]),
    "mg_walk_batch_level_chunks": (i32, [vp, vp, u32, P_u64, vp]),
    "mg_walk_batch_panel_info": (i32, [vp, u32, P_u32, P_u32, P_u32, vp]),
    "mg_walk_batch_set_level_order": (i32, [vp, u32, vp, u64]),
    # mlp
    "mg_mlp_create": (i32, [vp, u32, u32, u64, C.POINTER(vp)]),
    "mg_mlp_num_params": (i32, [vp, P_u32]),
    "mg_mlp_get_params": (i32, [vp, vp, vp, vp]),
    "mg_mlp_set_params": (i32, [vp, vp, vp]),
    "mg_mlp_forward": (i32, [vp, vp, vp, u32, vp]),
    "mg_mlp_loss_grad": (i32, [vp, vp, vp, vp, u32, C.POINTER(C.c_float), vp]),
    "mg_mlp_train": (i32, [vp, vp, vp, vp, u32, u32, u32, C.c_float, C.POINTER(C.c_float)]),
    "mg_mlp_destroy": (i32, [vp]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmatrixgraph_amd.so and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C matrixgraph_amd/csrc`"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
