"""MI355X-native bit-tiled Boolean walk engine (Python mirror of the C-ABI).

Objects here own C handles; every computation runs in libmatrixgraph_amd.so on the GPU.
Host-side containers (HostTiledMatrix) only hold arrays in the reference's layout
(core/data_structures/bit_tiled_matrix.cuh:17-72).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi

TILE = 64


class MGError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed (status {code}): {msg}")
        self.code = code


def _check(code: int, where: str) -> None:
    if code != 0:
        lib = _capi.load()
        raise MGError(code, where, lib.mg_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


def lib():
    return _capi.load()


def version() -> str:
    return lib().mg_version().decode()


# ----------------------------------------------------------------------------
# Host mirror of BitTiledMatrix
# ----------------------------------------------------------------------------
@dataclass
class HostTiledMatrix:
    """One block of bit tiles in the reference layout (tile_size 64).

    offset_row[n_strips+1], row_idx[n], col_idx[n] (uint32); data[n, 64] (uint64), word i of
    a tile = row i, bit j = column j (core/util/format_converter.cuh:408-410).
    """

    n_strips: int
    offset_row: np.ndarray
    row_idx: np.ndarray
    col_idx: np.ndarray
    data: np.ndarray
    tile_size: int = TILE

    @property
    def n_nz_tile(self) -> int:
        return int(self.row_idx.shape[0])

    @classmethod
    def empty(cls, n_strips: int) -> "HostTiledMatrix":
        return cls(n_strips, np.zeros(n_strips + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                   np.zeros((0, TILE), np.uint64))

    def nz_tile_bm(self) -> np.ndarray:
        """meta_buf/nz_tile_bm.bin words: n_strips^2/64 + 1 u64 (core/io/bit_tiled_matrix_io.cu:62-64)."""
        bits = self.n_strips * self.n_strips
        bm = np.zeros(bits // 64 + 1, np.uint64)
        pos = self.row_idx.astype(np.uint64) * np.uint64(self.n_strips) + self.col_idx.astype(np.uint64)
        np.bitwise_or.at(bm, (pos >> np.uint64(6)).astype(np.int64), np.uint64(1) << (pos & np.uint64(63)))
        return bm

    def to_dense(self, n: int | None = None) -> np.ndarray:
        """Dense uint8 n x n adjacency (n defaults to n_strips*64)."""
        n = self.n_strips * TILE if n is None else n
        full = np.zeros((self.n_strips * TILE, self.n_strips * TILE), np.uint8)
        if self.n_nz_tile:
            bits = np.unpackbits(self.data.view(np.uint8).reshape(self.n_nz_tile, TILE, 8), axis=2,
                                 bitorder="little").reshape(self.n_nz_tile, TILE, TILE)
            for t in range(self.n_nz_tile):
                x, y = int(self.row_idx[t]), int(self.col_idx[t])
                full[x * TILE:(x + 1) * TILE, y * TILE:(y + 1) * TILE] = bits[t]
        return full[:n, :n]

    @classmethod
    def from_dense(cls, dense: np.ndarray, keep_empty=None) -> "HostTiledMatrix":
        """Tiles with at least one set bit (plus any (x, y) in keep_empty), row-major."""
        n = dense.shape[0]
        ns = (n + TILE - 1) // TILE
        pad = np.zeros((ns * TILE, ns * TILE), np.uint8)
        pad[:n, :n] = dense
        blocks = pad.reshape(ns, TILE, ns, TILE).transpose(0, 2, 1, 3)
        occ = blocks.any(axis=(2, 3))
        if keep_empty is not None:
            for x, y in keep_empty:
                occ[x, y] = True
        xs, ys = np.nonzero(occ)
        data = np.packbits(blocks[xs, ys].astype(np.uint8), axis=2, bitorder="little").reshape(len(xs), TILE * 8)
        data = data.view(np.uint64).reshape(len(xs), TILE).copy()
        off = np.zeros(ns + 1, np.uint32)
        np.add.at(off, xs + 1, 1)
        off = np.cumsum(off).astype(np.uint32)
        return cls(ns, off, xs.astype(np.uint32), ys.astype(np.uint32), data)


# ----------------------------------------------------------------------------
# Context and device objects
# ----------------------------------------------------------------------------
class Context:
    def __init__(self, device: int = -1):
        h = C.c_void_p()
        _check(lib().mg_context_create(device, C.byref(h)), "mg_context_create")
        self.h = h

    def close(self):
        if self.h:
            lib().mg_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        _check(lib().mg_context_synchronize(self.h), "mg_context_synchronize")

    def timer_start(self):
        _check(lib().mg_context_timer_start(self.h), "mg_context_timer_start")

    def timer_stop(self) -> float:
        """Milliseconds of device time on the context's stream since timer_start (waits)."""
        ms = C.c_double()
        _check(lib().mg_context_timer_stop(self.h, C.byref(ms)), "mg_context_timer_stop")
        return ms.value

    def set_stream(self, stream_ptr: int | None):
        _check(lib().mg_context_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None),
               "mg_context_set_stream")

    @property
    def n_compute_units(self) -> int:
        d, n = C.c_int(), C.c_int()
        _check(lib().mg_context_device(self.h, C.byref(d), C.byref(n)), "mg_context_device")
        return n.value


    @property
    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().mg_context_stream(self.h, C.byref(s)), "mg_context_stream")
        return s.value or 0


class DeviceBuffer:
    """Zero-filled device allocation (DeviceOwnedBuffer semantics) for layer-1 calls."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        p = C.c_void_p()
        _check(lib().mg_buffer_alloc(ctx.h, nbytes, C.byref(p)), "mg_buffer_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx: Context, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(ctx, a.nbytes)
        _check(lib().mg_buffer_upload(ctx.h, C.c_void_p(b.ptr), _ptr(a), a.nbytes), "mg_buffer_upload")
        return b

    def to_array(self, dtype, count=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dt)
        _check(lib().mg_buffer_download(self.ctx.h, _ptr(out), C.c_void_p(self.ptr), out.nbytes), "mg_buffer_download")
        return out

    def close(self):
        if self.ptr:
            lib().mg_buffer_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TiledMatrix:
    """Device-resident bit-tiled matrix (one block)."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.h = handle

    @classmethod
    def upload(cls, ctx: Context, m: HostTiledMatrix) -> "TiledMatrix":
        off = np.ascontiguousarray(m.offset_row, np.uint32)
        row = np.ascontiguousarray(m.row_idx, np.uint32)
        col = np.ascontiguousarray(m.col_idx, np.uint32)
        data = np.ascontiguousarray(m.data, np.uint64)
        h = C.c_void_p()
        _check(lib().mg_tiled_matrix_upload(ctx.h, m.n_strips, m.n_nz_tile, m.tile_size, _ptr(off), _ptr(row),
                                            _ptr(col), _ptr(data), C.byref(h)), "mg_tiled_matrix_upload")
        return cls(ctx, h)

    def metadata(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib().mg_tiled_matrix_metadata(self.h, C.byref(a), C.byref(b), C.byref(c)), "mg_tiled_matrix_metadata")
        return a.value, b.value, c.value

    @property
    def n_strips(self) -> int:
        return self.metadata()[0]

    @property
    def n_nz_tile(self) -> int:
        return self.metadata()[1]

    def download(self) -> HostTiledMatrix:
        ns, n, _ = self.metadata()
        off = np.empty(ns + 1, np.uint32)
        row = np.empty(n, np.uint32)
        col = np.empty(n, np.uint32)
        data = np.empty((n, TILE), np.uint64)
        _check(lib().mg_tiled_matrix_download(self.ctx.h, self.h, _ptr(off), _ptr(row), _ptr(col), _ptr(data), None),
               "mg_tiled_matrix_download")
        return HostTiledMatrix(ns, off, row, col, data)

    def download_offsets(self) -> np.ndarray:
        """tile_offset_row only (n_strips + 1 entries)."""
        ns, _, _ = self.metadata()
        off = np.empty(ns + 1, np.uint32)
        _check(lib().mg_tiled_matrix_download(self.ctx.h, self.h, _ptr(off), None, None, None, None),
               "mg_tiled_matrix_download")
        return off

    def device_arrays(self):
        p = [C.c_void_p() for _ in range(4)]
        _check(lib().mg_tiled_matrix_device_arrays(self.h, *[C.byref(x) for x in p]), "mg_tiled_matrix_device_arrays")
        return tuple(x.value for x in p)

    def digest(self) -> int:
        d = C.c_uint64()
        _check(lib().mg_tiled_matrix_digest(self.ctx.h, self.h, C.byref(d)), "mg_tiled_matrix_digest")
        return d.value

    def popcount(self) -> int:
        d = C.c_uint64()
        _check(lib().mg_tiled_matrix_popcount(self.ctx.h, self.h, C.byref(d)), "mg_tiled_matrix_popcount")
        return d.value

    def prepare_operand(self) -> None:
        """Keeps this matrix's per-tile payloads and bit counts for the products it enters as an
        operand (mg_tiled_matrix_prepare_operand): a walk by products reuses A^T's at every hop."""
        _check(lib().mg_tiled_matrix_prepare_operand(self.ctx.h, self.h), "mg_tiled_matrix_prepare_operand")

    def to_edges(self) -> "EdgeList":
        """Every set bit as an edge (row, column), sorted, on n_strips*64 vertices."""
        h = C.c_void_p()
        _check(lib().mg_tiled_matrix_to_edges(self.ctx.h, self.h, C.byref(h)), "mg_tiled_matrix_to_edges")
        return EdgeList(self.ctx, h)

    def close(self):
        if self.h:
            lib().mg_tiled_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def multiply_symbolic(ctx: Context, a: TiledMatrix, bt: TiledMatrix):
    h = C.c_void_p()
    st = _capi.ProductStats()
    _check(lib().mg_multiply_symbolic(ctx.h, a.h, bt.h, C.byref(h), C.byref(st)), "mg_multiply_symbolic")
    return TiledMatrix(ctx, h), {f: getattr(st, f) for f, _ in _capi.ProductStats._fields_}


def multiply_numeric(ctx: Context, a: TiledMatrix, bt: TiledMatrix, c: TiledMatrix) -> None:
    _check(lib().mg_multiply_numeric(ctx.h, a.h, bt.h, c.h), "mg_multiply_numeric")


# Tile pairs whose two tiles both hold at least this many set bits are faster on the matrix
# cores than on the vector ALUs (C = A*A of planted communities: 0.48x at ~41 bits per tile,
# 1.32x at ~123, 1.76x at ~410).
DENSE_MIN_BITS = 128


def multiply_numeric_dense(ctx: Context, a: TiledMatrix, bt: TiledMatrix, c: TiledMatrix,
                           min_bits: int = DENSE_MIN_BITS) -> int:
    """Numeric pass with matrix-core products for tile pairs whose tiles both have at least
    min_bits set bits; returns the number of such products."""
    n = C.c_uint64()
    _check(lib().mg_multiply_numeric_dense(ctx.h, a.h, bt.h, c.h, min_bits, C.byref(n)), "mg_multiply_numeric_dense")
    return n.value


def multiply_numeric_noalloc(ctx: Context, a: TiledMatrix, bt: TiledMatrix, c: TiledMatrix) -> None:
    """The numeric pass without device allocation (static per-device payload scratch; A and
    Bt together at most NOALLOC_MAX_TILES tiles). Asynchronous on the context's stream."""
    _check(lib().mg_multiply_numeric_noalloc(ctx.h, a.h, bt.h, c.h), "mg_multiply_numeric_noalloc")


NOALLOC_MAX_TILES = (32 << 20) // 16 - 1

PRUNED_EXACT = 0
PRUNED_STRUCTURAL = (1 << 64) - 1
ROW_FEATURES = 8


def multiply_pruned(ctx: Context, a: TiledMatrix, bt: TiledMatrix, capacity: int = PRUNED_EXACT):
    """C = A * B keeping only numerically non-empty tiles (mg_multiply_pruned); capacity:
    PRUNED_EXACT (two passes), PRUNED_STRUCTURAL (one pass, structural scratch) or a predicted
    tile count (one pass, exact retry on overflow). Returns (C, stats dict)."""
    h = C.c_void_p()
    st = _capi.PrunedStats()
    _check(lib().mg_multiply_pruned(ctx.h, a.h, bt.h, int(capacity), C.byref(h), C.byref(st)), "mg_multiply_pruned")
    return TiledMatrix(ctx, h), {f: getattr(st, f) for f, _ in _capi.PrunedStats._fields_}


def multiply_pruned_model(ctx: Context, a: TiledMatrix, bt: TiledMatrix, model: "MLP", margin: float = 0.25):
    """mg_multiply_pruned sized by the cost model (device-side features and inference, exact
    retry on overflow). Returns (C, stats dict, per-row predicted non-empty tiles)."""
    h = C.c_void_p()
    st = _capi.PrunedStats()
    pred = np.zeros(a.n_strips, np.float32)
    _check(lib().mg_multiply_pruned_model(ctx.h, a.h, bt.h, model.h, margin, _ptr(pred), C.byref(h), C.byref(st)),
           "mg_multiply_pruned_model")
    return TiledMatrix(ctx, h), {f: getattr(st, f) for f, _ in _capi.PrunedStats._fields_}, pred


def multiply_row_features(ctx: Context, a: TiledMatrix, bt: TiledMatrix):
    """(features float32 (n_strips, ROW_FEATURES), structural tiles per C row, tile products)."""
    n = a.n_strips
    feat = np.zeros((n, ROW_FEATURES), np.float32)
    cnt = np.zeros(n, np.uint32)
    tp = C.c_uint64()
    _check(lib().mg_multiply_row_features(ctx.h, a.h, bt.h, _ptr(feat), _ptr(cnt),
                                          C.byref(tp)), "mg_multiply_row_features")
    return feat, cnt, tp.value


def multiply(ctx: Context, a: TiledMatrix, bt: TiledMatrix):
    """C = A * B (B supplied transposed): symbolic then numeric."""
    c, st = multiply_symbolic(ctx, a, bt)
    multiply_numeric(ctx, a, bt, c)
    return c, st


class EdgeList:
    """Device edge list, sorted by (src, dst), duplicate-free."""

    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self.h = handle

    @classmethod
    def rmat(cls, ctx, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
        h = C.c_void_p()
        _check(lib().mg_edges_rmat(ctx.h, scale, edge_factor, a, b, c, seed, C.byref(h)), "mg_edges_rmat")
        return cls(ctx, h)

    @classmethod
    def er(cls, ctx, log2_n, p_num, p_log2_den, seed):
        h = C.c_void_p()
        _check(lib().mg_edges_er(ctx.h, log2_n, p_num, p_log2_den, seed, C.byref(h)), "mg_edges_er")
        return cls(ctx, h)

    @classmethod
    def communities(cls, ctx, n_vertices, n_comm, comm_size, stride, p_num, p_log2_den, seed):
        h = C.c_void_p()
        _check(lib().mg_edges_communities(ctx.h, n_vertices, n_comm, comm_size, stride, p_num, p_log2_den, seed,
                                          C.byref(h)), "mg_edges_communities")
        return cls(ctx, h)

    @classmethod
    def upload(cls, ctx, n_vertices, src, dst):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = C.c_void_p()
        _check(lib().mg_edges_upload(ctx.h, n_vertices, len(src), _ptr(src), _ptr(dst), C.byref(h)), "mg_edges_upload")
        return cls(ctx, h)

    def union(self, other: "EdgeList") -> "EdgeList":
        h = C.c_void_p()
        _check(lib().mg_edges_union(self.ctx.h, self.h, other.h, C.byref(h)), "mg_edges_union")
        return EdgeList(self.ctx, h)

    def info(self):
        n, m = C.c_uint32(), C.c_uint64()
        _check(lib().mg_edges_info(self.h, C.byref(n), C.byref(m)), "mg_edges_info")
        return n.value, m.value

    def download(self):
        _, m = self.info()
        src = np.empty(m, np.uint32)
        dst = np.empty(m, np.uint32)
        _check(lib().mg_edges_download(self.ctx.h, self.h, _ptr(src), _ptr(dst)), "mg_edges_download")
        return src, dst

    def digest(self) -> int:
        d = C.c_uint64()
        _check(lib().mg_edges_digest(self.ctx.h, self.h, C.byref(d)), "mg_edges_digest")
        return d.value

    def two_hop(self) -> "EdgeList":
        """Distinct pairs (s, v), v != s, joined by a two-edge walk without self-loop steps."""
        h = C.c_void_p()
        _check(lib().mg_edges_two_hop(self.ctx.h, self.h, C.byref(h)), "mg_edges_two_hop")
        return EdgeList(self.ctx, h)

    def grid_block(self, n_chunks: int, bx: int, by: int) -> "EdgeList":
        """Block (bx, by) of the grid-cut partition in local ids, on ceil(n/n_chunks) vertices."""
        h = C.c_void_p()
        _check(lib().mg_edges_grid_block(self.ctx.h, self.h, n_chunks, bx, by, C.byref(h)), "mg_edges_grid_block")
        return EdgeList(self.ctx, h)

    def to_tiled_matrix(self, transpose: bool = False) -> TiledMatrix:
        h = C.c_void_p()
        _check(lib().mg_edges_to_tiled_matrix(self.ctx.h, self.h, 1 if transpose else 0, C.byref(h)),
               "mg_edges_to_tiled_matrix")
        return TiledMatrix(self.ctx, h)

    def close(self):
        if self.h:
            lib().mg_edges_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sample_sources(n_vertices: int, count: int, seed: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    _check(lib().mg_sample_sources(n_vertices, count, seed, _ptr(out)), "mg_sample_sources")
    return out


WALK_PRUNE = 1
WALK_SWEEP_FIRST_HOP = 2
WALK_VERTEX_ORDER = 4  # keep frontier strips in vertex order (no internal strip permutation)
HOP_SWEEP = 0
HOP_PUSH = 1
HOP_JOBS = 2
BATCH_SWEEP_FIRST_HOP = 1
BATCH_VERTEX_ORDER = 2


class Walk:
    """X_{k+1} = X_k * A from one-hot X_0, right operand A^T (device-resident frontiers)."""

    def __init__(self, ctx: Context, at: TiledMatrix, max_sources: int, max_hops: int, flags: int = 0):
        h = C.c_void_p()
        _check(lib().mg_walk_create(ctx.h, at.h, max_sources, max_hops, flags, C.byref(h)), "mg_walk_create")
        self.ctx, self.at, self.h = ctx, at, h
        self.n_strips = at.n_strips
        w = C.c_uint32()
        _check(lib().mg_walk_words_per_vertex(h, C.byref(w)), "mg_walk_words_per_vertex")
        self.words_per_vertex = w.value

    def set_timing(self, on: bool):
        _check(lib().mg_walk_set_timing(self.h, 1 if on else 0), "mg_walk_set_timing")

    def set_hop_kernel(self, segments: int) -> int:
        """Hop kernel of the 4096-source levels after the first: 0 = tile-decoding sweep,
        1/2/4/8 = job-list kernel with that many row segments (mg_walk_set_hop_kernel).
        Returns the choice in effect (0 where the job list does not apply)."""
        _check(lib().mg_walk_set_hop_kernel(self.ctx.h, self.h, segments), "mg_walk_set_hop_kernel")
        return self.hop_kernel()

    def hop_kernel(self) -> int:
        g = C.c_uint32()
        _check(lib().mg_walk_hop_kernel(self.h, C.byref(g)), "mg_walk_hop_kernel")
        return g.value

    def run(self, sources, hops: int) -> None:
        src = np.ascontiguousarray(sources, np.uint32)
        self._n_sources = len(src)
        _check(lib().mg_walk_run(self.ctx.h, self.h, len(src), _ptr(src), hops), "mg_walk_run")

    def hop_stats(self, hop: int) -> dict:
        st = _capi.HopStats()
        _check(lib().mg_walk_hop_stats(self.ctx.h, self.h, hop, C.byref(st)), "mg_walk_hop_stats")
        return {f: getattr(st, f) for f, _ in _capi.HopStats._fields_}

    def kernel_times(self, hop: int):
        """(symbolic_ms_sum, numeric_ms_sum, n_runs) over runs logged since set_timing(True)."""
        a, b, n = C.c_double(), C.c_double(), C.c_uint32()
        _check(lib().mg_walk_kernel_times(self.ctx.h, self.h, hop, C.byref(a), C.byref(b), C.byref(n)),
               "mg_walk_kernel_times")
        return a.value, b.value, n.value

    def frontier(self, hop: int) -> np.ndarray:
        """Vertex-major X_hop: shape (n_vertices, words_per_vertex) uint64."""
        out = np.empty((self.n_strips * TILE, self.words_per_vertex), np.uint64)
        _check(lib().mg_walk_frontier_download(self.ctx.h, self.h, hop, _ptr(out)), "mg_walk_frontier_download")
        return out

    def frontier_digest(self, hop: int) -> int:
        d = C.c_uint64()
        _check(lib().mg_walk_frontier_digest(self.ctx.h, self.h, hop, C.byref(d)), "mg_walk_frontier_digest")
        return d.value

    def export(self, hop: int) -> TiledMatrix:
        h = C.c_void_p()
        _check(lib().mg_walk_export(self.ctx.h, self.h, hop, C.byref(h)), "mg_walk_export")
        return TiledMatrix(self.ctx, h)

    def close(self):
        if self.h:
            lib().mg_walk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------
# On-disk format (host only): core/io/bit_tiled_matrix_io.cu:30-136
# ----------------------------------------------------------------------------
def _dir(path) -> bytes:
    p = str(path)
    if not p.endswith("/"):
        p += "/"
    return p.encode()


def write_block(path, m: HostTiledMatrix) -> None:
    off = np.ascontiguousarray(m.offset_row, np.uint32)
    row = np.ascontiguousarray(m.row_idx, np.uint32)
    col = np.ascontiguousarray(m.col_idx, np.uint32)
    data = np.ascontiguousarray(m.data, np.uint64)
    _check(lib().mg_io_write_block(_dir(path), m.n_strips, m.n_nz_tile, m.tile_size, _ptr(off), _ptr(row), _ptr(col),
                                   _ptr(data)), "mg_io_write_block")


def read_block(path) -> HostTiledMatrix:
    ns, nt, ts = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().mg_io_read_block_metadata(_dir(path), C.byref(ns), C.byref(nt), C.byref(ts)),
           "mg_io_read_block_metadata")
    off = np.empty(ns.value + 1, np.uint32)
    row = np.empty(nt.value, np.uint32)
    col = np.empty(nt.value, np.uint32)
    data = np.empty((nt.value, TILE), np.uint64)
    _check(lib().mg_io_read_block(_dir(path), _ptr(off), _ptr(row), _ptr(col), _ptr(data), None), "mg_io_read_block")
    return HostTiledMatrix(ns.value, off, row, col, data, ts.value)


def write_grid(path, blocks, n_vertices: int, n_edges: int) -> None:
    """blocks: list of n_chunks^2 HostTiledMatrix, gid = x*n_chunks + y."""
    n_chunks = int(round(len(blocks) ** 0.5))
    assert n_chunks * n_chunks == len(blocks)
    for gid, b in enumerate(blocks):
        write_block(f"{str(path).rstrip('/')}/block{gid}/", b)
    _check(lib().mg_io_write_grid_metadata(_dir(path), n_chunks, n_vertices, n_edges), "mg_io_write_grid_metadata")


def write_grid_from_edges(path, el: "EdgeList", n_chunks: int, transpose: bool = False) -> None:
    """Partitions el into an n_chunks x n_chunks grid on the device (grid cut, then one
    tile conversion per block) and writes it in the reference's on-disk grid format; with
    transpose the grid is that of A^T (block (x, y) = transpose of A's block (y, x))."""
    n_vertices, n_edges = el.info()
    blocks = []
    for x in range(n_chunks):
        for y in range(n_chunks):
            b = el.grid_block(n_chunks, y, x) if transpose else el.grid_block(n_chunks, x, y)
            tm = b.to_tiled_matrix(transpose=transpose)
            blocks.append(tm.download())
            tm.close()
            b.close()
    write_grid(path, blocks, n_vertices, n_edges)


def read_grid(path):
    nc, nv, ne = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().mg_io_read_grid_metadata(_dir(path), C.byref(nc), C.byref(nv), C.byref(ne)), "mg_io_read_grid_metadata")
    blocks = [read_block(f"{str(path).rstrip('/')}/block{g}/") for g in range(nc.value * nc.value)]
    return blocks, {"n_chunks": nc.value, "n_vertices": nv.value, "n_edges": ne.value}


# ----------------------------------------------------------------------------
# CSR-tiled grids (host only), the input format of the reference's enumerating walk:
# core/io/csr_tiled_matrix_io.cu:33-172, core/io/grid_csr_tiled_matrix_io.cu,
# tools/graph_converter/converter/to_grid_csr_tiled_matrix.cuh:38-85
# ----------------------------------------------------------------------------
def csr_tile_extent(block_scope: int, tile_size: int) -> int:
    """Tile extent (vertices per tile side) the reference converter uses for a tile_size
    argument: below 128 the argument counts strips and the extent is ceil(block_scope /
    tile_size) (core/util/format_converter.cuh:424-426; integer ceiling here, float there)."""
    if block_scope <= 0 or tile_size <= 0:
        raise MGError(1, "csr_tile_extent: block_scope and tile_size must be positive")
    return -(-block_scope // tile_size) if tile_size < 128 else tile_size


def write_csr_tiled_block(path, src, dst, block_scope: int, tile_extent: int) -> None:
    """One grid block's edges (global ids) -> path in CSRTiledMatrixIO's layout."""
    s = np.ascontiguousarray(src, np.uint32)
    d = np.ascontiguousarray(dst, np.uint32)
    if s.shape != d.shape:
        raise MGError(1, "write_csr_tiled_block: src and dst differ in length")
    _check(lib().mg_io_write_csr_tiled_block(_dir(path), block_scope, tile_extent, _ptr(s), _ptr(d), len(s)),
           "mg_io_write_csr_tiled_block")


def read_csr_tiled_block(path, with_tiles: bool = False):
    """Edges of one CSR-tiled block as (src, dst, info) in global ids, tile by tile; with
    with_tiles info also holds each edge's tile row and column."""
    ns, nt, ts, ne = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
    _check(lib().mg_io_read_csr_tiled_block_info(_dir(path), C.byref(ns), C.byref(nt), C.byref(ts), C.byref(ne)),
           "mg_io_read_csr_tiled_block_info")
    n = ne.value
    s, d = np.empty(n, np.uint32), np.empty(n, np.uint32)
    tr = np.empty(n, np.uint32) if with_tiles else None
    tc = np.empty(n, np.uint32) if with_tiles else None
    _check(lib().mg_io_read_csr_tiled_block(_dir(path), _ptr(s), _ptr(d), n, _ptr(tr) if with_tiles else None,
                                            _ptr(tc) if with_tiles else None), "mg_io_read_csr_tiled_block")
    info = {"n_strips": ns.value, "n_nz_tile": nt.value, "tile_size": ts.value, "n_edges": n}
    if with_tiles:
        info["tile_row"], info["tile_col"] = tr, tc
    return s, d, info


def write_csr_tiled_grid(path, n_vertices: int, src, dst, n_chunks: int = 1, tile_size: int = 64,
                         max_vid: int = None) -> None:
    """Edge list -> CSR-tiled grid as the reference's partitioner + converter lay it out:
    block scope S = ceil((max_vid + 1) / n_chunks) (tools/graph_partitioner/partitioner/
    grid_cut.cu:88-89), edge (u, v) in block (u // S) * n_chunks + v // S, every block written
    (empty ones as meta.yaml only), tile extent csr_tile_extent(S, tile_size)."""
    s = np.ascontiguousarray(src, np.uint32)
    d = np.ascontiguousarray(dst, np.uint32)
    if max_vid is None:
        max_vid = int(max(s.max(), d.max())) if len(s) else max(n_vertices - 1, 0)
    if n_chunks <= 0:
        raise MGError(1, "write_csr_tiled_grid: n_chunks must be positive")
    scope = -(-(max_vid + 1) // n_chunks)
    ext = csr_tile_extent(scope, tile_size)
    bx, by = s // scope, d // scope
    root = str(path).rstrip("/")
    for x in range(n_chunks):
        for y in range(n_chunks):
            m = (bx == x) & (by == y)
            write_csr_tiled_block(f"{root}/block{x * n_chunks + y}/", s[m], d[m], scope, ext)
    _check(lib().mg_io_write_csr_tiled_grid_metadata(_dir(path), n_chunks, n_vertices, len(s), max_vid),
           "mg_io_write_csr_tiled_grid_metadata")


def read_csr_tiled_grid(path):
    """All blocks of a CSR-tiled grid -> (src, dst, meta): the global edge list, ready for
    EdgeList.upload(ctx, max(n_vertices, max_vid + 1), src, dst)."""
    nc, nv, ne, mv = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib().mg_io_read_csr_tiled_grid_metadata(_dir(path), C.byref(nc), C.byref(nv), C.byref(ne), C.byref(mv)),
           "mg_io_read_csr_tiled_grid_metadata")
    root = str(path).rstrip("/")
    parts = [read_csr_tiled_block(f"{root}/block{g}/")[:2] for g in range(nc.value * nc.value)]
    s = np.concatenate([p[0] for p in parts]) if parts else np.empty(0, np.uint32)
    d = np.concatenate([p[1] for p in parts]) if parts else np.empty(0, np.uint32)
    return s, d, {"n_chunks": nc.value, "n_vertices": nv.value, "n_edges": ne.value, "max_vid": mv.value}


# ----------------------------------------------------------------------------
# Batched walks
# ----------------------------------------------------------------------------
class WalkBatch:
    """Many queries (workloads.QuerySpec: graph, sources, hops) over several graphs (A^T),
    run as one captured HIP graph (mg_walk_batch_*)."""

    def __init__(self, ctx: Context, graphs_at, queries, packing: int = 0, chunk_tiles: int = 0,
                 use_graph: bool = True, sweep_first_hop: bool = False, vertex_order: bool = False):
        self.ctx = ctx
        self.graphs = list(graphs_at)
        self.queries = list(queries)
        srcs = [np.ascontiguousarray(q.sources, np.uint32) for q in self.queries]
        arr = (_capi.WalkQuery * len(self.queries))()
        for i, (q, s) in enumerate(zip(self.queries, srcs)):
            arr[i].graph = q.graph
            arr[i].n_sources = len(s)
            arr[i].sources = s.ctypes.data_as(C.POINTER(C.c_uint32))
            arr[i].hops = q.hops
        gp = (C.c_void_p * len(self.graphs))(*[g.h.value if hasattr(g.h, "value") else g.h for g in self.graphs])
        opt = _capi.BatchOptions(packing, chunk_tiles, 1 if use_graph else 0,
                                 (BATCH_SWEEP_FIRST_HOP if sweep_first_hop else 0)
                                 | (BATCH_VERTEX_ORDER if vertex_order else 0))
        h = C.c_void_p()
        _check(lib().mg_walk_batch_create(ctx.h, gp, len(self.graphs), arr, len(self.queries), C.byref(opt),
                                          C.byref(h)), "mg_walk_batch_create")
        self.h = h

    def run(self):
        _check(lib().mg_walk_batch_run(self.ctx.h, self.h), "mg_walk_batch_run")

    def info(self):
        p, hmax, ch = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().mg_walk_batch_info(self.h, C.byref(p), C.byref(hmax), C.byref(ch)), "mg_walk_batch_info")
        return p.value, hmax.value, ch.value

    def pushed_panels(self) -> int:
        """Panels whose first hop is pushed from their source strips' columns."""
        n = C.c_uint32()
        _check(lib().mg_walk_batch_pushed_panels(self.h, C.byref(n)), "mg_walk_batch_pushed_panels")
        return n.value

    def totals(self):
        tp, by, ed = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().mg_walk_batch_totals(self.ctx.h, self.h, C.byref(tp), C.byref(by), C.byref(ed)),
               "mg_walk_batch_totals")
        return {"tile_products": tp.value, "alg_bytes": by.value, "edges_visited": ed.value}

    def query_digest(self, q: int) -> int:
        d = C.c_uint64()
        _check(lib().mg_walk_batch_query_digest(self.ctx.h, self.h, q, C.byref(d)), "mg_walk_batch_query_digest")
        return d.value

    def query_frontier(self, q: int) -> np.ndarray:
        g = self.graphs[self.queries[q].graph]
        nw = (len(self.queries[q].sources) + 63) // 64
        out = np.empty((g.n_strips * TILE, nw), np.uint64)
        _check(lib().mg_walk_batch_query_download(self.ctx.h, self.h, q, _ptr(out)), "mg_walk_batch_query_download")
        return out

    # --- scheduler control (cost models) ---------------------------------------------
    def set_logging(self, on: bool):
        _check(lib().mg_walk_batch_set_logging(self.h, 1 if on else 0), "mg_walk_batch_set_logging")

    def set_hop_kernel(self, segments: int):
        """Hop kernel of the levels after the first (see Walk.set_hop_kernel)."""
        _check(lib().mg_walk_batch_set_hop_kernel(self.ctx.h, self.h, segments), "mg_walk_batch_set_hop_kernel")

    def level_chunks(self, level: int) -> np.ndarray:
        """Chunks of hop level `level` in launch order (structured array of mg_chunk_record)."""
        n = C.c_uint64()
        _check(lib().mg_walk_batch_level_chunks(self.ctx.h, self.h, level, C.byref(n), None),
               "mg_walk_batch_level_chunks")
        arr = (_capi.ChunkRecord * max(1, n.value))()
        _check(lib().mg_walk_batch_level_chunks(self.ctx.h, self.h, level, C.byref(n), arr),
               "mg_walk_batch_level_chunks")
        dt = np.dtype([(f, np.uint32) for f, _ in _capi.ChunkRecord._fields_])
        return np.frombuffer(bytes(arr), dtype=dt, count=n.value).copy()

    def panel_info(self, p: int):
        g, h, w = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(lib().mg_walk_batch_panel_info(self.h, p, C.byref(g), C.byref(h), C.byref(w), None),
               "mg_walk_batch_panel_info")
        alive = np.zeros(max(1, h.value), np.uint64)
        _check(lib().mg_walk_batch_panel_info(self.h, p, None, None, None, _ptr(alive)), "mg_walk_batch_panel_info")
        return {"graph": g.value, "hops": h.value, "words_used": w.value, "alive": alive[:h.value]}

    def set_level_order(self, level: int, order):
        o = np.ascontiguousarray(order, np.uint32)
        _check(lib().mg_walk_batch_set_level_order(self.h, level, _ptr(o), len(o)), "mg_walk_batch_set_level_order")

    def query_export(self, q: int) -> TiledMatrix:
        h = C.c_void_p()
        _check(lib().mg_walk_batch_query_export(self.ctx.h, self.h, q, C.byref(h)), "mg_walk_batch_query_export")
        return TiledMatrix(self.ctx, h)

    def close(self):
        if self.h:
            lib().mg_walk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------
# MLP cost model
# ----------------------------------------------------------------------------
class MLP:
    """n_in -> 64 -> 64 -> n_out ReLU MLP, BF16 weights, FP32 accumulation (mg_mlp_*)."""

    def __init__(self, ctx: Context, n_in: int, n_out: int, seed: int = 0):
        self.ctx, self.n_in, self.n_out = ctx, n_in, n_out
        h = C.c_void_p()
        _check(lib().mg_mlp_create(ctx.h, n_in, n_out, seed, C.byref(h)), "mg_mlp_create")
        self.h = h
        n = C.c_uint32()
        _check(lib().mg_mlp_num_params(h, C.byref(n)), "mg_mlp_num_params")
        self.n_params = n.value

    def params(self):
        master = np.empty(self.n_params, np.float32)
        bf16 = np.empty(self.n_params, np.uint16)
        _check(lib().mg_mlp_get_params(self.ctx.h, self.h, _ptr(master), _ptr(bf16)), "mg_mlp_get_params")
        return master, bf16

    def set_params(self, master):
        m = np.ascontiguousarray(master, np.float32)
        _check(lib().mg_mlp_set_params(self.ctx.h, self.h, _ptr(m)), "mg_mlp_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        dx = DeviceBuffer.from_array(self.ctx, x)
        dy = DeviceBuffer(self.ctx, x.shape[0] * self.n_out * 4)
        _check(lib().mg_mlp_forward(self.ctx.h, self.h, C.c_void_p(dx.ptr), x.shape[0], C.c_void_p(dy.ptr)),
               "mg_mlp_forward")
        self.ctx.synchronize()
        return dy.to_array(np.float32).reshape(x.shape[0], self.n_out)

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceBuffer.from_array(self.ctx, x), DeviceBuffer.from_array(self.ctx, t)
        loss = C.c_float()
        grad = np.empty(self.n_params, np.float32)
        _check(lib().mg_mlp_loss_grad(self.ctx.h, self.h, C.c_void_p(dx.ptr), C.c_void_p(dt.ptr), x.shape[0],
                                      C.byref(loss), _ptr(grad)), "mg_mlp_loss_grad")
        return loss.value, grad

    def train(self, x, t, epochs=10, batch=1024, lr=1e-3) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dx, dt = DeviceBuffer.from_array(self.ctx, x), DeviceBuffer.from_array(self.ctx, t)
        loss = C.c_float()
        _check(lib().mg_mlp_train(self.ctx.h, self.h, C.c_void_p(dx.ptr), C.c_void_p(dt.ptr), x.shape[0], epochs,
                                  batch, lr, C.byref(loss)), "mg_mlp_train")
        return loss.value

    def close(self):
        if self.h:
            lib().mg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
