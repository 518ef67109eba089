"""CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or the timed CPU baseline). The product library never
links or calls it. The C restatement lives in oracle.c; see oracle.h for the pins.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
TILE = 64

_lib = None

u32, u64, i32, i64, f64, vp = C.c_uint32, C.c_uint64, C.c_int, C.c_int64, C.c_double, C.c_void_p

_SIG = {
    "or_sm64_mix": (u64, [u64]),
    "or_sm64_at": (u64, [u64, u64]),
    "or_stream_key": (u64, [u64, u64]),
    "or_digest_term": (u64, [u64, u64]),
    "or_rmat_edges": (i64, [u32, u32, f64, f64, f64, u64, vp, vp, u64]),
    "or_er_edges": (i64, [u32, u64, u32, u64, vp, vp, u64]),
    "or_community_edges": (i64, [u32, u32, u32, u32, u64, u32, u64, vp, vp, u64]),
    "or_sort_unique_edges": (u64, [u64, vp, vp]),
    "or_edges_digest": (u64, [u64, vp, vp]),
    "or_sample_sources": (None, [u32, u32, u64, vp]),
    "or_count_tiles": (u64, [u32, u64, vp, vp, i32]),
    "or_edges_to_tiles": (None, [u32, u64, vp, vp, i32, vp, vp, vp, vp]),
    "or_tiles_digest": (u64, [u32, u32, vp, vp, vp]),
    "or_bool_matmul_bruteforce": (None, [u32, vp, vp, vp]),
    "or_bitset_matmul": (None, [u32, u32, u32, vp, vp, vp, i32]),
    "or_frontier_step": (None, [u32, vp, vp, u32, vp, vp, i32]),
    "or_frontier_digest": (u64, [u64, u32, u32, vp]),
    "or_in_csr": (None, [u32, u64, vp, vp, vp, vp]),
    "or_tile_symbolic": (i64, [u32, vp, vp, vp, vp, vp, vp, vp, u64, C.POINTER(u64)]),
    "or_tile_numeric": (None, [u32, vp, vp, vp, vp, vp, vp, u32, vp, vp, vp, i32]),
    "or_walk_tile_hop": (u64, [u32, vp, vp, vp, u32, vp, vp, u32, u32, u32, i32]),
    "or_walk_symbolic": (u64, [u32, vp, vp, vp, vp]),
    "or_max_threads": (i32, []),
    "or_mlp_loss_grad": (None, [u32, u32, vp, vp, vp, u32, vp, vp]),
    "or_f32_to_bf16": (C.c_uint16, [C.c_float]),
    "or_bf16_to_f32": (C.c_float, [C.c_uint16]),
}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} not built: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        for n, (r, a) in _SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# --- generators --------------------------------------------------------------
def rmat_edges(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    cap = edge_factor << scale
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    n = lib().or_rmat_edges(scale, edge_factor, a, b, c, seed, _p(src), _p(dst), cap)
    assert n >= 0
    return src[:n].copy(), dst[:n].copy()


def er_edges(log2_n, p_num, p_log2_den, seed, cap=None):
    n = 1 << log2_n
    expect = (n * n * p_num) >> p_log2_den
    cap = cap or int(expect * 1.2 + 1024)
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    m = lib().or_er_edges(log2_n, p_num, p_log2_den, seed, _p(src), _p(dst), cap)
    assert m >= 0, "capacity too small"
    return src[:m].copy(), dst[:m].copy()


def community_edges(n_vertices, n_comm, comm_size, stride, p_num, p_log2_den, seed, cap=None):
    expect = (n_comm * comm_size * comm_size * p_num) >> p_log2_den
    cap = cap or int(expect * 1.2 + 1024)
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    m = lib().or_community_edges(n_vertices, n_comm, comm_size, stride, p_num, p_log2_den, seed, _p(src), _p(dst), cap)
    assert m >= 0, "capacity too small"
    return src[:m].copy(), dst[:m].copy()


def sort_unique_edges(src, dst):
    s = _c(src, np.uint32).copy()
    d = _c(dst, np.uint32).copy()
    n = lib().or_sort_unique_edges(len(s), _p(s), _p(d))
    return s[:n].copy(), d[:n].copy()


def edges_digest(src, dst) -> int:
    s, d = _c(src, np.uint32), _c(dst, np.uint32)
    return lib().or_edges_digest(len(s), _p(s), _p(d))


def sample_sources(n_vertices, count, seed):
    out = np.empty(count, np.uint32)
    lib().or_sample_sources(n_vertices, count, seed, _p(out))
    return out


# --- converter ---------------------------------------------------------------
def edges_to_tiles(n_vertices, src, dst, transpose=False):
    """Returns (n_strips, offset_row, row_idx, col_idx, data[n, 64])."""
    s, d = _c(src, np.uint32), _c(dst, np.uint32)
    ns = (n_vertices + TILE - 1) // TILE
    nt = lib().or_count_tiles(n_vertices, len(s), _p(s), _p(d), int(transpose))
    off = np.empty(ns + 1, np.uint32)
    row = np.empty(nt, np.uint32)
    col = np.empty(nt, np.uint32)
    data = np.empty((nt, TILE), np.uint64)
    lib().or_edges_to_tiles(n_vertices, len(s), _p(s), _p(d), int(transpose), _p(off), _p(row), _p(col), _p(data))
    return ns, off, row, col, data


def tiles_digest(n_strips, row, col, data) -> int:
    row, col, data = _c(row, np.uint32), _c(col, np.uint32), _c(data, np.uint64)
    return lib().or_tiles_digest(n_strips, len(row), _p(row), _p(col), _p(data))


# --- O1 / O2 / O3 --------------------------------------------------------------
def bool_matmul_bruteforce(a, b):
    a = _c(a, np.uint8)
    b = _c(b, np.uint8)
    n = a.shape[0]
    c = np.empty_like(a)
    lib().or_bool_matmul_bruteforce(n, _p(a), _p(b), _p(c))
    return c


def pack_rows(dense):
    """uint8 (m, k) -> uint64 (m, ceil(k/64)) with bit j of word w = column 64w+j."""
    m, k = dense.shape
    kw = (k + 63) // 64
    pad = np.zeros((m, kw * 64), np.uint8)
    pad[:, :k] = dense
    return np.packbits(pad, axis=1, bitorder="little").view(np.uint64).reshape(m, kw).copy()


def unpack_rows(words, k):
    m = words.shape[0]
    bits = np.unpackbits(np.ascontiguousarray(words).view(np.uint8).reshape(m, -1), axis=1, bitorder="little")
    return bits[:, :k]


def bitset_matmul(a_rows, b_rows, m, k, n, threads=0):
    a_rows, b_rows = _c(a_rows, np.uint64), _c(b_rows, np.uint64)
    c = np.empty((m, (n + 63) // 64), np.uint64)
    lib().or_bitset_matmul(m, k, n, _p(a_rows), _p(b_rows), _p(c), threads)
    return c


def in_csr(n_vertices, src, dst):
    """In-neighbour CSR: in_off[v]..in_off[v+1] lists u with edge u -> v."""
    src = _c(src, np.uint32)
    dst = _c(dst, np.uint32)
    off = np.empty(n_vertices + 1, np.uint64)
    in_src = np.empty(len(src), np.uint32)
    lib().or_in_csr(n_vertices, len(src), _p(src), _p(dst), _p(off), _p(in_src))
    return off, in_src


def frontier_step(n_vertices, in_off, in_src, x, threads=0):
    x = _c(x, np.uint64)
    W = x.shape[1]
    y = np.empty_like(x)
    lib().or_frontier_step(n_vertices, _p(_c(in_off, np.uint64)), _p(_c(in_src, np.uint32)), W, _p(x), _p(y), threads)
    return y


def one_hot_frontier(n_vertices, sources, W=64):
    x = np.zeros((n_vertices, W), np.uint64)
    for s, v in enumerate(np.asarray(sources, np.int64)):
        x[v, s // 64] |= np.uint64(1) << np.uint64(s % 64)
    return x


def frontier_digest(x, n_sources) -> int:
    x = _c(x, np.uint64)
    return lib().or_frontier_digest(x.shape[0], x.shape[1], (n_sources + 63) // 64, _p(x))


def walk_frontiers(n_vertices, src, dst, sources, hops, W=64, threads=0):
    """O3 walk: list of vertex-major X_0..X_hops."""
    off, isrc = in_csr(n_vertices, src, dst)
    xs = [one_hot_frontier(n_vertices, sources, W)]
    for _ in range(hops):
        xs.append(frontier_step(n_vertices, off, isrc, xs[-1], threads))
    return xs


# --- O4 --------------------------------------------------------------------------
def tile_symbolic(n_strips, a_off, a_col, bt_off, bt_col):
    a_off, a_col, bt_off, bt_col = (_c(v, np.uint32) for v in (a_off, a_col, bt_off, bt_col))
    cap = n_strips * n_strips
    c_off = np.empty(n_strips + 1, np.uint32)
    c_row = np.empty(cap, np.uint32)
    c_col = np.empty(cap, np.uint32)
    tp = C.c_uint64()
    n = lib().or_tile_symbolic(n_strips, _p(a_off), _p(a_col), _p(bt_off), _p(bt_col), _p(c_off), _p(c_row),
                               _p(c_col), cap, C.byref(tp))
    assert n >= 0
    return c_off, c_row[:n].copy(), c_col[:n].copy(), tp.value


def tile_numeric(n_strips, a, bt, c_row, c_col, threads=0):
    """a, bt: (off, row, col, data) tuples; returns C data (n_c, 64)."""
    c_row, c_col = _c(c_row, np.uint32), _c(c_col, np.uint32)
    out = np.zeros((len(c_row), TILE), np.uint64)
    lib().or_tile_numeric(n_strips, _p(_c(a[0], np.uint32)), _p(_c(a[2], np.uint32)), _p(_c(a[3], np.uint64)),
                          _p(_c(bt[0], np.uint32)), _p(_c(bt[2], np.uint32)), _p(_c(bt[3], np.uint64)), len(c_row),
                          _p(c_row), _p(c_col), _p(out), threads)
    return out


def walk_tile_hop(n_strips, at_off, at_col, at_data, x_in, strip_begin=0, strip_end=None, threads=0, x_out=None,
                  stride=1):
    """CPU tile algorithm for one hop (the timed CPU baseline) on strips
    strip_begin, strip_begin+stride, ... < strip_end. Returns (x_out, visited)."""
    x_in = _c(x_in, np.uint64)
    W = x_in.shape[1]
    strip_end = n_strips if strip_end is None else strip_end
    if x_out is None:
        x_out = np.zeros_like(x_in)
    v = lib().or_walk_tile_hop(n_strips, _p(_c(at_off, np.uint32)), _p(_c(at_col, np.uint32)),
                               _p(_c(at_data, np.uint64)), W, _p(x_in), _p(x_out), strip_begin, strip_end, stride,
                               threads)
    return x_out, v


def walk_symbolic(n_strips, at_off, at_col, s_in):
    s_in = _c(s_in, np.uint64)
    s_out = np.empty_like(s_in)
    tp = lib().or_walk_symbolic(n_strips, _p(_c(at_off, np.uint32)), _p(_c(at_col, np.uint32)), _p(s_in), _p(s_out))
    return s_out, tp


def max_threads() -> int:
    return lib().or_max_threads()


# --- MLP ------------------------------------------------------------------------
def mlp_n_params(n_in, n_out):
    return 64 * n_in + 64 + 64 * 64 + 64 + n_out * 64 + n_out


def mlp_loss_grad(n_in, n_out, w, x, t):
    """FP64 loss and gradient (oracle) for weights w (float64 array)."""
    w = _c(w, np.float64)
    x = _c(x, np.float32)
    t = _c(t, np.float32)
    grad = np.empty(mlp_n_params(n_in, n_out), np.float64)
    loss = C.c_double()
    lib().or_mlp_loss_grad(n_in, n_out, _p(w), _p(x), _p(t), x.shape[0], C.byref(loss), _p(grad))
    return loss.value, grad


def f32_to_bf16(f: float) -> int:
    return lib().or_f32_to_bf16(f)


def bf16_to_f32(h: int) -> float:
    return lib().or_bf16_to_f32(h)
