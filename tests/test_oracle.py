"""Pins of the CPU oracle (no GPU): closed forms, brute force, cross-checks between the
independent oracles, generator invariants and the reference's toy-graph fixture."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
RNG = np.random.default_rng(12345)


def dense_from_edges(n, src, dst):
    a = np.zeros((n, n), np.uint8)
    a[np.asarray(src, np.int64), np.asarray(dst, np.int64)] = 1
    return a


def edges_from_tiles(ns, off, row, col, data):
    out = []
    for t in range(len(row)):
        for r in range(64):
            w = int(data[t, r])
            while w:
                c = (w & -w).bit_length() - 1
                w &= w - 1
                out.append((int(row[t]) * 64 + r, int(col[t]) * 64 + c))
    return sorted(out)


def random_graph(n, m, rng=RNG):
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    return O.sort_unique_edges(src, dst)


# --- O1: closed forms ---------------------------------------------------------
def test_bruteforce_identity_zero_permutations():
    n = 37
    a = (RNG.random((n, n)) < 0.2).astype(np.uint8)
    eye = np.eye(n, dtype=np.uint8)
    assert np.array_equal(O.bool_matmul_bruteforce(a, eye), a)
    assert np.array_equal(O.bool_matmul_bruteforce(eye, a), a)
    assert not O.bool_matmul_bruteforce(a, np.zeros_like(a)).any()
    p1, p2 = RNG.permutation(n), RNG.permutation(n)
    P1 = np.zeros((n, n), np.uint8); P1[np.arange(n), p1] = 1
    P2 = np.zeros((n, n), np.uint8); P2[np.arange(n), p2] = 1
    comp = np.zeros((n, n), np.uint8); comp[np.arange(n), p2[p1]] = 1
    assert np.array_equal(O.bool_matmul_bruteforce(P1, P2), comp)


def test_bruteforce_cycle_powers():
    n = 29
    c = np.zeros((n, n), np.uint8)
    c[np.arange(n), (np.arange(n) + 1) % n] = 1
    p = c.copy()
    for k in range(2, 6):
        p = O.bool_matmul_bruteforce(p, c)
        expect = np.zeros((n, n), np.uint8)
        expect[np.arange(n), (np.arange(n) + k) % n] = 1
        assert np.array_equal(p, expect)


def test_bruteforce_single_bits_transpose():
    n = 20
    for _ in range(20):
        i, j, k = RNG.integers(0, n, 3)
        a = np.zeros((n, n), np.uint8); a[i, j] = 1
        b = np.zeros((n, n), np.uint8); b[j, k] = 1
        e = np.zeros((n, n), np.uint8); e[i, k] = 1
        assert np.array_equal(O.bool_matmul_bruteforce(a, b), e)
    a = (RNG.random((n, n)) < 0.3).astype(np.uint8)
    b = (RNG.random((n, n)) < 0.3).astype(np.uint8)
    assert np.array_equal(O.bool_matmul_bruteforce(a, b).T, O.bool_matmul_bruteforce(b.T.copy(), a.T.copy()))


# --- O2 vs O1 -------------------------------------------------------------------
@pytest.mark.parametrize("m,k,n,dens", [(5, 70, 130, 0.1), (64, 64, 64, 0.05), (100, 129, 65, 0.3), (1, 1, 1, 1.0)])
def test_bitset_matches_bruteforce(m, k, n, dens):
    size = max(m, k, n)
    a = np.zeros((size, size), np.uint8); a[:m, :k] = RNG.random((m, k)) < dens
    b = np.zeros((size, size), np.uint8); b[:k, :n] = RNG.random((k, n)) < dens
    c1 = O.bool_matmul_bruteforce(a, b)[:m, :n]
    c2 = O.unpack_rows(O.bitset_matmul(O.pack_rows(a[:m, :k]), O.pack_rows(b[:k, :n]), m, k, n), n)
    assert np.array_equal(c1, c2)


# --- O3 vs O2 and closed forms ---------------------------------------------------
def test_frontier_matches_bitset_powers():
    n = 300
    src, dst = random_graph(n, 1500)
    a = dense_from_edges(n, src, dst)
    sources = RNG.choice(n, 70, replace=False).astype(np.uint32)
    xs = O.walk_frontiers(n, src, dst, sources, 4, W=2)
    rows = np.zeros((len(sources), n), np.uint8)
    rows[np.arange(len(sources)), sources] = 1
    cur = O.pack_rows(rows)
    apk = O.pack_rows(a)
    for k in range(1, 5):
        cur = O.bitset_matmul(cur, apk, len(sources), n, n)
        expect = O.unpack_rows(cur, n)  # (sources, n)
        got = O.unpack_rows(xs[k], 128)[:, :len(sources)].T  # vertex-major -> (sources, n)
        assert np.array_equal(expect, got), k


def test_frontier_cycle_and_path_closed_forms():
    n = 1000
    v = np.arange(n, dtype=np.uint32)
    cyc = O.walk_frontiers(n, v, (v + 1) % n, np.arange(64, dtype=np.uint32) * 7, 5, W=1)
    path = O.walk_frontiers(n, v[:-1], v[1:], np.arange(64, dtype=np.uint32) * 15 + 40, 3, W=1)
    for k in range(6):
        x = cyc[k]
        assert int(np.unpackbits(x.view(np.uint8)).sum()) == 64  # exactly one vertex per source per hop
        for s in range(64):
            assert (int(x[(7 * s + k) % n, 0]) >> s) & 1
    for k in range(4):
        x = path[k]
        for s in range(64):
            t = 15 * s + 40 + k
            if t < n:
                assert (int(x[t, 0]) >> s) & 1
    # DAG hop bound: on a path of 1000 vertices nothing survives 1000 hops from vertex 0
    short = O.walk_frontiers(5, np.array([0, 1, 2, 3], np.uint32), np.array([1, 2, 3, 4], np.uint32),
                             np.array([0], np.uint32), 5, W=1)
    assert short[4][4, 0] == 1 and not short[5].any()


# --- O4 vs dense ------------------------------------------------------------------
def test_tile_oracle_matches_dense_product():
    n = 500
    src, dst = random_graph(n, 4000)
    a = dense_from_edges(n, src, dst)
    ns, a_off, a_row, a_col, a_data = O.edges_to_tiles(n, src, dst)
    _, t_off, t_row, t_col, t_data = O.edges_to_tiles(n, src, dst, transpose=True)
    c_off, c_row, c_col, tp = O.tile_symbolic(ns, a_off, a_col, t_off, t_col)
    # structural set = tile occupancy product (numpy integer matmul, independent)
    occ = np.zeros((ns, ns), np.int64); occ[a_row, a_col] = 1
    prod = occ @ occ
    assert set(zip(c_row.tolist(), c_col.tolist())) == set(zip(*np.nonzero(prod)))
    assert tp == int(prod.sum())
    c_data = O.tile_numeric(ns, (a_off, a_row, a_col, a_data), (t_off, t_row, t_col, t_data), c_row, c_col)
    dense = np.zeros((ns * 64, ns * 64), np.uint8)
    bits = np.unpackbits(c_data.view(np.uint8).reshape(-1, 64, 8), axis=2, bitorder="little")
    for t in range(len(c_row)):
        dense[c_row[t] * 64:(c_row[t] + 1) * 64, c_col[t] * 64:(c_col[t] + 1) * 64] = bits[t]
    expect = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(a), n, n, n), n)
    assert np.array_equal(dense[:n, :n], expect)


def test_walk_tile_hop_and_symbolic():
    n = 64 * 9 + 5
    src, dst = random_graph(n, 5000)
    ns, t_off, t_row, t_col, t_data = O.edges_to_tiles(n, src, dst, transpose=True)
    nv = ns * 64
    sources = RNG.choice(n, 200, replace=False).astype(np.uint32)
    xs = O.walk_frontiers(nv, src, dst, sources, 3, W=4)
    occ_at = np.zeros((ns, ns), bool); occ_at[t_row, t_col] = True
    s = np.zeros(ns, np.uint64)
    for i, v in enumerate(sources):
        s[v // 64] |= np.uint64(1) << np.uint64(i // 64)
    for k in range(3):
        got, visited = O.walk_tile_hop(ns, t_off, t_col, t_data, xs[k])
        assert np.array_equal(got, xs[k + 1])
        assert visited == len(src)
        s_next, tp = O.walk_symbolic(ns, t_off, t_col, s)
        expect = np.array([np.bitwise_or.reduce(s[occ_at[y]]) if occ_at[y].any() else 0 for y in range(ns)],
                          np.uint64)
        assert np.array_equal(s_next, expect)
        assert tp == sum(bin(int(s[t])).count("1") for t in t_col)
        s = s_next


# --- the reference's toy graph ------------------------------------------------------
def test_toy_graph_fixture():
    g = json.loads((GOLDEN / "toy_graph.json").read_text())
    n = g["n_vertices"]
    e = np.array(g["edges"], np.uint32)
    src, dst = O.sort_unique_edges(e[:, 0], e[:, 1])
    ns, off, row, col, data = O.edges_to_tiles(n, src, dst)
    assert ns == 1 and list(off) == [0, 1] and list(row) == [0] and list(col) == [0]
    assert [str(int(w)) for w in data[0]] == g["tile_words_A"]
    a = dense_from_edges(n, src, dst)
    xs = O.walk_frontiers(n, src, dst, np.arange(n, dtype=np.uint32), 4, W=1)
    cur = a
    for k in range(1, 5):
        expect = set(map(tuple, g["powers"][str(k)]))
        if k > 1:
            cur = O.bool_matmul_bruteforce(cur, a)
        assert set(zip(*np.nonzero(cur))) == expect
        got = {(s, v) for v in range(n) for s in range(n) if (int(xs[k][v, 0]) >> s) & 1}
        assert got == expect
        assert len(expect) == g["total_count"][str(k)]
    _, t_off, t_row, t_col, t_data = O.edges_to_tiles(n, src, dst, transpose=True)
    c_off, c_row, c_col, _ = O.tile_symbolic(ns, off, col, t_off, t_col)
    c = O.tile_numeric(ns, (off, row, col, data), (t_off, t_row, t_col, t_data), c_row, c_col)
    got2 = {(r, cc) for r in range(64) for cc in range(64) if (int(c[0, r]) >> cc) & 1}
    assert got2 == set(map(tuple, g["powers"]["2"]))


# --- generators and converter ---------------------------------------------------------
def test_rmat_invariants_and_determinism():
    s1, d1 = O.rmat_edges(12, 16, seed=1)
    s2, d2 = O.rmat_edges(12, 16, seed=1)
    s3, d3 = O.rmat_edges(12, 16, seed=2)
    assert np.array_equal(s1, s2) and np.array_equal(d1, d2)
    assert O.edges_digest(s1, d1) != O.edges_digest(s3, d3)
    key = s1.astype(np.uint64) << np.uint64(32) | d1
    assert np.all(np.diff(key.astype(np.int64)) > 0)  # sorted, unique
    assert not np.any(s1 == d1)
    assert 0.5 * 65536 < len(s1) < 65536
    top = (s1 >> 11) & 1
    assert 0.15 < top.mean() < 0.35  # P(src top bit = 1) = c + d = 0.24 before de-duplication


def test_er_and_communities_counts():
    s, d = O.er_edges(12, 1, 6, seed=5)  # p = 1/64
    n = 4096
    mean = n * (n - 1) / 64
    assert abs(len(s) - mean) < 6 * np.sqrt(mean)
    assert not np.any(s == d)
    key = s.astype(np.uint64) << np.uint64(32) | d
    assert np.all(np.diff(key.astype(np.int64)) > 0)
    n_comm, size, stride = 4, 256, 1024
    p_num = 1288490189  # round(0.3 * 2^32)
    cs, cd = O.community_edges(8192, n_comm, size, stride, p_num, 32, seed=3)
    assert np.all(cs // stride == cd // stride) and np.all(cs % stride < size) and np.all(cd % stride < size)
    mean = n_comm * size * (size - 1) * 0.3
    assert abs(len(cs) - mean) < 6 * np.sqrt(mean)
    # p = 1/2 with one draw per word: every word is the plain draw (exact Bernoulli(1/2))
    hs, hd = O.er_edges(7, 1, 1, seed=9)
    assert 0.4 < len(hs) / (128 * 127) < 0.6


def test_sample_sources():
    a = O.sample_sources(1000, 300, 7)
    assert len(set(a.tolist())) == 300 and a.max() < 1000
    assert np.array_equal(a, O.sample_sources(1000, 300, 7))
    perm = O.sample_sources(50, 50, 1)
    assert sorted(perm.tolist()) == list(range(50))


def test_converter_round_trip():
    n = 64 * 5 + 17
    src, dst = random_graph(n, 3000)
    ns, off, row, col, data = O.edges_to_tiles(n, src, dst)
    assert ns == 6 and off[0] == 0 and off[-1] == len(row)
    keys = row.astype(np.int64) * ns + col
    assert np.all(np.diff(keys) > 0)  # row-major order, unique
    assert np.array_equal(np.diff(off), np.bincount(row, minlength=ns))
    assert edges_from_tiles(ns, off, row, col, data) == sorted(zip(src.tolist(), dst.tolist()))
    _, toff, trow, tcol, tdata = O.edges_to_tiles(n, src, dst, transpose=True)
    assert edges_from_tiles(ns, toff, trow, tcol, tdata) == sorted(zip(dst.tolist(), src.tolist()))
    assert all((data[t] != 0).any() for t in range(len(row)))
