"""The host-pointer surface in the reference's Go-API style (matrixgraph_bit_matmult,
matrixgraph_bit_walk) called through the C-ABI, against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def lib():
    import matrixgraph_amd

    return matrixgraph_amd.lib()


def pack_bits(dense):
    """r x s 0/1 matrix -> one bit string, element (i, j) at bit i*s + j, u64 words."""
    bits = np.packbits(dense.astype(np.uint8).reshape(-1), bitorder="little")
    words = np.zeros((len(bits) + 7) // 8 * 8, np.uint8)
    words[:len(bits)] = bits
    return words.view(np.uint64)


def unpack_bits(words, r, s):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[: r * s].reshape(r, s)


@pytest.mark.parametrize("m,k,n,density", [(70, 130, 45, 0.05), (64, 64, 64, 0.3), (1, 200, 3, 0.5)])
def test_bit_matmult_matches_oracle(m, k, n, density):
    rng = np.random.default_rng(m * 1000 + k)
    a = (rng.random((m, k)) < density).astype(np.uint8)
    b = (rng.random((k, n)) < density).astype(np.uint8)
    A, Bt = pack_bits(a), pack_bits(np.ascontiguousarray(b.T))
    Cw = np.full((m * n + 63) // 64, 0xFFFF, np.uint64)  # overwritten, not accumulated
    rc = lib().matrixgraph_bit_matmult(A.ctypes.data, Bt.ctypes.data, Cw.ctypes.data, m, k, n)
    assert rc == 0
    want = O.unpack_rows(O.bitset_matmul(O.pack_rows(a), O.pack_rows(b), m, k, n), n)
    assert np.array_equal(unpack_bits(Cw, m, n), want)


def test_bit_walk_matches_oracle_and_reports_errors():
    scale, hops, n_src = 10, 3, 100
    n = 1 << scale
    s, d = O.rmat_edges(scale, 8, 0.57, 0.19, 0.19, 5)
    sources = O.sample_sources(n, n_src, 5)
    used = (n_src + 63) // 64
    x = np.zeros(n * used, np.uint64)
    pc = C.c_uint64()
    rc = lib().matrixgraph_bit_walk(s.ctypes.data, d.ctypes.data, len(s), n, sources.ctypes.data, n_src, hops,
                                    x.ctypes.data, C.byref(pc))
    assert rc == 0
    want = O.walk_frontiers(n, s, d, sources, hops, W=2)[hops][:, :used]
    assert np.array_equal(x.reshape(n, used), want)
    assert pc.value == int(np.unpackbits(want.view(np.uint8)).sum())
    # errors come back as status 1 (message on stderr), never exit()
    assert lib().matrixgraph_bit_walk(s.ctypes.data, d.ctypes.data, len(s), n, sources.ctypes.data, 0, hops,
                                      x.ctypes.data, None) == 1
    dup = np.array([3, 3], np.uint32)
    assert lib().matrixgraph_bit_walk(s.ctypes.data, d.ctypes.data, len(s), n, dup.ctypes.data, 2, 1,
                                      x.ctypes.data, None) == 1
