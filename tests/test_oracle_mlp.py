"""Pins of the MLP oracle (no GPU): FP64 gradients against central finite differences,
and the BF16 round-to-nearest-even conversion against bit-pattern cases."""
import struct

import numpy as np
import pytest

import oracle as O


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def bits_f32(u):
    return struct.unpack("<f", struct.pack("<I", u))[0]


@pytest.mark.parametrize("u,expect", [
    (0x3F800000, 0x3F80),  # 1.0 exact
    (0x3F808000, 0x3F80),  # tie, lower kept bit even -> down
    (0x3F818000, 0x3F82),  # tie, lower kept bit odd -> up
    (0x3F808001, 0x3F81),  # above the tie -> up
    (0x3F807FFF, 0x3F80),  # below the tie -> down
    (0xBF818000, 0xBF82),  # negative tie -> magnitude up (to even)
    (0x7F7FFFFF, 0x7F80),  # largest finite float rounds to +inf
    (0x7F800000, 0x7F80),  # +inf
    (0x00000001, 0x0000),  # smallest subnormal rounds to +0
    (0x00018000, 0x0002),  # subnormal tie -> even
])
def test_bf16_rounding_bit_cases(u, expect):
    assert O.f32_to_bf16(bits_f32(u)) == expect


def test_bf16_nan_and_round_trip():
    assert (O.f32_to_bf16(float("nan")) & 0x7FC0) == 0x7FC0
    rng = np.random.default_rng(3)
    for h in rng.integers(0, 0x7F80, 200):
        h = int(h)
        assert O.f32_to_bf16(O.bf16_to_f32(h)) == h  # exact BF16 values are fixed points


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    n_in, n_out, n = 8, 2, 40
    npar = O.mlp_n_params(n_in, n_out)
    w = rng.normal(0, 0.3, npar)
    x = rng.normal(0, 1, (n, n_in)).astype(np.float32)
    t = rng.normal(0, 1, (n, n_out)).astype(np.float32)
    loss, grad = O.mlp_loss_grad(n_in, n_out, w, x, t)
    # mean squared error, directly
    assert loss > 0
    idx = rng.choice(npar, 60, replace=False)
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (O.mlp_loss_grad(n_in, n_out, wp, x, t)[0] - O.mlp_loss_grad(n_in, n_out, wm, x, t)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_mlp_oracle_loss_closed_form():
    # all weights zero except the output bias -> y = b3 for every sample
    n_in, n_out = 4, 2
    w = np.zeros(O.mlp_n_params(n_in, n_out))
    w[-2:] = [0.5, -1.0]
    x = np.ones((3, n_in), np.float32)
    t = np.array([[1, 1], [0, 0], [2, -1]], np.float32)
    loss, grad = O.mlp_loss_grad(n_in, n_out, w, x, t)
    y = np.array([0.5, -1.0])
    assert loss == pytest.approx(float(((y - t) ** 2).mean()))
    assert grad[-2:] == pytest.approx(list((2 * (y - t)).mean(axis=0) / 2))
