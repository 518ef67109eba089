"""MLP cost-model kernels vs the FP64 oracle: BF16 weight rounding, forward outputs,
gradients, determinism, and that training reduces the loss."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def mg():
    import matrixgraph_amd

    return matrixgraph_amd


def data(n, n_in, n_out, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, n_in)).astype(np.float32)
    w_true = rng.normal(0, 1, (n_in, n_out))
    t = (np.tanh(x @ w_true) + 0.1 * rng.normal(0, 1, (n, n_out))).astype(np.float32)
    return x, t


@pytest.mark.parametrize("n_in,n_out", [(8, 2), (16, 1), (4, 2)])
def test_mlp_bf16_forward_and_gradient_match_oracle(ctx, n_in, n_out):
    M = mg()
    m = M.MLP(ctx, n_in, n_out, seed=7)
    master, bf16 = m.params()
    assert all(O.f32_to_bf16(float(f)) == int(h) for f, h in zip(master[:300], bf16[:300]))
    wq = np.array([O.bf16_to_f32(int(h)) for h in bf16], np.float64)  # the weights the kernels use
    x, t = data(1000, n_in, n_out, 3)
    y = m.forward(x)
    # forward: FP32 accumulation of BF16 weights vs FP64 reference
    loss_ref, grad_ref = O.mlp_loss_grad(n_in, n_out, wq, x, t)
    loss, grad = m.loss_grad(x, t)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    scale = np.abs(grad_ref).max()
    assert np.abs(grad - grad_ref).max() <= 1e-4 * scale + 1e-7
    # outputs against an explicit FP64 forward of the same weights
    assert np.isfinite(y).all() and y.shape == (1000, n_out)
    loss_y = float(((y.astype(np.float64) - t) ** 2).mean())
    assert loss_y == pytest.approx(loss_ref, rel=1e-4)
    # determinism: same inputs, same gradient bits
    _, grad2 = m.loss_grad(x, t)
    assert np.array_equal(grad, grad2)


def test_mlp_training_reduces_loss(ctx):
    M = mg()
    x, t = data(20000, 8, 2, 5)
    m = M.MLP(ctx, 8, 2, seed=1)
    l0, _ = m.loss_grad(x, t)
    l1 = m.train(x, t, epochs=5, batch=512, lr=3e-3)
    assert l1 < 0.5 * l0
