"""Grow-proposal MLP: the FP64 oracle pinned by finite differences; the GPU
forward (MFMA) and backward kernels checked against it."""
import numpy as np
import pytest

import oracle as O


def _problem(seed=0, M=64):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(O.MLP_PARAMS) * 0.2).astype(np.float32)
    X = rng.standard_normal((M, 16)).astype(np.float32)
    y = (rng.random(M) < 0.5).astype(np.float32)
    return p, X, y


def test_oracle_gradient_matches_finite_differences():
    p, X, y = _problem(1)
    _, g = O.mlp_grad(p, X, y, quantize=False)
    rng = np.random.default_rng(2)
    # central differences with steps that are exact in float32 around p
    for i in rng.choice(O.MLP_PARAMS, 40, replace=False):
        h = np.float32(2.0 ** -16)
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (O.mlp_grad(pp, X, y, False)[0] - O.mlp_grad(pm, X, y, False)[0]) / (
            float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) <= 1e-5 + 1e-3 * abs(g[i]), (i, fd, g[i])


def test_oracle_forward_is_the_network():
    p, X, _ = _problem(3, 5)
    out = O.mlp_forward(p, X, quantize=False)
    W1 = p[:1024].reshape(64, 16).astype(np.float64)
    b1 = p[1024:1088].astype(np.float64)
    W2 = p[1088:5184].reshape(64, 64).astype(np.float64)
    b2 = p[5184:5248].astype(np.float64)
    W3 = p[5248:5312].astype(np.float64)
    b3 = float(p[5312])
    h1 = np.maximum(X.astype(np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    np.testing.assert_allclose(out, h2 @ W3 + b3, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_gpu_scores_match_oracle_bf16_model():
    from snc_amd.grow import GrowMLP
    p, _, _ = _problem(4)
    rng = np.random.default_rng(5)
    for M in (1, 15, 16, 17, 1000, 4099):
        X = rng.standard_normal((M, 16)).astype(np.float32)
        m = GrowMLP(7)
        m.params = p
        got = m.score(X).astype(np.float64)
        want = O.mlp_forward(p, X, quantize=True)
        err = np.abs(got - want)
        # FP32 MFMA accumulation vs FP64; BF16 rounding of the hidden layer
        # may land one BF16 ulp apart.
        assert err.max() <= 2e-2 * (1 + np.abs(want).max()), (M, err.max())
        assert err.mean() <= 2e-3 * (1 + np.abs(want).mean()), (M, err.mean())
        m.close()


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    from snc_amd.grow import GrowMLP
    p, X, y = _problem(6, 300)
    m = GrowMLP(1)
    m.params = p
    loss, g = m.grad(X, y)
    lo, go = O.mlp_grad(p, X, y, quantize=True)
    assert abs(loss - lo) <= 1e-4 * (1 + abs(lo))
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-5)


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    from snc_amd.grow import GrowMLP
    rng = np.random.default_rng(8)
    X = rng.standard_normal((2048, 16)).astype(np.float32)
    y = ((X[:, 0] + 0.5 * X[:, 3] - X[:, 7]) > 0).astype(np.float32)
    m = GrowMLP(2)
    l0, _ = m.grad(X, y)
    m.train(X, y, epochs=30, batch=128, lr=3e-3, seed=3)
    l1, _ = m.grad(X, y)
    assert l1 < 0.5 * l0
    acc = ((m.score(X) > 0) == (y > 0.5)).mean()
    assert acc > 0.9
