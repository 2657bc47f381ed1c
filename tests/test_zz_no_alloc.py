"""The allocation-free MLP gradient (snc_hip_mlp_grad_noalloc) against the
allocating one (snc_hip_mlp_grad), through the C ABI.

Both run the same two kernels over the same examples in one chunk, so the
gradient and the loss agree bit for bit. Two calls in a row, the second with
fewer examples, check that what the first call left in the shared scratch
array does not leak into the second; a second model on the same GPU shares
the array; sizes beyond the stated maximum are rejected.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _examples(rng, M):
    X = rng.standard_normal((M, 16)).astype(np.float32)
    y = (rng.random(M) < 0.4).astype(np.float32)
    return X, y


def test_noalloc_gradient_equals_allocating_gradient():
    from snc_amd._lib import HipError
    from snc_amd.grow import NOALLOC_MAX_EXAMPLES, GrowMLP
    rng = np.random.default_rng(21)
    p = (rng.standard_normal(O.MLP_PARAMS) * 0.2).astype(np.float32)
    m = GrowMLP(3)
    m.params = p
    for M in (NOALLOC_MAX_EXAMPLES, 300, 1):  # each call smaller than the one before
        X, y = _examples(rng, M)
        print(f"noalloc gradient, M={M}", flush=True)
        l_ref, g_ref = m.grad(X, y)
        l_na, g_na = m.grad_noalloc(X, y)
        assert np.array_equal(g_na, g_ref), M
        assert l_na == l_ref, M
        # and both follow the FP64 oracle of the BF16 model
        l_o, g_o = O.mlp_grad(p, X, y, quantize=True)
        assert abs(l_na - l_o) <= 1e-4 * (1 + abs(l_o))
        np.testing.assert_allclose(g_na, g_o, rtol=1e-3, atol=1e-5)
    # a second model on the same GPU shares the scratch array
    m2 = GrowMLP(4)
    m2.params = p * np.float32(0.5)
    X, y = _examples(rng, 77)
    l2, g2 = m2.grad_noalloc(X, y)
    assert np.array_equal(g2, m2.grad(X, y)[1]) and l2 == m2.grad(X, y)[0]
    l1, g1 = m.grad_noalloc(X, y)
    assert np.array_equal(g1, m.grad(X, y)[1]) and l1 == m.grad(X, y)[0]
    # beyond the stated maximum: a clear error, nothing launched
    X, y = _examples(rng, NOALLOC_MAX_EXAMPLES + 1)
    with pytest.raises(HipError, match="SNC_MLP_NOALLOC_MAX_EXAMPLES"):
        m.grad_noalloc(X, y)
    m.close()
    m2.close()
