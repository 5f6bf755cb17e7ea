"""Pins of the learned-start network's oracle: bfloat16 rounding against torch's own
float32 -> bfloat16 conversion, the backward pass against central finite differences,
and the Adam step against the paper's algorithm."""
import math

import numpy as np

from oracle.mlp_oracle import MlpOracle, adam_step, bf16_round, gelu, gelu_grad


def test_bf16_round_matches_torch():
    import torch
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        # exact ties between two bfloat16 neighbours (even and odd upper halves)
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -(1.0 + 2 ** -8), 0.0, -0.0], np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(bf16_round(x), ref)


def test_gelu_derivative():
    x = np.linspace(-6, 6, 101)
    h = 1e-6
    np.testing.assert_allclose(gelu_grad(x), (gelu(x + h) - gelu(x - h)) / (2 * h), atol=1e-8)
    # GELU(x) = x Phi(x): Phi(0) = 1/2, GELU(-inf) -> 0
    assert gelu(np.array([0.0]))[0] == 0.0 and abs(gelu(np.array([-40.0]))[0]) < 1e-300


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    dims = [7, 11, 9, 5]
    n = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(3))
    flat = rng.standard_normal(n) * 0.5
    X = rng.standard_normal((6, 7))
    T = rng.standard_normal((6, 5))
    net = MlpOracle(dims, flat)
    loss, g = net.grads_flat(X, T)
    eps = 1e-6
    for i in rng.choice(n, 40, replace=False):
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp = MlpOracle(dims, fp).grads_flat(X, T)[0]
        lm = MlpOracle(dims, fm).grads_flat(X, T)[0]
        assert abs((lp - lm) / (2 * eps) - g[i]) < 1e-7 * max(1.0, abs(g[i]))


def test_adam_matches_algorithm():
    """Kingma & Ba, Algorithm 1 with the efficient step size of their section 2:
    alpha_t = alpha sqrt(1 - b2^t) / (1 - b1^t), eps applied to sqrt(v_t)."""
    rng = np.random.default_rng(9)
    w = rng.standard_normal(10)
    m = np.zeros(10)
    v = np.zeros(10)
    w_ref, m_ref, v_ref = w.copy(), m.copy(), v.copy()
    for t in range(1, 6):
        g = rng.standard_normal(10)
        w, m, v = adam_step(w, g, m, v, t, 1e-2)
        m_ref = 0.9 * m_ref + 0.1 * g
        v_ref = 0.999 * v_ref + 0.001 * g * g
        mh = m_ref / (1 - 0.9 ** t)
        vh = v_ref / (1 - 0.999 ** t)
        w_ref = w_ref - 1e-2 * mh / (np.sqrt(vh) + 1e-8 / math.sqrt(1 - 0.999 ** t))
    np.testing.assert_allclose(w, w_ref, rtol=1e-12, atol=1e-14)
