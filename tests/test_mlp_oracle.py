"""The FP64 restatement of the step-size network (oracle/mlp_oracle.py): its
hand-written gradient against central finite differences of its loss, and the
BF16 rounding against hand-worked values."""
import numpy as np

from oracle import mlp_oracle as M


def _case(seed=0, n=40):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(M.NP) * 0.3
    norm = np.concatenate([rng.standard_normal(8), 0.5 + rng.random(8), [0.3], [1.7]])
    X = rng.standard_normal((n, 8)) * 2.0
    T = rng.standard_normal(n)
    return p, norm, X, T


def test_gradient_matches_central_differences():
    p, norm, X, T = _case()
    _, g = M.loss_grad(p, norm, X, T)
    rng = np.random.default_rng(1)
    # every bias, the output layer, and a random sample of weights of each layer
    idx = np.concatenate([np.arange(M.O_B1, M.O_W2), np.arange(M.O_B2, M.NP),
                          rng.choice(np.arange(M.O_W1, M.O_B1), 40, replace=False),
                          rng.choice(np.arange(M.O_W2, M.O_B2), 80, replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, norm, X, T) - M.loss(pm, norm, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_gradient_sign_error_is_caught():
    """A planted mistake (the output-layer gradient with the wrong sign) fails the
    same finite-difference comparison."""
    p, norm, X, T = _case(seed=2)
    _, g = M.loss_grad(p, norm, X, T)
    i = M.O_W3 + 5
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (M.loss(pp, norm, X, T) - M.loss(pm, norm, X, T)) / (2 * h)
    assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i]))
    assert abs(fd + g[i]) > 1e-6 * max(1.0, abs(g[i]))


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1.0)
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.140625], np.float32)
    out = M.bf16(vals)
    np.testing.assert_array_equal(out, [1.0, 1.0, 1.0 + 2 * 2.0 ** -7, -2.5, 3.140625])


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 3.0])
    # GELU(0) = 0, GELU(z) - GELU(-z) = z, GELU'(0) = 1/2
    assert M.gelu(z)[0] == 0.0
    np.testing.assert_allclose(M.gelu(z[1:2]) - M.gelu(-z[1:2]), z[1:2], rtol=1e-15)
    assert M.gelu_d(np.array([0.0]))[0] == 0.5
