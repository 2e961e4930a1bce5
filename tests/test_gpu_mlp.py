"""Device kernels of the step-size proposal network (every product on the FP32
matrix cores, the gradient summed over workgroups in a fixed order) against the
FP64 numpy restatement (oracle/mlp_oracle.py, itself pinned by finite
differences) evaluated at the device's BF16-rounded weights.  Tolerances cover
FP32 accumulation: 1e-5 relative on outputs and the loss, 1e-4 on the gradient."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _case(seed=0, n=3000):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(M.NP) * 0.25).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(8), 0.5 + rng.random(8), [0.3], [1.7]]).astype(np.float32)
    X = (rng.standard_normal((n, 8)) * 2.0).astype(np.float32)
    T = rng.standard_normal(n).astype(np.float32)
    return p, norm, X, T


def _mlp():
    from demystification_dfn_amd.device import DeviceMlp
    print("[gpu] mlp", flush=True)
    return DeviceMlp()


def test_forward_loss_and_gradient_match_fp64_oracle():
    p, norm, X, T = _case()
    m = _mlp()
    m.set_params(p, norm)
    pb = M.bf16(p)
    y_or, _ = M.forward(pb, norm, X)
    y_or = norm[16] + norm[17] * y_or
    y_dev = m.forward(X).astype(np.float64)
    assert np.max(np.abs(y_dev - y_or)) <= 1e-5 * np.max(np.abs(y_or))
    l_dev, g_dev = m.loss_grad(X, T)
    l_or, g_or = M.loss_grad(pb, norm, X, T)
    assert abs(l_dev - l_or) <= 1e-5 * l_or
    g_dev = g_dev.astype(np.float64)
    for a, b in ((M.O_W1, M.O_B1), (M.O_B1, M.O_W2), (M.O_W2, M.O_B2), (M.O_B2, M.O_W3), (M.O_W3, M.NP)):
        assert np.linalg.norm(g_dev[a:b] - g_or[a:b]) <= 1e-4 * np.linalg.norm(g_or[a:b]), (a, b)


def test_training_fits_a_smooth_target_and_is_reproducible():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((16384, 8)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    norm = np.concatenate([X.mean(0), X.std(0), [T.mean()], [T.std()]]).astype(np.float32)
    p0, _ = _mlp().get_params()
    runs = []
    for _ in range(2):
        m = _mlp()
        m.set_params(p0, norm)
        l0, _ = m.loss_grad(X, T)
        lf = m.train(X, T, epochs=15, batch=512, lr=2e-3, seed=9)
        runs.append((l0, lf, m.get_params()[0]))
    (l0, lf, pa), (_, lf2, pb) = runs
    assert lf < 0.1 * l0, (l0, lf)
    assert lf == lf2
    np.testing.assert_array_equal(pa, pb)
