"""The option selector's network.

CPU (the oracle pinned): the numpy restatement's gradient equals central
finite differences of its loss on a small network, masked outputs get no
gradient, and Adam's first step moves every parameter by lr against the
sign of its gradient.

GPU (the product against the oracle, through the C ABI): forward outputs
within 1e-5 relative, every parameter's gradient within 1e-4 relative to
the largest, and 50 training steps that follow the oracle's Adam
trajectory (losses within 1e-3 relative) and decrease the loss; training
is deterministic (two runs, the same parameters bit for bit).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _data(B, n_in, n_out, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    M = (rng.random((B, n_out)) < 0.7).astype(np.float64)
    return X, T, M


def _init(n_in, h, n_out, seed):
    lay, total = MO.layout(n_in, h, n_out)
    rng = np.random.default_rng(seed)
    p = np.zeros(total)
    for (w, I, J, b) in lay:
        p[w:w + I * J] = rng.standard_normal(I * J) * np.sqrt(2.0 / I)
        p[b:b + J] = rng.standard_normal(J) * 0.1
    return p


def test_oracle_gradient_matches_finite_differences():
    n_in, h, n_out = 5, 7, 4
    X, T, M = _data(9, n_in, n_out, 1)
    p = _init(n_in, h, n_out, 2)
    loss, g = MO.loss_and_grad(p, X, T, M, h, n_out)
    eps = 1e-6
    num = np.empty_like(p)
    for i in range(p.size):
        q = p.copy(); q[i] += eps
        lp, _ = MO.loss_and_grad(q, X, T, M, h, n_out)
        q[i] -= 2 * eps
        lm, _ = MO.loss_and_grad(q, X, T, M, h, n_out)
        num[i] = (lp - lm) / (2 * eps)
    np.testing.assert_allclose(g, num, rtol=1e-5, atol=1e-7)


def test_oracle_masked_outputs_have_no_gradient():
    n_in, h, n_out = 4, 6, 3
    X, T, M = _data(8, n_in, n_out, 3)
    M[:, 1] = 0.0
    p = _init(n_in, h, n_out, 4)
    lay, _ = MO.layout(n_in, h, n_out)
    _, g = MO.loss_and_grad(p, X, T, M, h, n_out)
    w, I, J, b = lay[3]
    assert np.all(g[w:w + I * J].reshape(I, J)[:, 1] == 0) and g[b + 1] == 0
    # and changing the masked target changes nothing
    T2 = T.copy(); T2[:, 1] += 10
    l1, _ = MO.loss_and_grad(p, X, T, M, h, n_out)
    l2, _ = MO.loss_and_grad(p, X, T2, M, h, n_out)
    assert l1 == l2


def test_oracle_adam_first_step():
    p = np.array([1.0, -2.0, 0.5]); g = np.array([0.3, -4.0, 1e-3])
    q, m, v = MO.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(q, p - 0.01 * np.sign(g), rtol=0, atol=1e-6)


torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sap():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import saplibrary_amd
    saplibrary_amd.lib()
    return saplibrary_amd


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(sap):
    n_in, h, n_out, B = 24, 64, 30, 200
    X, T, M = _data(B, n_in, n_out, 11)
    net = sap.Mlp(n_in, n_out, h, seed=5)
    p = net.params().astype(np.float64)
    Y = net.forward(_dev(X)).cpu().numpy()
    Yo, _, _ = MO.forward(p, X.astype(np.float32), h, n_out)
    np.testing.assert_allclose(Y, Yo, rtol=0, atol=1e-5 * np.abs(Yo).max())
    g, loss = net.gradient(_dev(X), _dev(T), _dev(M))
    lo, go = MO.loss_and_grad(p, X.astype(np.float32), T.astype(np.float32), M, h, n_out)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-4 * np.abs(go).max())


@pytest.mark.gpu
def test_gpu_training_follows_oracle_adam(sap):
    n_in, h, n_out, B = 10, 32, 6, 120
    X, T, M = _data(B, n_in, n_out, 21)
    net = sap.Mlp(n_in, n_out, h, seed=7)
    p = net.params().astype(np.float64)
    losses = net.train(_dev(X), _dev(T), _dev(M), 50, lr=1e-3)
    m = np.zeros_like(p); v = np.zeros_like(p)
    ref = []
    Xf, Tf = X.astype(np.float32), T.astype(np.float32)
    for t in range(1, 51):
        lo, go = MO.loss_and_grad(p, Xf, Tf, M, h, n_out)
        ref.append(lo)
        p, m, v = MO.adam(p, go, m, v, t, lr=1e-3)
    np.testing.assert_allclose(losses, ref, rtol=1e-3)
    assert losses[-1] < 0.8 * losses[0]
    np.testing.assert_allclose(net.params(), p, rtol=0, atol=1e-3 * np.abs(p).max())
    net2 = sap.Mlp(n_in, n_out, h, seed=7)
    net2.train(_dev(X), _dev(T), _dev(M), 50, lr=1e-3)
    assert np.array_equal(net.params(), net2.params())
