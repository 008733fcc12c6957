"""GPU kernels of the learned closure (kmm_mlp_*) against the oracle
(oracle/mlp_oracle.py): FP32 gradients vs analytic FP64 gradients, bf16
gradients vs the oracle's bf16 mode, matrix-core inference vs the oracle's
bf16 forward, and Adam training on a smooth regression target."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

kmm = pytest.importorskip("polymer_dns_amd.kmm")


def setup_net(seed=5, nscale=True):
    net = kmm.Mlp(seed=seed)
    rng = np.random.default_rng(seed)
    mx = rng.standard_normal(22).astype(np.float32) * 0.1
    sx = (1.0 + rng.uniform(0, 1, 22)).astype(np.float32)
    my = rng.standard_normal(6).astype(np.float32) * 0.1
    sy = (1.0 + rng.uniform(0, 1, 6)).astype(np.float32)
    if nscale:
        net.set_norm(mx, sx, my, sy)
    else:
        mx, sx, my, sy = np.zeros(22, np.float32), np.ones(22, np.float32), np.zeros(6, np.float32), np.ones(6, np.float32)
    # non-zero biases so that every parameter group is exercised
    p = net.params()
    W, b = net.unpack(p)
    for l in range(5):
        b[l][:] = 0.05 * rng.standard_normal(b[l].shape)
    net.set_params(np.concatenate([np.concatenate([w.ravel(), bb]) for w, bb in zip(W, b)]))
    return net, (mx, sx, my, sy)


def flat(gW, gb):
    return np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(gW, gb)])


def test_fp32_gradients_match_oracle():
    net, nm = setup_net()
    rng = np.random.default_rng(1)
    X = rng.standard_normal((300, 22)).astype(np.float32)
    T = rng.standard_normal((300, 6)).astype(np.float32)
    loss, g = net.grad(X, T, mode=0)
    W, b = net.unpack()
    W64 = [w.astype(np.float64) for w in W]
    b64 = [x.astype(np.float64) for x in b]
    l_ref, gW, gb = M.loss_and_grads(W64, b64, X.astype(np.float64), T.astype(np.float64),
                                     *[x.astype(np.float64) for x in nm], mode="fp64")
    gr = flat(gW, gb)
    assert abs(loss - l_ref) < 1e-5 * abs(l_ref)
    assert np.abs(g - gr).max() < 1e-4 * np.abs(gr).max()
    net.close()


def test_bf16_gradients_and_inference_match_oracle():
    net, nm = setup_net(seed=6)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((1000, 22)).astype(np.float32)
    T = rng.standard_normal((1000, 6)).astype(np.float32)
    W, b = net.unpack()
    loss, g = net.grad(X, T, mode=1)
    l_ref, gW, gb = M.loss_and_grads(W, b, X, T, *nm, mode="bf16")
    gr = flat(gW, gb)
    assert abs(loss - l_ref) < 1e-3 * abs(l_ref)
    assert np.linalg.norm(g - gr) < 2e-2 * np.linalg.norm(gr)
    # matrix-core inference vs the bf16 forward (rounding points identical;
    # FP32 sums in another order can flip single bf16 roundings)
    Y = net.infer(X)
    Yr = M.forward(W, b, X, *nm, mode="bf16")
    err = np.abs(Y - Yr) / np.abs(Yr).max()
    assert np.median(err) < 1e-3 and err.max() < 3e-2
    # odd sizes: partial tiles
    Y7 = net.infer(X[:7])
    assert np.array_equal(Y7, Y[:7])
    net.close()


def test_training_reduces_loss():
    net = kmm.Mlp(seed=9)
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (4096, 22)).astype(np.float32)
    A = rng.standard_normal((22, 6)).astype(np.float32) / np.sqrt(22)
    T = (np.tanh(X @ A) + 0.1 * (X[:, :6] ** 2)).astype(np.float32)
    l0, _ = net.grad(X, T, mode=1)
    losses = net.train(X, T, steps=400, batch=256, lr=2e-3, seed=4)
    l1, _ = net.grad(X, T, mode=1)
    assert np.isfinite(losses).all()
    assert l1 < 0.1 * l0, (l0, l1)
    net.close()


@pytest.mark.parametrize("kind", ["gelu", "gelu_poly"])
def test_gelu_gradients_and_inference_match_oracle(kind):
    """The GELU networks (x sigma(1.702 x); x s(t) with the polynomial
    sigmoid): FP32 gradients vs the oracle's FP64 gradients, bf16 gradients
    and matrix-core inference vs the oracle's bf16 mode."""
    net, nm = setup_net(seed=16)
    net.set_activation(kind)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((600, 22)).astype(np.float32)
    T = rng.standard_normal((600, 6)).astype(np.float32)
    W, b = net.unpack()
    M.ACT = kind
    try:
        loss, g = net.grad(X, T, mode=0)
        l_ref, gW, gb = M.loss_and_grads([w.astype(np.float64) for w in W], [x.astype(np.float64) for x in b],
                                         X.astype(np.float64), T.astype(np.float64),
                                         *[x.astype(np.float64) for x in nm], mode="fp64")
        gr = flat(gW, gb)
        assert abs(loss - l_ref) < 1e-5 * abs(l_ref)
        assert np.abs(g - gr).max() < 1e-4 * np.abs(gr).max()
        loss, g = net.grad(X, T, mode=1)
        l_ref, gW, gb = M.loss_and_grads(W, b, X, T, *nm, mode="bf16")
        gr = flat(gW, gb)
        assert abs(loss - l_ref) < 1e-3 * abs(l_ref)
        assert np.linalg.norm(g - gr) < 2e-2 * np.linalg.norm(gr)
        Y = net.infer(X)
        Yr = M.forward(W, b, X, *nm, mode="bf16")
        err = np.abs(Y - Yr) / np.abs(Yr).max()
        assert np.median(err) < 1e-3 and err.max() < 3e-2
    finally:
        M.ACT = "relu"
    net.close()
