"""Pins for the MLP-closure oracle (oracle/mlp_oracle.py): bfloat16 rounding
against its bit-level definition, the activation (closed form and derivative) and the analytic
gradients against central finite differences."""
import math

import numpy as np

from oracle import mlp_oracle as M


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = 2.0 ** -7                       # bf16 spacing at 1
    x = np.array([1.0, 1.0 + ulp / 2, 1.0 + 3 * ulp / 2, 1.0 + ulp / 4, -2.5, 3.0e38], dtype=np.float32)
    got = M.bf16(x)
    # ties go to the even mantissa: 1 + ulp/2 -> 1, 1 + 3ulp/2 -> 1 + 2ulp
    exp = np.array([1.0, 1.0, 1.0 + 2 * ulp, 1.0, -2.5, 3.0e38], dtype=np.float32)
    assert np.array_equal(got[:5], exp[:5])
    assert (got.view(np.uint32) & 0xFFFF).max() == 0
    assert abs(got[5] / exp[5] - 1) < 2 ** -8
    assert M.bf16(one) == one


def test_activation_and_derivative():
    x = np.array([-3.0, -0.5, 0.0, 0.5, 2.5])
    assert np.array_equal(M.act(x), [0.0, 0.0, 0.0, 0.5, 2.5])
    # bf16 rounding and the activation commute (the matrix-core path rounds first)
    z = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    assert np.array_equal(M.bf16(M.act(z)), M.act(M.bf16(z)))
    xs = np.linspace(-3, 3, 120)          # no sample on the kink
    h = 1e-6
    fd = (M.act(xs + h) - M.act(xs - h)) / (2 * h)
    assert np.abs(fd - M.act_grad(xs)).max() < 1e-8


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    W, b = M.init_params(7)
    W = [w.astype(np.float64) for w in W]
    b = [bb.astype(np.float64) + 0.01 * rng.standard_normal(bb.shape) for bb in b]
    X = rng.standard_normal((5, 22))
    T = rng.standard_normal((5, 6))
    mx, sx, my, sy = np.zeros(22), np.ones(22), np.zeros(6), np.ones(6)
    loss, gW, gb = M.loss_and_grads(W, b, X, T, mx, sx, my, sy, mode="fp64")
    h = 1e-6
    for l, (i, j) in [(0, (3, 5)), (1, (10, 100)), (2, (0, 0)), (3, (127, 64)), (4, (5, 17))]:
        Wp = [w.copy() for w in W]
        Wm = [w.copy() for w in W]
        Wp[l][i, j] += h
        Wm[l][i, j] -= h
        fd = (M.loss_and_grads(Wp, b, X, T, mx, sx, my, sy, "fp64")[0]
              - M.loss_and_grads(Wm, b, X, T, mx, sx, my, sy, "fp64")[0]) / (2 * h)
        assert abs(fd - gW[l][i, j]) < 1e-7 * max(1.0, abs(fd)), (l, i, j, fd, gW[l][i, j])
    for l, i in [(0, 7), (4, 2)]:
        bp = [x.copy() for x in b]
        bm = [x.copy() for x in b]
        bp[l][i] += h
        bm[l][i] -= h
        fd = (M.loss_and_grads(W, bp, X, T, mx, sx, my, sy, "fp64")[0]
              - M.loss_and_grads(W, bm, X, T, mx, sx, my, sy, "fp64")[0]) / (2 * h)
        assert abs(fd - gb[l][i]) < 1e-7 * max(1.0, abs(fd))


def test_gelu_sigmoid_form_against_the_exact_gelu_and_its_derivative():
    """GELU as x sigma(1.702 x): within 0.021 of x Phi(x) (the exact GELU,
    Phi the standard normal CDF), 0 at 0, -> x for large x, -> 0 for large
    -x; the derivative against central differences; the gradients of the
    whole network against finite differences with GELU."""
    x = np.linspace(-8, 8, 4001)
    phi = 0.5 * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in x]))
    g = M.act(x, "gelu")
    assert np.abs(g - x * phi).max() < 0.021
    assert M.act(np.array([0.0]), "gelu")[0] == 0.0
    assert abs(M.act(np.array([30.0]), "gelu")[0] - 30.0) < 1e-12
    assert abs(M.act(np.array([-30.0]), "gelu")[0]) < 1e-12
    h = 1e-6
    fd = (M.act(x + h, "gelu") - M.act(x - h, "gelu")) / (2 * h)
    assert np.abs(fd - M.act_grad(x, "gelu")).max() < 1e-8
    rng = np.random.default_rng(4)
    W, b = M.init_params(8)
    W = [w.astype(np.float64) for w in W]
    b = [bb.astype(np.float64) + 0.01 * rng.standard_normal(bb.shape) for bb in b]
    X = rng.standard_normal((5, 22))
    T = rng.standard_normal((5, 6))
    nm = (np.zeros(22), np.ones(22), np.zeros(6), np.ones(6))
    M.ACT = "gelu"
    try:
        loss, gW, gb = M.loss_and_grads(W, b, X, T, *nm, mode="fp64")
        for l, (i, j) in [(0, (3, 5)), (2, (1, 2)), (4, (5, 17))]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (M.loss_and_grads(Wp, b, X, T, *nm, "fp64")[0] - M.loss_and_grads(Wm, b, X, T, *nm, "fp64")[0]) / (2 * h)
            assert abs(fd - gW[l][i, j]) < 1e-7 * max(1.0, abs(fd)), (l, i, j)
    finally:
        M.ACT = "relu"


def test_gelu_polynomial_sigmoid_against_the_exact_gelu_and_its_derivative():
    """GELU as x s(t), t = clamp(1/2 + x/5.28, 0, 1), s the quintic smoothstep:
    within 0.0195 of x Phi(x) (closer than the sigmoid form), exactly x for
    x >= 2.64 and exactly 0 for x <= -2.64, 0 and slope 1/2 at 0 (as x Phi(x));
    the derivative against central differences (continuous at the clamps,
    where s' = 0); the whole network's gradients against finite differences."""
    x = np.linspace(-8, 8, 4001)
    phi = 0.5 * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in x]))
    g = M.act(x, "gelu_poly")
    assert np.abs(g - x * phi).max() < 0.0195
    assert np.abs(g - x * phi).max() < np.abs(M.act(x, "gelu") - x * phi).max()
    big = x[x >= 2.64]
    assert np.array_equal(M.act(big, "gelu_poly"), big)
    assert not np.any(M.act(x[x <= -2.64], "gelu_poly"))
    assert M.act(np.array([0.0]), "gelu_poly")[0] == 0.0
    assert abs(M.act_grad(np.array([0.0]), "gelu_poly")[0] - 0.5) < 1e-15
    h = 1e-6
    xs = x[np.abs(np.abs(x) - 2.64) > 1e-3]
    fd = (M.act(xs + h, "gelu_poly") - M.act(xs - h, "gelu_poly")) / (2 * h)
    assert np.abs(fd - M.act_grad(xs, "gelu_poly")).max() < 1e-7
    rng = np.random.default_rng(5)
    W, b = M.init_params(9)
    W = [w.astype(np.float64) for w in W]
    b = [bb.astype(np.float64) + 0.01 * rng.standard_normal(bb.shape) for bb in b]
    X = rng.standard_normal((5, 22))
    T = rng.standard_normal((5, 6))
    nm = (np.zeros(22), np.ones(22), np.zeros(6), np.ones(6))
    M.ACT = "gelu_poly"
    try:
        loss, gW, gb = M.loss_and_grads(W, b, X, T, *nm, mode="fp64")
        for l, (i, j) in [(0, (3, 5)), (2, (1, 2)), (4, (5, 17))]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (M.loss_and_grads(Wp, b, X, T, *nm, "fp64")[0] - M.loss_and_grads(Wm, b, X, T, *nm, "fp64")[0]) / (2 * h)
            assert abs(fd - gW[l][i, j]) < 1e-7 * max(1.0, abs(fd)), (l, i, j)
    finally:
        M.ACT = "relu"
