"""CPU restatement (numpy) of the learned polymer-stress closure: a 22-input,
4 x 128 ReLU or GELU, 6-output multilayer perceptron (no reference counterpart; the
network is new here).

TEST INFRASTRUCTURE ONLY: only tests/ may import this module.

Conventions shared with polymer_dns_amd/csrc/kmm_mlp.hip:
  * parameters: W[l] (n_out x n_in, row-major), b[l] (n_out); layer sizes
    22 -> 128 -> 128 -> 128 -> 128 -> 6; ReLU (or GELU, module ACT) after
    the four hidden layers, none after the last;
  * input and output affine normalisation: x_n = (x - mu_x)/sd_x,
    y = y_n sd_y + mu_y;
  * "bf16" mode rounds weights and the input of every layer to bfloat16
    (round to nearest even), accumulates in FP32;
  * loss = mean over samples of sum over the 6 outputs of
    (y_n - t_n)^2 with t_n = (t - mu_y)/sd_y.

Pins: analytic gradients vs central finite differences (FP64); bf16
rounding vs the bit pattern definition; the activation's values and
derivative vs differences away from the kink.
"""
from __future__ import annotations

import math

import numpy as np

SIZES = (22, 128, 128, 128, 128, 6)
NL = len(SIZES) - 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even) and return float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a)


GELU_A = 1.702   # GELU in the sigmoid form x sigma(1.702 x) (Hendrycks & Gimpel 2016, section 2)
GELU_C = 2.64    # GELU with the polynomial sigmoid: t = clamp(1/2 + x/(2 c), 0, 1)
ACT = "relu"     # hidden-layer activation of the functions below: "relu", "gelu" or "gelu_poly"


def _poly_t(x):
    return np.clip(0.5 + x / (2.0 * GELU_C), 0.0, 1.0)


def act(x, kind=None):
    """ReLU; GELU as x sigma(1.702 x); or GELU as x s(t) with the quintic
    smoothstep s(t) = t^3 (t (6 t - 15) + 10) of t = clamp(1/2 + x/(2 c), 0, 1)."""
    kind = kind or ACT
    if kind == "gelu":
        x = np.asarray(x)
        return x / (1.0 + np.exp(-GELU_A * x))
    if kind == "gelu_poly":
        x = np.asarray(x)
        t = _poly_t(x)
        return x * (t ** 3 * (t * (6.0 * t - 15.0) + 10.0))
    return np.maximum(x, 0.0)


def act_grad(x, kind=None):
    """d act / dx (ReLU: 0 at the kink, as the GPU training kernel takes it;
    GELU: sigma + 1.702 x sigma (1 - sigma); polynomial GELU:
    s(t) + x s'(t) / (2 c), s'(t) = 30 t^2 (t - 1)^2)."""
    kind = kind or ACT
    x = np.asarray(x)
    if kind == "gelu":
        s = 1.0 / (1.0 + np.exp(-GELU_A * x))
        return s + GELU_A * x * s * (1.0 - s)
    if kind == "gelu_poly":
        t = _poly_t(x)
        return t ** 3 * (t * (6.0 * t - 15.0) + 10.0) + x / (2.0 * GELU_C) * 30.0 * (t * (t - 1.0)) ** 2
    return (x > 0).astype(x.dtype)


def init_params(seed):
    """He-style uniform initialisation from a seed (same recipe as the host
    library: U(-a, a), a = sqrt(6/n_in), zero biases)."""
    rng = np.random.default_rng(seed)
    W, b = [], []
    for l in range(NL):
        a = math.sqrt(6.0 / SIZES[l])
        W.append(rng.uniform(-a, a, (SIZES[l + 1], SIZES[l])).astype(np.float32))
        b.append(np.zeros(SIZES[l + 1], dtype=np.float32))
    return W, b


def forward(W, b, X, mu_x, sd_x, mu_y, sd_y, mode="bf16", keep=False):
    """Returns y (N x 6) and, with keep, the list of layer inputs and
    pre-activations (for the backward pass)."""
    dt = np.float64 if mode == "fp64" else np.float32
    a = ((np.asarray(X, dtype=dt) - mu_x) / sd_x).astype(dt)
    acts, pres = [], []
    for l in range(NL):
        if mode == "bf16":
            a_in, w = bf16(a), bf16(W[l])
            z = (a_in.astype(np.float64) @ w.astype(np.float64).T).astype(np.float32) + b[l]
        else:
            a_in, w = a, np.asarray(W[l], dtype=dt)
            z = a_in @ w.T + np.asarray(b[l], dtype=dt)
        acts.append(a_in)
        pres.append(z)
        a = act(z).astype(z.dtype) if l < NL - 1 else z
    y = a * sd_y + mu_y
    return (y, acts, pres) if keep else y


def loss_and_grads(W, b, X, T, mu_x, sd_x, mu_y, sd_y, mode="fp64"):
    """MSE loss on normalised outputs and its gradients w.r.t. W and b."""
    y, acts, pres = forward(W, b, X, mu_x, sd_x, mu_y, sd_y, mode=mode, keep=True)
    n = X.shape[0]
    yn = (y - mu_y) / sd_y
    tn = (np.asarray(T, dtype=yn.dtype) - mu_y) / sd_y
    loss = float(((yn - tn) ** 2).sum() / n)
    d = 2.0 * (yn - tn) / n                      # d loss / d z_last
    gW, gb = [None] * NL, [None] * NL
    for l in range(NL - 1, -1, -1):
        if l < NL - 1:
            d = d * act_grad(pres[l])
        gW[l] = d.T @ acts[l]
        gb[l] = d.sum(axis=0)
        if l > 0:
            w = bf16(W[l]) if mode == "bf16" else np.asarray(W[l], dtype=d.dtype)
            d = d @ w
    return loss, gW, gb


def adam(W, b, gW, gb, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update of the FP32 master parameters (Kingma & Ba 2015)."""
    t = state.setdefault("t", 0) + 1
    state["t"] = t
    for l in range(NL):
        for key, P, G in (("W", W, gW), ("b", b, gb)):
            m = state.setdefault(key + "m%d" % l, np.zeros_like(P[l]))
            v = state.setdefault(key + "v%d" % l, np.zeros_like(P[l]))
            m[:] = beta1 * m + (1 - beta1) * G[l]
            v[:] = beta2 * v + (1 - beta2) * G[l] * G[l]
            mh = m / (1 - beta1 ** t)
            vh = v / (1 - beta2 ** t)
            P[l] = (P[l] - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
