"""CPU oracle of the energy-surrogate MLP (in -> hidden x L, GELU -> 1).

TEST INFRASTRUCTURE: the tests check the GPU forward and backward kernels
against this file; nothing in the shipped library imports it.

Two variants share the same formulas:
  exact=True   float64 throughout (pinned by finite differences in the tests);
  exact=False  rounds to bfloat16 exactly where the GPU path stores bf16 (input,
               weights, hidden activations, back-propagated pre-activation
               gradients), with float64 accumulation standing in for the
               GPU's fp32 MFMA accumulation.
Loss: mean squared error over the batch.  GELU: 0.5 x (1 + erf(x / sqrt 2)).
No reference counterpart (the reference has no learned component).
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476


def bf16(x):
    """round-to-nearest-even to bfloat16, returned as float64"""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + np.uint32(0x7FFF)
    return ((u + r) & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x * SQRT1_2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x * SQRT1_2)) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def forward(params, x, exact=True):
    W, b, w_out, b_out = params
    rnd = (lambda a: a) if exact else bf16
    h = rnd(x)
    pres, acts = [], [h]
    for Wl, bl in zip(W, b):
        z = h @ rnd(Wl) + bl
        pres.append(z)
        h = rnd(gelu(z))
        acts.append(h)
    y = h @ w_out + b_out
    return y, pres, acts


def loss_and_grads(params, x, t, exact=True):
    W, b, w_out, b_out = params
    rnd = (lambda a: a) if exact else bf16
    y, pres, acts = forward(params, x, exact)
    B = x.shape[0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / B
    dw_out = acts[-1].T @ dy
    db_out = float(dy.sum())
    dP = rnd(np.outer(dy, w_out) * gelu_grad(pres[-1]))
    dW, db = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        dW[l] = acts[l].T @ dP
        db[l] = dP.sum(0)
        if l > 0:
            dP = rnd((dP @ rnd(W[l]).T) * gelu_grad(pres[l - 1]))
    return loss, (dW, db, dw_out, db_out)


def init_params(rng, in_dim=128, hidden=512, layers=3):
    W, b = [], []
    k = in_dim
    for _ in range(layers):
        W.append(rng.standard_normal((k, hidden)) * np.sqrt(2.0 / k))
        b.append(np.zeros(hidden))
        k = hidden
    return W, b, rng.standard_normal(hidden) * np.sqrt(1.0 / hidden), 0.0
