"""Pins the MLP oracle (oracle/mlp_np.py): its analytic gradients equal central
finite differences of its own loss, and its bf16 variant stays within bf16
error of the exact one."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_np  # noqa: E402


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    W, b, wo, bo = mlp_np.init_params(rng, in_dim=8, hidden=6, layers=3)
    x = rng.standard_normal((5, 8))
    t = rng.standard_normal(5)
    params = (W, b, wo, bo)
    loss, (dW, db, dwo, dbo) = mlp_np.loss_and_grads(params, x, t, exact=True)
    h = 1e-6

    def L(p):
        return mlp_np.loss_and_grads(p, x, t, exact=True)[0]

    for l in range(3):
        for (i, j) in [(0, 0), (3, 2), (5, 5)] if l else [(0, 0), (7, 5), (4, 1)]:
            Wp = [w.copy() for w in W]; Wp[l][i, j] += h
            Wm = [w.copy() for w in W]; Wm[l][i, j] -= h
            fd = (L((Wp, b, wo, bo)) - L((Wm, b, wo, bo))) / (2 * h)
            assert abs(fd - dW[l][i, j]) < 1e-6 * max(1.0, abs(fd))
        bp = [v.copy() for v in b]; bp[l][2] += h
        bm = [v.copy() for v in b]; bm[l][2] -= h
        fd = (L((W, bp, wo, bo)) - L((W, bm, wo, bo))) / (2 * h)
        assert abs(fd - db[l][2]) < 1e-6
    wp = wo.copy(); wp[3] += h
    wm = wo.copy(); wm[3] -= h
    assert abs((L((W, b, wp, bo)) - L((W, b, wm, bo))) / (2 * h) - dwo[3]) < 1e-6
    assert abs((L((W, b, wo, bo + h)) - L((W, b, wo, bo - h))) / (2 * h) - dbo) < 1e-6


def test_bf16_rounding_and_variant_error():
    assert mlp_np.bf16(np.array([1.0, 1.00390625, 1.005859375]))[0] == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert mlp_np.bf16(np.array([1.00390625]))[0] == 1.0
    assert mlp_np.bf16(np.array([1.005859375]))[0] == 1.0078125
    rng = np.random.default_rng(1)
    p = mlp_np.init_params(rng, in_dim=32, hidden=64, layers=3)
    x = rng.standard_normal((16, 32))
    ye, _, _ = mlp_np.forward(p, x, exact=True)
    yb, _, _ = mlp_np.forward(p, x, exact=False)
    assert np.abs(ye - yb).max() < 0.05 * np.abs(ye).max() + 1e-3
