"""Pin of the closure network's CPU oracle (oracle/mlp_ref.py): the
analytic gradient of the float64 form against central finite differences
of the loss, on every parameter block."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    n_in, W, n_out, rows = 7, 12, 3, 9
    n = W * n_in + W + W * W + W + n_out * W + n_out
    p = rng.standard_normal(n) * 0.4
    X = rng.standard_normal((rows, n_in))
    T = rng.standard_normal((rows, n_out))
    loss, g = mlp_ref.loss_and_grad(p, X, T, n_in, W, n_out, rb=False)
    h = 1e-6
    # a few indices in each block: W1, b1, W2, b2, W3, b3
    offs = np.cumsum([0, W * n_in, W, W * W, W, n_out * W, n_out])
    for blk in range(6):
        for i in rng.choice(np.arange(offs[blk], offs[blk + 1]), size=3, replace=False):
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            lp, _ = mlp_ref.loss_and_grad(pp, X, T, n_in, W, n_out, rb=False)
            lm, _ = mlp_ref.loss_and_grad(pm, X, T, n_in, W, n_out, rb=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (blk, i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], dtype=np.float32)
    r = mlp_ref.bf16(x)
    # 8 significant bits: 1 + 2^-8 ties to even -> 1.0; 1 + 1.5*2^-8 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125)
    assert abs(r[3] - (-3.140625)) < 1e-7
    assert r[4] == 65536.0
