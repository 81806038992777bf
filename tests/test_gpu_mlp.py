"""The closure network's HIP kernels (lsd_mlp_*) against the CPU oracle
(oracle/mlp_ref.py, pinned by finite differences), at the closure's sizes.

The oracle rounds every matrix-product operand to bfloat16 as the matrix
cores do and accumulates in float32; the GPU sums in another order, so a
value that lands on the other side of a bfloat16 rounding boundary before
the next layer can move single outputs by about one bfloat16 step (2^-8).
Bounds: outputs and gradients to 2e-2 of their largest magnitude, 2e-3 in
relative L2 over all entries.
"""
import os
import sys

import numpy as np
import pytest

from lsd.gpu import Mlp

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu

N_IN, WIDTH, N_OUT = 92, 256, 10


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_and_gradient_vs_oracle():
    rng = np.random.default_rng(3)
    m = Mlp(N_IN, WIDTH, N_OUT, seed=11)
    p = m.params()
    X = rng.standard_normal((777, N_IN)).astype(np.float32)
    T = rng.standard_normal((777, N_OUT)).astype(np.float32)
    print("mlp forward/grad rows=777", flush=True)
    y = m.forward(X)
    yo, _ = mlp_ref.forward(p, X, N_IN, WIDTH, N_OUT, rb=True)
    assert np.abs(y - yo).max() <= 2e-2 * np.abs(yo).max()
    assert rel_l2(y, yo) <= 2e-3
    loss, g = m.grad(X, T)
    lo, go = mlp_ref.loss_and_grad(p, X, T, N_IN, WIDTH, N_OUT, rb=True)
    assert abs(loss - lo) <= 1e-3 * lo
    assert rel_l2(g, go) <= 2e-3
    assert np.abs(g - go).max() <= 2e-2 * np.abs(go).max()
    m.close()


def test_training_fits_a_smooth_map():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((N_IN, N_OUT)) / np.sqrt(N_IN)
    X = rng.standard_normal((8192, N_IN)).astype(np.float32)
    T = np.tanh(X @ A).astype(np.float32)
    m = Mlp(N_IN, WIDTH, N_OUT, seed=1)
    print("mlp training 400 steps", flush=True)
    loss = m.train(X, T, batch=1024, steps=400, lr=2e-3)
    assert loss[-20:].mean() < 0.1 * loss[:5].mean()
    y = m.forward(X[:512])
    assert rel_l2(y, T[:512]) < 0.3
    m.close()
