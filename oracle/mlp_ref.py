"""CPU oracle of the closure network -- TEST INFRASTRUCTURE: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.

A 3-layer perceptron n_in -> W -> W -> n_out with exact-erf GELU between
the layers and a mean-squared-error loss, restated from the textbook
chain rule.  With bf16=True every matrix-product operand is rounded to
bfloat16 (round to nearest even) as the GPU's matrix cores see it, with
float32 accumulation; with bf16=False everything is float64 (the form the
finite-difference check pins).

Parameter vector, as in include/lsd.h: W1[W][n_in] b1[W] W2[W][W] b2[W]
W3[n_out][W] b3[n_out].
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def unpack(p, n_in, W, n_out):
    o = 0
    out = []
    for shape in ((W, n_in), (W,), (W, W), (W,), (n_out, W), (n_out,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def _mm(a, b, rb):
    if rb:
        return bf16(a).astype(np.float32) @ bf16(b).astype(np.float32)
    return a @ b


def forward(p, X, n_in, W, n_out, rb=True):
    dt = np.float32 if rb else np.float64
    W1, b1, W2, b2, W3, b3 = [a.astype(dt) for a in unpack(p, n_in, W, n_out)]
    X = X.astype(dt)
    Z1 = _mm(X, W1.T, rb) + b1
    H1 = gelu(Z1).astype(dt)
    Z2 = _mm(H1, W2.T, rb) + b2
    H2 = gelu(Z2).astype(dt)
    Y = _mm(H2, W3.T, rb) + b3
    return Y, (X, Z1, H1, Z2, H2)


def loss_and_grad(p, X, T, n_in, W, n_out, rb=True):
    dt = np.float32 if rb else np.float64
    W1, b1, W2, b2, W3, b3 = [a.astype(dt) for a in unpack(p, n_in, W, n_out)]
    Y, (Xc, Z1, H1, Z2, H2) = forward(p, X, n_in, W, n_out, rb)
    d = Y - T.astype(dt)
    ny = d.size
    loss = float((d.astype(np.float64) ** 2).sum() / ny)
    dY = (2.0 * d / ny).astype(dt)
    gW3 = _mm(dY.T, H2, rb)
    gb3 = dY.sum(axis=0)
    dH2 = _mm(dY, W3, rb)
    dZ2 = (dH2 * gelu_grad(Z2)).astype(dt)
    gW2 = _mm(dZ2.T, H1, rb)
    gb2 = dZ2.sum(axis=0)
    dH1 = _mm(dZ2, W2, rb)
    dZ1 = (dH1 * gelu_grad(Z1)).astype(dt)
    gW1 = _mm(dZ1.T, Xc, rb)
    gb1 = dZ1.sum(axis=0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g
