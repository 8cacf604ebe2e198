"""FP64 numpy restatement of the h-predictor MLP -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference (its seed for the h-rho
iteration is h = h_in + dt dh/dt, src/stepND_leapfrog_mhd.f90:153).  This
module is the checker of the HIP forward / backward / Adam kernels
(libndsp_hip.so, ndsp_mlp_*): same parameter layout, same GELU (tanh form),
same loss, computed in FP64 from the BF16-rounded weights the device uses.

Parameter vector (float32[1313]):
  W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3[1]    (row-major, out x in)
Forward:  a1 = gelu(W1 x + b1), a2 = gelu(W2 a1 + b2), y = W3 a2 + b3
          (forward_bf16: the matrix-core inference, x and a1 in BF16)
Loss:     L = sum_b (y_b - t_b)^2 / B
Adam:     m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
          w -= lr * (m / (1-b1^k)) / (sqrt(v / (1-b2^k)) + eps)
"""
from __future__ import annotations

import numpy as np

NIN, NH = 6, 32
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1
C_GELU = 0.7978845608028654  # sqrt(2/pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN); o += NH * NIN
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o:o + 1]; o += 1
    assert o == NPARAM
    return W1, b1, W2, b2, W3, b3


def bf16_round(p):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    u = np.asarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def dgelu(x):
    u = C_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    du = C_GELU * (1.0 + 3 * 0.044715 * x ** 2)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * du


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = X @ W1.T + b1
    a1 = gelu(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu(z2)
    y = a2 @ W3 + b3[0]
    return y, (X, z1, a1, z2, a2)


def forward_bf16(p, X):
    """the matrix-core inference of the seed: weights, inputs and the first
    hidden layer's activations rounded to BF16, sums in FP64 here (FP32 on
    the device), the second hidden layer and the output unrounded"""
    W1, b1, W2, b2, W3, b3 = unpack(bf16_round(p))
    Xb = bf16_round(np.asarray(X, dtype=np.float32))
    a1 = bf16_round(gelu(Xb @ W1.T + b1).astype(np.float32))
    a2 = gelu(a1 @ W2.T + b2)
    return a2 @ W3 + b3[0]


def loss_and_grad(p, X, T):
    """mean squared error and its gradient with respect to every parameter"""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (X, z1, a1, z2, a2) = forward(p, X)
    B = X.shape[0]
    r = y - T
    L = float(np.sum(r * r) / B)
    dy = 2.0 * r / B
    gW3 = a2.T @ dy
    gb3 = np.array([dy.sum()])
    da2 = np.outer(dy, W3)
    dz2 = da2 * dgelu(z2)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(axis=0)
    da1 = dz2 @ W2
    dz1 = da1 * dgelu(z1)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return L, g


def adam(p, m, v, g, k, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** k)
    vh = v / (1 - beta2 ** k)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (float32 vector in the layout above)"""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    o = 0
    for nout, nin in ((NH, NIN), (NH, NH), (1, NH)):
        lim = np.sqrt(6.0 / (nout + nin))
        p[o:o + nout * nin] = rng.uniform(-lim, lim, nout * nin)
        o += nout * nin + nout  # skip the biases
    assert o == NPARAM
    return p.astype(np.float32)
