"""FP64 numpy oracle of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Same network and parameter layout as the library (include/diskfft_mlp.h):
4 -> 128 -> 128 -> 128 -> 3, GELU(z) = z/2 (1 + erf(z / sqrt 2)), loss =
mean over points and outputs of (y - t)^2.  Analytic backward pass in FP64;
pinned by central finite differences (tests/test_mlp_oracle.py).
`forward_bf16` repeats the device's roundings (inputs, weights and hidden
activations to BF16, round to nearest even) for forward comparisons.
"""
import numpy as np
from scipy.special import erf

H = 128
NPAR = 34051
OFF = dict(W1=(0, (H, 4)), b1=(512, (H,)), W2=(640, (H, H)), b2=(17024, (H,)),
           W3=(17152, (H, H)), b3=(33536, (H,)), W4=(33664, (3, H)), b4=(34048, (3,)))


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return {k: w[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def pack(p):
    w = np.zeros(NPAR)
    for k, (o, s) in OFF.items():
        w[o:o + int(np.prod(s))] = np.asarray(p[k]).ravel()
    return w


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(w, x):
    p = unpack(w)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    z3 = h2 @ p["W3"].T + p["b3"]
    h3 = gelu(z3)
    y = h3 @ p["W4"].T + p["b4"]
    return y, (x, z1, h1, z2, h2, z3, h3)


def loss_grad(w, x, t):
    """Mean squared error and its gradient (FP64, analytic)."""
    p = unpack(w)
    y, (x, z1, h1, z2, h2, z3, h3) = forward(w, x)
    n = x.shape[0]
    d = y - t
    loss = np.sum(d * d) / (3 * n)
    gy = 2.0 * d / (3 * n)
    g = {}
    g["W4"] = gy.T @ h3
    g["b4"] = gy.sum(0)
    dz3 = (gy @ p["W4"]) * gelu_d(z3)
    g["W3"] = dz3.T @ h2
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ p["W3"]) * gelu_d(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * gelu_d(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def bf16(a):
    """Round float values to BF16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward_bf16(w, x):
    """Forward pass with the device's roundings (accumulation in FP64)."""
    p = unpack(w)
    a = bf16(x)
    for k in (1, 2, 3):
        z = bf16(a @ bf16(p[f"W{k}"]).T + p[f"b{k}"].astype(np.float32))
        a = bf16(gelu(z))
    return a @ bf16(p["W4"]).T + p["b4"].astype(np.float32)
