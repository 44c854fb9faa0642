"""NumPy reference of the warm-start network (TEST INFRASTRUCTURE).

Forward pass, mean-squared-error loss and its analytic gradient, in float64
from the float32 parameter blob of mc_mlp_set (layout in
include/mochii_gpu.h).  The gradient here is itself pinned by central finite
differences (tests/test_mlp_ref.py); the HIP training kernels are checked
against it, and the BF16 matrix-core inference against its forward pass.
The network has no counterpart in the reference program.
"""
import numpy as np
from scipy.special import erf

NI, NH, NO = 8, 64, 4
SHAPES = [("W1", (NI, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)),
          ("W3", (NH, NH)), ("b3", (NH,)), ("W4", (NH, NO)), ("b4", (NO,)),
          ("in_mu", (NI,)), ("in_sd", (NI,)), ("out_mu", (NO,)), ("out_sd", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
NTRAIN = sum(int(np.prod(s)) for n, s in SHAPES if n[0] in "Wb")


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    assert P.size == NPARAM
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = P[o:o + k].reshape(shp)
        o += k
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(P, X, standardised=False):
    """Outputs for feature rows X (n x 8): de-standardised targets, or the
    standardised network outputs when standardised=True."""
    p = unpack(P)
    h = (np.asarray(X, dtype=np.float64) - p["in_mu"]) / p["in_sd"]
    for l in (1, 2, 3):
        h = gelu(h @ p[f"W{l}"] + p[f"b{l}"])
    z = h @ p["W4"] + p["b4"]
    return z if standardised else z * p["out_sd"] + p["out_mu"]


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(P, X):
    """The inference arithmetic of the matrix-core kernel: standardised
    inputs, weights and hidden activations rounded to bfloat16, products
    accumulated exactly here (FP32 on the device), biases and GELU in full
    precision; de-standardised outputs."""
    p = unpack(np.asarray(P, dtype=np.float32))
    x = (np.asarray(X, dtype=np.float32) - p["in_mu"].astype(np.float32)) / p["in_sd"].astype(np.float32)
    h = bf16(x)
    for l in (1, 2, 3):
        z = h @ bf16(p[f"W{l}"]) + p[f"b{l}"]
        h = bf16(gelu(z.astype(np.float32).astype(np.float64)))
    z = h @ bf16(p["W4"]) + p["b4"]
    return z * p["out_sd"] + p["out_mu"]


def loss_and_grad(P, X, Y):
    """Mean of the squared standardised output errors over n x 4 entries and
    its gradient in the blob layout (standardisation entries zero)."""
    p = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    Yn = (np.asarray(Y, dtype=np.float64) - p["out_mu"]) / p["out_sd"]
    a = [(X - p["in_mu"]) / p["in_sd"]]
    zs = []
    for l in (1, 2, 3):
        z = a[-1] @ p[f"W{l}"] + p[f"b{l}"]
        zs.append(z)
        a.append(gelu(z))
    out = a[-1] @ p["W4"] + p["b4"]
    e = out - Yn
    loss = float(np.mean(e * e))
    dz = 2.0 * e / e.size
    grads = {"W4": a[3].T @ dz, "b4": dz.sum(axis=0)}
    for l in (3, 2, 1):
        da = dz @ p[f"W{l + 1}"].T
        dz = da * gelu_grad(zs[l - 1])
        grads[f"W{l}"] = a[l - 1].T @ dz
        grads[f"b{l}"] = dz.sum(axis=0)
    g = np.zeros(NPARAM)
    o = 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        if name in grads:
            g[o:o + k] = grads[name].ravel()
        o += k
    return loss, g
