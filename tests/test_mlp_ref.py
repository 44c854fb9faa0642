"""The NumPy reference of the warm-start network: its analytic gradient
against central finite differences (float64), and its forward pass against
a direct evaluation of the formula layer by layer."""
import numpy as np

from oracle import mlp_ref as M


def _random_blob(rng, scale=0.3):
    P = np.zeros(M.NPARAM)
    P[:M.NTRAIN] = rng.normal(0.0, scale, M.NTRAIN)
    p = M.unpack(P)
    o = M.NTRAIN
    P[o:o + 8] = rng.normal(0, 1, 8)               # in_mu
    P[o + 8:o + 16] = rng.uniform(0.5, 2.0, 8)     # in_sd
    P[o + 16:o + 20] = rng.normal(0, 1, 4)         # out_mu
    P[o + 20:o + 24] = rng.uniform(0.5, 2.0, 4)    # out_sd
    return P


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = _random_blob(rng)
    X = rng.normal(0, 1.5, (40, 8))
    Y = rng.normal(0, 1.0, (40, 4))
    loss, g = M.loss_and_grad(P, X, Y)
    idx = np.concatenate([rng.choice(M.NTRAIN, 300, replace=False),
                          np.arange(M.NTRAIN - 8, M.NTRAIN)])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (M.loss_and_grad(Pp, X, Y)[0] - M.loss_and_grad(Pm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    assert np.all(g[M.NTRAIN:] == 0.0)


def test_forward_layer_by_layer():
    rng = np.random.default_rng(1)
    P = _random_blob(rng)
    X = rng.normal(0, 1, (5, 8))
    p = M.unpack(P)
    from math import erf, sqrt
    for r in range(5):
        h = [(X[r, k] - p["in_mu"][k]) / p["in_sd"][k] for k in range(8)]
        for l in (1, 2, 3):
            W, b = p[f"W{l}"], p[f"b{l}"]
            z = [b[n] + sum(h[k] * W[k, n] for k in range(len(h))) for n in range(64)]
            h = [0.5 * v * (1 + erf(v / sqrt(2))) for v in z]
        out = [p["b4"][n] + sum(h[k] * p["W4"][k, n] for k in range(64)) for n in range(4)]
        want = np.array(out) * p["out_sd"] + p["out_mu"]
        np.testing.assert_allclose(M.forward(P, X[r:r + 1])[0], want, rtol=1e-12, atol=1e-12)
