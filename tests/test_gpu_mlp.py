"""Learned warm start on the GPU: the HIP training gradient against the NumPy
reference, BF16 matrix-core inference against the FP32 forward pass,
training on the device, and the thermal solve started from the network
against the plain bisection (same answer to the bisection's tolerance, fewer
net-rate evaluations)."""
import numpy as np
import pytest

import mochii_amd as mc
from mochii_amd import surrogate
from mochii_amd.workloads import HI_COOLING_FIT, hii_thermal
from oracle import mlp_ref as M
from tests.cases import gpu_model

pytestmark = pytest.mark.gpu


def _ctx():
    params, nH = hii_thermal(n=4, nphotons=100)
    return gpu_model(params, nH)


def _blob(rng, scale=0.25):
    P = np.zeros(M.NPARAM, dtype=np.float32)
    P[:M.NTRAIN] = rng.normal(0.0, scale, M.NTRAIN)
    o = M.NTRAIN
    P[o:o + 8] = rng.normal(0, 1, 8)
    P[o + 8:o + 16] = rng.uniform(0.5, 2.0, 8)
    P[o + 16:o + 20] = rng.normal(0, 1, 4)
    P[o + 20:o + 24] = rng.uniform(0.5, 2.0, 4)
    return P


def test_training_gradient_matches_reference():
    """FP32 forward/backward kernels vs the float64 analytic gradient (itself
    pinned by finite differences)."""
    rng = np.random.default_rng(2)
    g = _ctx()
    P = _blob(rng)
    X = rng.normal(0, 1.5, (700, 8)).astype(np.float32)
    Y = rng.normal(0, 1.0, (700, 4)).astype(np.float32)
    g.mlp_set(P)
    loss, grad = g.mlp_grad(X, Y)
    lr, gr = M.loss_and_grad(P, X, Y)
    assert abs(loss / lr - 1.0) < 1e-5
    scale = np.abs(gr[:M.NTRAIN]).max()
    err = np.abs(grad[:M.NTRAIN] - gr[:M.NTRAIN])
    assert err.max() <= 1e-4 * scale, err.max() / scale
    assert np.all(grad[M.NTRAIN:] == 0.0)


def test_bf16_inference_matches_fp32_forward():
    """BF16 weights and activations with FP32 accumulation on the matrix
    cores: equal to a NumPy emulation of the same roundings up to the
    accumulation order, and within BF16 rounding of the FP32 network; ragged
    row counts (not a multiple of the 32-leaf tile) included."""
    rng = np.random.default_rng(3)
    g = _ctx()
    P = _blob(rng)
    g.mlp_set(P)
    for n in (1, 31, 33, 1000, 4099):
        X = rng.normal(0, 1.5, (n, 8)).astype(np.float32) * P[M.NTRAIN + 8:M.NTRAIN + 16] \
            + P[M.NTRAIN:M.NTRAIN + 8]
        got = g.mlp_infer(X).astype(np.float64)
        sd = P[M.NTRAIN + 20:M.NTRAIN + 24].astype(np.float64)
        err = np.abs(got - M.forward_bf16(P, X)) / sd      # in standardised units
        assert err.max() < 1e-2 and err.mean() < 1e-3, (n, err.max(), err.mean())
        err = np.abs(got - M.forward(P, X)) / sd
        assert err.max() < 0.15 and err.mean() < 0.025, (n, err.max(), err.mean())


def test_training_on_device_fits_a_smooth_map():
    rng = np.random.default_rng(4)
    g = _ctx()
    X = rng.normal(0, 1, (20000, 8)).astype(np.float32)
    Y = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), X[:, 2] * X[:, 3], np.sin(X[:, 4]),
                  X[:, 5] - 0.3 * X[:, 6] ** 2], axis=1).astype(np.float32)
    P, hist = surrogate.train(g, X, Y, schedule=((1500, 2e-3),), batch=1024, seed=5)
    assert hist[-100:].mean() < 0.1 * hist[:20].mean()
    pred = M.forward(P, X[:2000])
    rel = np.mean((pred - Y[:2000]) ** 2) / np.mean((Y[:2000] - Y[:2000].mean(0)) ** 2)
    assert rel < 0.05


@pytest.fixture(scope="module")
def trained():
    """A network trained on the library's own H II region iterations."""
    params, nH = hii_thermal(n=16, nphotons=60_000, diffuse=True)
    g = gpu_model(params, nH)
    g.cooling_setup(HI_COOLING_FIT)
    X, Y = surrogate.collect(g, 30)
    X, Y = surrogate.balance(X, Y, max_rows=200_000)
    P, hist = surrogate.train(g, X, Y, schedule=((2000, 2e-3), (1000, 4e-4)), batch=2048)
    return P, hist


def _state_and_rates(model):
    return model.state("xHI", "xHeI", "xHeII", "ne", "Te")


@pytest.mark.parametrize("iters,levels", [(1, 9), (12, 9), (1, 0), (12, 0)])
def test_warm_start_solve_matches_bisection(trained, iters, levels):
    """From the same state and rates, the network-started Newton solve and the
    reference bisection return T_e within the bisection's bracket width and
    the same ionisation state; the warm start needs fewer net-rate
    evaluations, and the Newton start from the current T_e is reported
    beside it.  levels: bisection levels run before the warm solve (9, the
    default, bounds a leaf with several crossings to the bisection's level-9
    cell; 0 starts the warm solve on the full range, where the network's
    gain over the Newton start shows)."""
    P, _ = trained
    params, nH = hii_thermal(n=16, nphotons=60_000, diffuse=True, seed=777)
    g = gpu_model(params, nH)
    g.cooling_setup(HI_COOLING_FIT)
    g.mlp_set(P)
    for _ in range(iters - 1):
        g.ionizing_field_and_rates()
        g.gas_thermal_update()
    g.ionizing_field_and_rates()
    s0 = _state_and_rates(g)
    res, ev = {}, {}
    g.set_warm_levels(levels)
    for mode in (mc.lib.MC_WARM_NONE, mc.lib.MC_WARM_NEWTON, mc.lib.MC_WARM_MLP):
        g.set_state(**s0)
        g.set_warm_start(mode)
        m = g.gas_thermal_update()
        res[mode] = (g.state(), m)
        ev[mode] = g.thermal_stats()
    T0 = res[0][0]["Te"]
    for mode in (1, 2):
        # two searches that each end in a verified sign-change bracket of
        # relative width 1e-5 agree to that width wherever net(T) crosses
        # zero once.  Where it crosses more than once, the warm forms run the
        # reference's bisection for their first MC_WARM_LEVELS_DEFAULT = 9
        # levels and solve inside that level's cell, which holds the
        # bisection's crossing: any other crossing they end on lies within
        # the cell, a factor (te_max / te_min)^(2^-9) = 1.0077 of it.
        rel = np.abs(res[mode][0]["Te"] / T0 - 1.0)
        bound = (params["te_max"] / params["te_min"]) ** (2.0 ** -9) - 1.0
        far = int(np.sum(rel >= 5e-3))
        print(f"mode {mode}: T_e rel diff max {rel.max():.2e}, "
              f"fraction within 2.1e-5: {np.mean(rel < 2.1e-5):.4f}, leaves beyond 5e-3: {far}")
        assert np.mean(rel < 2.1e-5) > 0.99
        if levels == 9:
            assert rel.max() < bound, (rel.max(), bound)
        else:
            # without the levels a start can end on another crossing (13% away
            # for 1 leaf in 4096 seen): bounded to 1 in 1000
            assert far <= rel.size // 1000
        x0, x = 1 - res[0][0]["xHI"], 1 - res[mode][0]["xHI"]
        assert np.max(np.abs(x - x0)) < 1e-4
        assert ev[mode]["fallbacks"] <= 0.01 * max(ev[mode]["bracketed"], 1)
    assert ev[0]["bracketed"] > 20
    print("net-rate evaluations per bracketed leaf:",
          {k: v["evaluations"] / max(v["bracketed"], 1) for k, v in ev.items()})
    assert ev[2]["evaluations"] < (0.8 if levels else 0.6) * ev[0]["evaluations"]
    if iters == 1 and levels == 0:
        # first update: the current T_e is the uniform initial guess, the
        # network's guess is informed by the rates
        assert ev[2]["evaluations"] < ev[1]["evaluations"]
