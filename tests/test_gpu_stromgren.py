"""The Stromgren sphere converged on the GPU (the reference's G1 gate,
tests/g1_stromgren/check_g1_stromgren.py, and the iteration of
src/main.f90:124-227): against the oracle iterated to convergence on the
same model (same Sobol set, same criterion) -- the same iteration count,
update history and converged x_HI -- and against the analytic radius
R_S = (3 Q_H / (4 pi n_H^2 alpha_B))^(1/3) through the ionised-volume radius
R_eff = (3 V_ion / 4 pi)^(1/3) of the leaves."""
import numpy as np
import pytest

import mochii_amd as mc
from mochii_amd.workloads import PC2CM, stromgren
from oracle import pyoracle as po
from tests.cases import gpu_model, oracle_model

pytestmark = pytest.mark.gpu


def _oracle_run(params, nH, niter):
    o = oracle_model(params, nH)
    hist = []
    for it in range(1, niter + 1):
        o.transport_pass(nthreads=8)
        o.rates_compute()
        max_dx, dx_vol = o.equilibrium_update()
        hist.append((it, max_dx, dx_vol))
        if dx_vol < params["gas_tol"]:
            break
    return o, hist


def test_converged_sphere_matches_oracle_iteration_by_iteration():
    params, nH = stromgren(n=16, nphotons=20000, niter=60)
    g = gpu_model(params, nH)
    res = mc.run_model(g, final_pass=False)
    o, ohist = _oracle_run(params, nH, 60)
    assert res["converged"] and len(res["history"]) == len(ohist) > 5
    np.testing.assert_allclose(np.array(res["history"])[:, 1:], np.array(ohist)[:, 1:],
                               rtol=1e-6, atol=1e-12)
    gx, ox = g.state()["xHI"], o.state()["xHI"]
    np.testing.assert_allclose(gx, ox, rtol=1e-6, atol=1e-12)


def test_ionised_volume_radius_against_analytic():
    n, half, nHv = 48, 4.0, 100.0
    params, nH = stromgren(n=n, nphotons=400000, niter=80, half=half, nH=nHv)
    g = gpu_model(params, nH)
    res = mc.run_model(g)
    assert res["converged"]
    x = g.state()["xHI"]
    dc = 2.0 * half / n
    v_ion = np.sum(1.0 - x) * dc ** 3
    r_eff = (3.0 * v_ion / (4.0 * np.pi)) ** (1.0 / 3.0)
    alpha_b = po.lib().orc_alphaB_HII(1.0e4)
    r_s = (3.0 * 1.0e49 / (4.0 * np.pi * nHv ** 2 * alpha_b)) ** (1.0 / 3.0) / PC2CM
    print("R_eff", r_eff, "R_S", r_s, "ratio", r_eff / r_s)
    # the front's finite width and the partial ionisation inside it leave
    # R_eff within a few per cent of the sharp-edge radius (the reference's
    # gate compares with a 1D marching profile at 1%; the sharp edge differs
    # from that profile by ~4%)
    assert abs(r_eff / r_s - 1.0) < 0.06
    c = -half + (np.arange(n) + 0.5) * dc
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).transpose(2, 1, 0).ravel()   # raster order, x fastest
    assert np.all(x[r < 0.6 * r_s] < 1e-2)
    out = x[r > 1.25 * r_s]
    assert np.median(out) > 0.9 and np.all(out > 0.5)
