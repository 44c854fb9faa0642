"""HIP thermal balance (gas_thermal_update) against the FP64 oracle, through
the C-ABI: same pass, same cooling tables, bisection on log T per leaf."""
import numpy as np
import pytest

from mochii_amd import MC_FP64, run_model
from mochii_amd.workloads import HI_COOLING_FIT, hii_thermal
from tests.cases import gpu_model, oracle_model, rel_close

pytestmark = pytest.mark.gpu


def _pair(n=16, nphot=20000, **kw):
    params, nH = hii_thermal(n=n, nphotons=nphot, **kw)
    orc, gm = oracle_model(params, nH), gpu_model(params, nH)
    orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
    gm.cooling_setup(HI_COOLING_FIT)
    return params, nH, orc, gm


def _compare_update(orc, gm, tol_T=1e-4):
    mo = orc.thermal_update()
    mg = gm.gas_thermal_update()
    so, sg = orc.state(), gm.state()
    # the cell solve and the balance run without FMA contraction, as the
    # oracle and the reference compute them, and the rates agree to their
    # reduction order: every bisection takes the same path and T_e is
    # bit-equal leaf for leaf
    rel = np.abs(sg["Te"] / so["Te"] - 1.0)
    print("T_e: max rel diff", rel.max(), "bit-equal fraction", np.mean(rel == 0.0))
    assert np.all(rel == 0.0), rel.max()
    tol_x = max(5e-5, tol_T)
    ion = so["xHI"] < 0.5
    ok, err = rel_close(1.0 - sg["xHI"][ion], 1.0 - so["xHI"][ion], 1e-6)
    assert ok, err
    ok, err = rel_close(sg["xHI"][ion], so["xHI"][ion], tol_x)
    assert ok, err
    np.testing.assert_allclose(sg["ne"], so["ne"], rtol=tol_x, atol=1e-9)
    np.testing.assert_allclose(mg, mo, rtol=1e-4, atol=5e-5)
    return so, sg


def test_thermal_update_matches_oracle():
    """Two iterations (pass -> thermal update -> opacity feedback -> pass ->
    update) on a 16^3 H + He H II region, FP64 transport."""
    n = 16
    params, nH, orc, gm = _pair(n=n)
    c = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X**2 + Y**2 + Z**2).ravel()
    xHI = np.where(r < 0.5, 1e-3, 1.0)                      # ionised core, radius 4 pc
    orc.set_state(xHI=xHI)
    gm.set_state(xHI=xHI)
    for it in range(2):
        orc.transport_pass(nthreads=4)
        orc.rates_compute()
        gm.ionizing_field_and_rates()
        ok, err = rel_close(gm.rates()["hHI"], orc.rates()["hHI"], 1e-5 if it else 1e-9)
        assert ok, (it, err)
        so, sg = _compare_update(orc, gm, 1e-4 if it == 0 else 2e-3)
    hot = so["Te"] > 5e3
    assert hot.sum() > 200, hot.sum()


def test_thermal_converges_to_warm_hii_region():
    """GPU-only iteration to the volume criterion: photo-heated metal-free gas
    sits at 1.2e4-3e4 K inside the ionised zone, neutral gas at te_min, and
    the pass still closes its energy budget."""
    params, nH = hii_thermal(n=24, nphotons=200_000, niter=80)
    gm = gpu_model(params, nH)
    gm.cooling_setup(HI_COOLING_FIT)
    res = run_model(gm)
    assert res["converged"], res["history"][-3:]
    st = gm.state()
    ion = st["xHI"] < 0.1
    assert ion.sum() > 100
    Tm = np.median(st["Te"][ion])
    assert 1.2e4 < Tm < 3.0e4, Tm
    led = gm.energy_ledger()
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]


def test_thermal_update_requires_cooling_setup():
    params, nH = hii_thermal(n=8, nphotons=1000)
    gm = gpu_model(params, nH)
    gm.ionizing_field_and_rates()
    from mochii_amd.lib import MochiiError
    with pytest.raises(MochiiError, match="cooling_setup"):
        gm.gas_thermal_update()


def test_local_ne_cooling_matches_oracle_and_suppresses_line_cooling():
    """cooling_model 'local_ne' (the reference default): the H I and metal
    line-cooling fits suppressed by the n-level (T, n_e) tables.  At
    n_H = 1500 cm^-3 the metal lines are partly collisionally de-excited:
    the thermal update matches the oracle with the tables, and runs hotter
    than with the unsuppressed 'low_density' fits, on both sides."""
    from mochii_amd.lib import MC_COOL_LOW_DENSITY
    n = 10
    res = {}
    for model in (1, MC_COOL_LOW_DENSITY):
        params, nH = hii_thermal(n=n, nphotons=20000, metals=True, cooling_model=model,
                                 nH=np.full(n ** 3, 1500.0), half=2.0)
        orc, gm = oracle_model(params, nH), gpu_model(params, nH)
        orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
        gm.cooling_setup(HI_COOLING_FIT)
        c = (np.arange(n) + 0.5) / n * 2 - 1
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        xHI = np.where(np.sqrt(X**2 + Y**2 + Z**2).ravel() < 0.35, 1e-3, 1.0)
        orc.set_state(xHI=xHI)
        orc.opacity_fill()
        gm.set_state(xHI=xHI)
        orc.transport_pass(nthreads=8)
        orc.rates_compute()
        gm.ionizing_field_and_rates()
        orc.thermal_update()
        gm.gas_thermal_update()
        so, sg = orc.state(), gm.state()
        rel = np.abs(sg["Te"] / so["Te"] - 1.0)
        print(model, "T_e max rel", rel.max(), "bit-equal", np.mean(rel == 0.0))
        assert np.all(rel == 0.0)
        res[model] = (so["Te"], sg["Te"], so["xHI"])
    lit = res[1][2] < 0.5
    assert lit.sum() >= 5
    for side in (0, 1):
        dT = res[1][side][lit] / res[MC_COOL_LOW_DENSITY][side][lit] - 1.0
        print("local_ne / low_density - 1 on lit leaves:", np.median(dT), dT.max())
        assert np.median(dT) > 1e-3 and np.all(dT > -1e-4)
