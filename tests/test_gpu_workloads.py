"""The two large workloads against the oracle at reduced size (the same
generators, fewer leaves and packets) and at full size through
size-independent properties: the energy closure, the emitted luminosity, the
scored segments and a sane thermal update.

  multifractal: the 10-level cascade octree (here lmax 6, ~1.1e4 leaves),
                16 seeded sources + the external sphere field, dust with HG
                scattering, H + He thermal balance with local_ne cooling;
  turbulent:    the lognormal car cloud (here 24^3) with 4 sources and the
                trace metals C N O Ne S Ar."""
import numpy as np
import pytest

from mochii_amd.lib import MC_FP64, MC_MIXED
from mochii_amd.workloads import (HI_COOLING_FIT, multifractal_cloud, turbulent_cloud)
from tests.cases import gpu_model, oracle_model, rel_close

pytestmark = pytest.mark.gpu
A_FIT = [a for a, _ in HI_COOLING_FIT]
T_FIT = [t for _, t in HI_COOLING_FIT]


def _ionised_near_sources(params, x, y, z, nH, radius, seed=7):
    """x_HI 1e-3 (+ He singly ionised, T_e 8000-15000 K) within radius of any
    source, neutral and cold elsewhere."""
    near = np.zeros(x.size, bool)
    for k in range(params["nsource"]):
        r = np.sqrt((x - params["src_x"][k]) ** 2 + (y - params["src_y"][k]) ** 2 +
                    (z - params["src_z"][k]) ** 2)
        near |= r < radius
    rng = np.random.default_rng(seed)
    xHI = np.where(near, 10 ** rng.uniform(-4, -2.5, x.size), 1.0)
    xHeI = np.where(near, 10 ** rng.uniform(-3, -1, x.size), 1.0)
    xHeII = np.where(near, rng.uniform(0.3, 0.9, x.size), 0.0)
    Te = np.where(near, rng.uniform(8e3, 1.5e4, x.size), 1e3)
    ne = nH * ((1 - xHI) + 0.1 * (xHeII + 2 * np.maximum(0, 1 - xHeI - xHeII)))
    return dict(xHI=xHI, xHeI=xHeI, xHeII=xHeII, Te=Te, ne=ne), near


def _check_pass(orc, gm, prec, dust):
    to, tg = orc.tallies(), gm.tallies()
    tol, floor = (1e-9, 1e-6) if prec == MC_FP64 else (2e-4, 1e-4)
    keys = ("hhe", "jt", "dust_abs") if dust else ("hhe", "jt")
    for k in keys:
        ok, err = rel_close(tg[k], to[k], tol, floor_frac=floor)
        assert ok, (k, err)
    if prec == MC_FP64:
        assert (tg["seg_direct"], tg["seg_scattered"]) == (to["seg_direct"], to["seg_scattered"])
    led = gm.energy_ledger()
    np.testing.assert_allclose(led, to["ledger"], rtol=1e-12 if prec == MC_FP64 else 1e-5,
                               atol=1e-12 * led[0])
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]
    ro, rg = orc.rates(), gm.rates()
    for k in ("gHI", "gHeI", "gHeII", "hHI", "hHeI", "hHeII"):
        ok, err = rel_close(rg[k], ro[k], tol, floor_frac=floor)
        assert ok, (k, err)
    if dust:
        ok, err = rel_close(gm.dust("heat_dust"), orc.dust_heat(), tol, floor_frac=floor)
        assert ok, ("heat_dust", err)
    return to, tg


def _check_thermal(orc, gm, near):
    mo = orc.thermal_update()
    mg = gm.gas_thermal_update()
    so, sg = orc.state(), gm.state()
    rel = np.abs(sg["Te"] / so["Te"] - 1.0)
    print("T_e: max rel", rel.max(), "bit-equal", np.mean(rel == 0.0))
    # the cell solve and the balance run without FMA contraction, as the
    # oracle and the reference compute them: T_e is bit-equal leaf for leaf,
    # also in faint neutral leaves where the last bit of x_HI decides the sign
    assert np.all(rel == 0.0)
    ion = so["xHI"] < 0.5
    assert ion.sum() > 50
    ok, err = rel_close(1.0 - sg["xHI"][ion], 1.0 - so["xHI"][ion], 1e-5)
    assert ok, err
    np.testing.assert_allclose(mg, mo, rtol=2e-3, atol=5e-5)


@pytest.mark.parametrize("prec,rows", [(MC_FP64, 0), (MC_MIXED, 1), (MC_MIXED, 2)])
def test_multifractal_reduced_matches_oracle(prec, rows):
    """rows: the mixed walk's moment rows issued per lane (1) or transposed (2)."""
    params, nH, octree, rhokap, tab = multifractal_cloud(nphotons=20000, precision=prec, lmax=6,
                                                         mass_thr=3e-4, ir_packets=1e5)
    assert 5000 < nH.size < 30000 and params["nsource"] == 16 and params["ext_on"]
    orc = oracle_model(params, nH, rhokap=rhokap, dust_table=tab, octree=octree)
    gm = gpu_model(params, nH, rhokap=rhokap, dust_table=tab, octree=octree)
    gm.set_row_issue(rows)
    st, near = _ionised_near_sources(params, *octree[:3], nH, 6.0)
    for m in (orc, gm):
        m.set_state(**st)
    orc.opacity_fill()
    gm.gas_opacity_fill()
    orc.cooling_setup(A_FIT, T_FIT)
    gm.cooling_setup(HI_COOLING_FIT)
    orc.transport_pass(nthreads=8)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    to, _ = _check_pass(orc, gm, prec, dust=True)
    assert to["seg_scattered"] > 0
    if prec == MC_FP64:
        _check_thermal(orc, gm, near)


def test_multifractal_moment_cache_matches_per_coefficient_sums():
    """The mixed-form walk of the reduced multifractal model (16 sources, the
    external sphere field, grains with scattering) with its energy-moment
    tallies merged in the per-wave cache agrees with the per-coefficient sums
    of the same walk in every rate sum, band J bin and the grain absorption,
    in one launch batch and in batches of 4096 records, with the cache's rows
    issued per lane and transposed."""
    from mochii_amd.lib import MC_ROWS_LANE, MC_ROWS_TOGETHER
    params, nH, octree, rhokap, tab = multifractal_cloud(nphotons=200_000, precision=MC_MIXED,
                                                         lmax=6, mass_thr=3e-4, ir_packets=1e5)
    st, _ = _ionised_near_sources(params, *octree[:3], nH, 6.0)
    out = {}
    cases = [(False, None, MC_ROWS_LANE)] + [(True, b, r) for b in (None, 4096)
                                             for r in (MC_ROWS_LANE, MC_ROWS_TOGETHER)]
    for mom, batch, rows in cases:
        gm = gpu_model(params, nH, rhokap=rhokap, dust_table=tab, octree=octree)
        gm.set_state(**st)
        gm.gas_opacity_fill()
        gm.set_tally_moments(mom)
        gm.set_row_issue(rows)
        if batch:
            gm.set_launch_batch(batch)
        gm.ionizing_field_and_rates()
        if mom:
            assert gm.tally_moment_info()[0] > 16
            assert gm.row_issue()[1] == (1 if rows == MC_ROWS_TOGETHER else 0)
        out[(mom, batch, rows)] = gm.tallies()
        gm.close()
    ref = out[cases[0]]
    for key in cases[1:]:
        t = out[key]
        assert (t["seg_direct"], t["seg_scattered"]) == (ref["seg_direct"], ref["seg_scattered"])
        for k in ("hhe", "jt", "dust_abs"):
            ok, e = rel_close(t[k], ref[k], 1e-6, floor_frac=1e-9)
            assert ok, (key, k, e)


def test_turbulent_reduced_matches_oracle():
    n = 24
    params, nH = turbulent_cloud(n=n, nphotons=20000, half=4.0)
    orc = oracle_model(params, nH)
    gm = gpu_model(params, nH)
    c = -4.0 + (np.arange(n) + 0.5) * (8.0 / n)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")            # raster order, x fastest
    st, near = _ionised_near_sources(params, X.ravel(), Y.ravel(), Z.ravel(), nH, 1.5)
    for m in (orc, gm):
        m.set_state(**st)
    orc.opacity_fill()
    gm.gas_opacity_fill()
    orc.cooling_setup(A_FIT, T_FIT)
    gm.cooling_setup(HI_COOLING_FIT)
    orc.transport_pass(nthreads=8)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    _check_pass(orc, gm, MC_FP64, dust=False)
    _check_thermal(orc, gm, near)


@pytest.mark.parametrize("which", ["multifractal", "turbulent"])
def test_full_size_properties(which):
    """The full-size models (2.0e7 leaves / 256^3) for one pass of 1e7
    packets and one thermal update in the mixed form: closure <= 5e-13, the
    emitted luminosity is the band total, every packet scores, the state
    stays physical."""
    if which == "multifractal":
        params, nH, octree, rhokap, tab = multifractal_cloud(nphotons=10_000_000,
                                                             precision=MC_MIXED)
        gm = gpu_model(params, nH, rhokap=rhokap, dust_table=tab, octree=octree)
        assert nH.size > 1.5e7
    else:
        params, nH = turbulent_cloud(nphotons=10_000_000, precision=MC_MIXED)
        gm = gpu_model(params, nH)
        assert nH.size == 256 ** 3
    gm.cooling_setup(HI_COOLING_FIT)
    gm.ionizing_field_and_rates()
    led = gm.energy_ledger()
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]
    assert led[0] == pytest.approx(gm.ion_Ltot, rel=1e-9)
    nd, ns = gm.segment_counts()
    assert nd > 10_000_000
    r = gm.rates()
    assert np.all(np.isfinite(r["gHI"])) and (r["gHI"] > 0).sum() > 0.001 * nH.size
    gm.gas_thermal_update()
    s = gm.state()
    assert np.all((s["xHI"] >= 0) & (s["xHI"] <= 1))
    assert np.all((s["Te"] >= params["te_min"]) & (s["Te"] <= params["te_max"]))
    assert np.all(np.isfinite(s["ne"]) & (s["ne"] >= 0))
