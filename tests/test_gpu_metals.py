"""HIP trace-metal path against the FP64 oracle, through the C-ABI: metal
band edges, stage cache, metal opacity and path tallies in the walk, metal
Gammas, the cell solve with metal electrons and two-way charge exchange, the
thermal balance with metal cooling / heating, the hydrogen residual."""
import numpy as np
import pytest

from mochii_amd import MC_FP64, MC_MIXED
from mochii_amd.lib import MC_CX_METAL_ONLY, MC_WARM_NEWTON, MC_WARM_NONE
from mochii_amd.atomic import metal_thresholds, registry
from mochii_amd.workloads import HI_COOLING_FIT, elements_of, hii_thermal
from tests.cases import gpu_model, oracle_model, rel_close

pytestmark = pytest.mark.gpu


def _core(n):
    c = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X**2 + Y**2 + Z**2).ravel()
    return np.where(r < 0.5, 1e-3, 1.0)          # ionised core, radius 4 pc


def _pair(n=8, nphot=20000, xhi=True, cooling=True, **kw):
    params, nH = hii_thermal(n=n, nphotons=nphot, metals=True, **kw)
    orc, gm = oracle_model(params, nH), gpu_model(params, nH)
    if cooling:
        orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
        gm.cooling_setup(HI_COOLING_FIT)
    if xhi:
        x = _core(n)
        orc.set_state(xHI=x)
        orc.opacity_fill()
        gm.set_state(xHI=x)
    return params, orc, gm


def _pass(orc, gm):
    orc.transport_pass(nthreads=4)
    orc.rates_compute()
    gm.ionizing_field_and_rates()


def test_metal_band_and_stage_cache_match_oracle():
    params, orc, gm = _pair(xhi=False)
    np.testing.assert_allclose(gm.band_edges(), orc.band()["eedge"], rtol=1e-15, atol=0)
    assert gm.band_edges().size == 17
    for ie, el in enumerate(elements_of(params)):
        fo, fg = orc.species_fractions(ie), gm.species_fractions(ie)
        np.testing.assert_allclose(fg.sum(1), 1.0, rtol=1e-13)
        ok, err = rel_close(fg, fo, 1e-12, floor_frac=1e-12)
        assert ok, (el["name"], err)


@pytest.mark.parametrize("prec", [MC_FP64, MC_MIXED])
def test_metal_pass_matches_oracle(prec):
    """One pass with metal opacity and path tallies: H/He rates, the metal
    raw(2, ntr, nleaf) tallies and the metal Gammas / photoheating."""
    params, orc, gm = _pair(n=10, precision=prec)
    _pass(orc, gm)
    tol, floor = (1e-9, 1e-6) if prec == MC_FP64 else (2e-4, 1e-4)
    ro, rg = orc.rates(), gm.rates()
    for k in ("gHI", "gHeI", "hHI", "hHeI"):
        ok, err = rel_close(rg[k], ro[k], tol, floor_frac=floor)
        assert ok, (k, err)
    led = gm.energy_ledger()
    assert abs(led[0] - led[1] - led[2]) <= 5e-13 * led[0]
    n_tested = 0
    for ie, el in enumerate(elements_of(params)):
        to, tg = orc.species_tallies(ie), gm.species_tallies(ie)
        ok, err = rel_close(tg, to, tol, floor_frac=floor)
        assert ok, (el["name"], "raw", err)
        (go, ho), (gg, hg) = orc.species_rates(ie), gm.species_rates(ie)
        ok, err = rel_close(gg, go, tol, floor_frac=floor)
        assert ok, (el["name"], "gamma", err)
        ok, err = rel_close(hg, ho, tol, floor_frac=floor)
        assert ok, (el["name"], "heat", err)
        n_tested += int((go > 0).sum())
    assert n_tested > 1000


# registries: C N O Ne S Ar and C N O Ne S take the cell-solve kernels with
# the element layout compiled in, C O S the generic registry path, and the
# reference's eleven elements (32 transitions, solar abundances of Asplund et
# al. 2009 for those whose reference default is 0) the 32-transition kernels
REGISTRIES = {"CNONeSAr": {"ar": 2.5e-6}, "CNONeS": {}, "COS": {"n": 0.0, "ne": 0.0},
              "ALL": {"ar": 2.5e-6, "mg": 3.98e-5, "fe": 3.16e-5, "si": 3.24e-5,
                      "cl": 3.16e-7, "ca": 2.19e-6}}


@pytest.mark.parametrize("cx,metal_ne,reg", [(0, 1, "CNONeSAr"), (MC_CX_METAL_ONLY, 1, "CNONeSAr"),
                                             (0, 0, "CNONeSAr"), (0, 1, "CNONeS"),
                                             (0, 1, "COS"), (0, 1, "ALL")])
def test_metal_equilibrium_update_matches_oracle(cx, metal_ne, reg):
    """Fixed-T update: solve_ion_cell with the metal electrons and
    (two_way) charge exchange, relaxation, charge neutrality of the written
    state, then the stage cache and the hydrogen residual."""
    els = registry(REGISTRIES[reg])
    params, nH = hii_thermal(n=8, nphotons=20000, metals=True)
    params.update(solve_te=0, charge_exchange=cx, metal_ne=metal_ne,
                  metal_eth=metal_thresholds(els))
    orc, gm = oracle_model(params, nH, elements=els), gpu_model(params, nH, elements=els)
    x = _core(8)
    orc.set_state(xHI=x)
    orc.opacity_fill()
    gm.set_state(xHI=x)
    _pass(orc, gm)
    mo = orc.equilibrium_update()
    orc.opacity_fill()
    mg = gm.gas_equilibrium_update()
    so, sg = orc.state(), gm.state()
    for k in ("xHI", "xHeI", "xHeII", "ne"):
        ok, err = rel_close(sg[k], so[k], 1e-8, floor_frac=1e-10)
        assert ok, (k, err)
    np.testing.assert_allclose(mg, mo, rtol=1e-8, atol=1e-12)
    for ie, el in enumerate(els):
        ok, err = rel_close(gm.species_fractions(ie), orc.species_fractions(ie), 1e-7,
                            floor_frac=1e-10)
        assert ok, (el["name"], err)
    eo, eg = orc.residual_grid(), gm.hydrogen_balance_residual_grid()
    np.testing.assert_allclose(eg, eo, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("prec", [MC_FP64, MC_MIXED])
def test_full_registry_pass_and_thermal_update_match_oracle(prec):
    """The reference's eleven elements, 32 metal transitions (past the
    16-transition register cache: the 32-transition cell-solve kernels and a
    walk with 32 cross sections per packet in LDS): H/He rates, every
    element's raw tallies, Gammas and photoheating after one pass; in the
    FP64 form the thermal update's T_e bit-equal to the oracle."""
    els = registry(REGISTRIES["ALL"])
    assert len(els) == 11 and sum(e["nstage"] - 1 for e in els) == 32
    params, nH = hii_thermal(n=8, nphotons=20000, metals=True, precision=prec)
    params["metal_eth"] = metal_thresholds(els)
    orc, gm = oracle_model(params, nH, elements=els), gpu_model(params, nH, elements=els)
    orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
    gm.cooling_setup(HI_COOLING_FIT)
    x = _core(8)
    orc.set_state(xHI=x)
    orc.opacity_fill()
    gm.set_state(xHI=x)
    _pass(orc, gm)
    tol, floor = (1e-9, 1e-6) if prec == MC_FP64 else (2e-4, 1e-4)
    ro, rg = orc.rates(), gm.rates()
    for k in ("gHI", "gHeI", "hHI", "hHeI"):
        ok, err = rel_close(rg[k], ro[k], tol, floor_frac=floor)
        assert ok, (k, err)
    for ie, el in enumerate(els):
        ok, err = rel_close(gm.species_tallies(ie), orc.species_tallies(ie), tol,
                            floor_frac=floor)
        assert ok, (el["name"], "raw", err)
        (go, ho), (gg, hg) = orc.species_rates(ie), gm.species_rates(ie)
        ok, err = rel_close(gg, go, tol, floor_frac=floor)
        assert ok, (el["name"], "gamma", err)
    if prec != MC_FP64:
        return
    orc.thermal_update()
    orc.opacity_fill()
    gm.gas_thermal_update()
    so, sg = orc.state(), gm.state()
    rel = np.abs(sg["Te"] / so["Te"] - 1.0)
    print("full registry T_e: max rel diff", rel.max())
    assert np.all(rel == 0.0)
    for ie, el in enumerate(els):
        ok, err = rel_close(gm.species_fractions(ie), orc.species_fractions(ie), 1e-7,
                            floor_frac=1e-10)
        assert ok, (el["name"], err)


def test_metal_thermal_update_matches_oracle():
    """T_e by bisection with metal line cooling, H-impact cooling, metal
    free-free and photoheating (and the metal electrons / two-way charge
    exchange in every trial solve)."""
    params, orc, gm = _pair(n=8)
    _pass(orc, gm)
    mo = orc.thermal_update()
    orc.opacity_fill()
    mg = gm.gas_thermal_update()
    so, sg = orc.state(), gm.state()
    rel = np.abs(sg["Te"] / so["Te"] - 1.0)
    print("T_e: max rel diff", rel.max(), "bit-equal fraction", np.mean(rel == 0.0))
    assert np.all(rel == 0.0)             # the solve's arithmetic is the oracle's
    ion = so["xHI"] < 0.5
    ok, err = rel_close(1.0 - sg["xHI"][ion], 1.0 - so["xHI"][ion], 1e-6)
    assert ok, err
    # n_e where n_HII = n_H (1 - x_HI) is resolved.  In unlit leaves at te_min
    # every ionising channel is charge exchange, 1 - x_HI ~ 1e-13 sits below
    # the 1e-15 absolute resolution of hydrogen_neutral_fraction's root, and
    # the reference algorithm itself returns n_e ~ 1e-7 that moves by factors
    # of 2 with the warm start (the oracle does, called from different
    # states); there only the bound n_e << n_H is a property of the solve
    resolved = 1.0 - so["xHI"] >= 1e-8
    assert resolved.sum() > 50
    np.testing.assert_allclose(sg["ne"][resolved], so["ne"][resolved], rtol=1e-4)
    dark = ~resolved
    assert np.all(sg["ne"][dark] < 1e-6 * sg["nH"][dark])
    assert np.all(so["ne"][dark] < 1e-6 * so["nH"][dark])
    np.testing.assert_allclose(mg, mo, rtol=1e-3, atol=5e-5)
    hot = so["Te"] > 5e3
    assert hot.sum() > 20


def test_dere_hybrid_thermal_update_matches_oracle():
    """ci_model 'dere_hybrid' (src/recomb_mod.f90:263-311): the H / He and
    every metal Voronov rate times the Dere ratio, in the pass's stage cache,
    the cell solves and the collisional-ionisation cooling of the thermal
    update -- T_e bit-equal to the oracle under the same switch, and the
    stage fractions of the collisionally ionised hot gas moved against the
    Voronov model."""
    n = 8
    out = {}
    for ci in (0, 1):
        params, nH = hii_thermal(n=n, nphotons=20000, metals=True)
        params["ci_model"] = ci
        orc, gm = oracle_model(params, nH), gpu_model(params, nH)
        orc.cooling_setup([a for a, _ in HI_COOLING_FIT], [t for _, t in HI_COOLING_FIT])
        gm.cooling_setup(HI_COOLING_FIT)
        x = _core(n)
        orc.set_state(xHI=x)
        orc.opacity_fill()
        gm.set_state(xHI=x)
        _pass(orc, gm)
        mo = orc.thermal_update()
        orc.opacity_fill()
        mg = gm.gas_thermal_update()
        so, sg = orc.state(), gm.state()
        rel = np.abs(sg["Te"] / so["Te"] - 1.0)
        print("ci_model", ci, "T_e: max rel diff", rel.max(), "bit-equal", np.mean(rel == 0.0))
        assert np.all(rel == 0.0)
        np.testing.assert_allclose(mg, mo, rtol=1e-3, atol=5e-5)
        els = elements_of(params)
        # the stage fractions follow n_e, which is resolved where n_HII is
        # (the dark leaves' n_e is noise-limited, as in the test above)
        resolved = 1.0 - so["xHI"] >= 1e-8
        assert resolved.sum() > 50
        for ie, el in enumerate(els):
            ok, err = rel_close(gm.species_fractions(ie)[resolved],
                                orc.species_fractions(ie)[resolved], 1e-4, floor_frac=1e-6)
            assert ok, (ci, el["name"], err)
        out[ci] = (sg["Te"], [gm.species_fractions(ie) for ie in range(len(els))])
    assert not np.array_equal(out[0][0], out[1][0])
    assert any(not np.allclose(a, b, rtol=1e-6, atol=0) for a, b in zip(out[0][1], out[1][1]))


def test_metal_warm_start_matches_bisection():
    """The safeguarded Newton warm start with the metal terms accepts the
    bisection's bracket: same T_e to the bracket width on nearly all leaves."""
    params, nH = hii_thermal(n=10, nphotons=30000, metals=True)
    a, b = gpu_model(params, nH), gpu_model(params, nH)
    for m in (a, b):
        m.cooling_setup(HI_COOLING_FIT)
        m.set_state(xHI=_core(10))
        m.ionizing_field_and_rates()
    a.set_warm_start(MC_WARM_NONE)
    b.set_warm_start(MC_WARM_NEWTON)
    a.gas_thermal_update()
    b.gas_thermal_update()
    sa, sb = a.state(), b.state()
    ta, tb = sa["Te"], sb["Te"]
    rel = np.abs(tb / ta - 1.0)
    print("warm vs bisection: max", rel.max(), "within 2.1e-5:", np.mean(rel < 2.1e-5),
          "stats", a.thermal_stats(), b.thermal_stats())
    assert rel.max() < 5e-3 and np.mean(rel < 2.1e-5) > 0.99
    for m in (a, b):              # the work counters are consistent
        st = m.thermal_stats()
        assert 0 < st["bracketed"] <= st["leaves"] == m.nleaf, st
        assert 4 * st["bracketed"] <= st["evaluations"] <= 64 * st["bracketed"], st
        assert 0 <= st["fallbacks"] <= st["bracketed"], st
    # leaves whose hydrogen is resolved: the same ionisation state; leaves
    # without it (where the warm form stops the n_e iteration once it is
    # inside its noise band) stay inside that band
    res = (1.0 - sa["xHI"] >= 1e-8) & (rel < 2.1e-5)
    assert res.sum() > 100
    ok, err = rel_close(1.0 - sb["xHI"][res], 1.0 - sa["xHI"][res], 1e-3, floor_frac=1e-6)
    assert ok, err
    ok, err = rel_close(sb["ne"][res], sa["ne"][res], 1e-3, floor_frac=1e-6)
    assert ok, err
    dark = 1.0 - sa["xHI"] < 1e-10
    assert np.all(1.0 - sb["xHI"][dark] < 1e-9)
    assert np.all(sb["ne"][dark] < 1e-6 * sb["nH"][dark])


@pytest.mark.parametrize("rows", [1, 2])
def test_energy_moment_tallies_match_per_coefficient_sums(rows):
    """MC_MIXED metal walk: the energy-moment tallies (four Chebyshev moments
    of the packet's ln-E sub-bin per segment, every coefficient sum rebuilt
    from its interpolant on the sub-bin) against the per-coefficient sums of
    the same walk, with the moment rows issued per lane (MC_ROWS_LANE) and
    transposed (MC_ROWS_TOGETHER).  The walk is the same (same segments);
    H/He rate sums, band J and the metal raw tallies agree to the
    interpolation tolerance."""
    params, nH = hii_thermal(n=10, nphotons=40000, metals=True, precision=MC_MIXED, diffuse=True)
    on, off = gpu_model(params, nH), gpu_model(params, nH)
    on.set_row_issue(rows)
    off.set_tally_moments(False)
    nsub, err = on.tally_moment_info()
    print("sub-bins", nsub, "interpolation error", err)
    assert nsub > 16 and err <= 1e-7
    x = _core(10)
    for m in (on, off):
        m.set_state(xHI=x)
        m.ionizing_field_and_rates()
    ta, tb = on.tallies(), off.tallies()
    assert ta["seg_direct"] == tb["seg_direct"] and ta["seg_scattered"] == tb["seg_scattered"]
    for k in ("hhe", "jt"):
        ok, e = rel_close(ta[k], tb[k], 1e-6, floor_frac=1e-9)
        assert ok, (k, e)
    n_tested = 0
    for ie, el in enumerate(elements_of(params)):
        a, b = on.species_tallies(ie), off.species_tallies(ie)
        ok, e = rel_close(a, b, 1e-6, floor_frac=1e-9)
        assert ok, (el["name"], e)
        n_tested += int((b > 0).sum())
    assert n_tested > 1000
    ra, rb = on.rates(), off.rates()
    for k in ("gHI", "gHeI", "hHI", "hHeI"):
        ok, e = rel_close(ra[k], rb[k], 1e-6, floor_frac=1e-9)
        assert ok, (k, e)
