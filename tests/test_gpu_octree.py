"""HIP octree (face-exit) walk against the FP64 oracle on refined trees
(src/raytrace_amr.f90:348-486, src/octree_mod.f90:309-559)."""
import numpy as np
import pytest

from mochii_amd import MC_FP64, MC_GRID_AMR, MC_MIXED
from mochii_amd.workloads import lognormal_field, octree_from_field, stromgren, synthetic_kext
from tests.cases import gpu_model, oracle_model, rel_close

pytestmark = pytest.mark.gpu


def _tree(n=16, lmax=4, seed=2, sigma=0.8):
    f = lognormal_field(n, 100.0, sigma, seed=seed)
    x, y, z, lev, nh = octree_from_field(f, 4.0, lmax=lmax, thresh_dex=0.3)
    return (x, y, z, lev, 4.0), nh


def _models(nphot=20000, precision=MC_FP64, dust=False, **tree_kw):
    octree, nh = _tree(**tree_kw)
    params, _ = stromgren(n=2, nphotons=nphot, precision=precision)
    params.update(grid_type=MC_GRID_AMR)
    extra = {}
    if dust:
        params.update(ion_add_dust=1, ion_dust_scatter=1, lambda_ref=0.55)
        extra = dict(rhokap=nh * 4.868e-22 * 3.0856776e18, dust_table=synthetic_kext())
    orc = oracle_model(params, nh, octree=octree, **extra)
    gm = gpu_model(params, nh, octree=octree, **extra)
    return octree, nh, orc, gm


def test_octree_fixed_rays_bit_exact():
    octree, nh, orc, gm = _models(nphot=10)
    rng = np.random.default_rng(4)
    rays = [((0.0, 0.0, 0.0), (0.3, -0.5, 0.81)), ((0.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
            ((0.01, 0.02, -0.03), (-0.6, 0.0, 0.8))]
    for _ in range(60):
        k = rng.normal(size=3)
        rays.append((tuple(rng.uniform(-3.9, 3.9, 3)), tuple(k / np.linalg.norm(k))))
    for pos, d in rays:
        want, _ = orc.trace_ray(*pos, *d)
        for prec in (MC_FP64, MC_MIXED):
            got = gm.trace_ray(pos, d, precision=prec)
            assert np.array_equal(got, want), (prec, pos, d, got[:6], want[:6])


@pytest.mark.parametrize("state", ["neutral", "partial"])
def test_octree_pass_fp64_matches_oracle(state):
    octree, nh, orc, gm = _models()
    if state == "partial":
        x, y, z = octree[:3]
        xHI = np.where(np.sqrt(x**2 + y**2 + z**2) < 2.5, 1e-3, 1.0)
        orc.set_state(xHI=xHI)
        gm.set_state(xHI=xHI)
    orc.transport_pass(nthreads=4)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    to, tg = orc.tallies(), gm.tallies()
    for key in ("hhe", "jt"):
        ok, err = rel_close(tg[key], to[key], 1e-9)
        assert ok, (key, err)
    assert tg["seg_direct"] == to["seg_direct"]
    np.testing.assert_allclose(gm.energy_ledger(), to["ledger"], rtol=1e-12)
    ok, err = rel_close(gm.rates()["gHI"], orc.rates()["gHI"], 1e-9)
    assert ok, err


def test_octree_pass_mixed_matches_oracle():
    octree, nh, orc, gm = _models(precision=MC_MIXED)
    x, y, z = octree[:3]
    xHI = np.where(np.sqrt(x**2 + y**2 + z**2) < 2.5, 1e-3, 1.0)
    orc.set_state(xHI=xHI)
    gm.set_state(xHI=xHI)
    orc.transport_pass(nthreads=4)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    ro, rg = orc.rates(), gm.rates()
    for k in ("gHI", "hHI"):
        ok, err = rel_close(rg[k], ro[k], 1e-4, floor_frac=1e-4)
        assert ok, (k, err)
        assert abs(rg[k].sum() / ro[k].sum() - 1) < 1e-5


def test_octree_dust_scattering_matches_oracle():
    octree, nh, orc, gm = _models(dust=True)
    x, y, z = octree[:3]
    xHI = np.where(np.sqrt(x**2 + y**2 + z**2) < 2.5, 1e-3, 1.0)
    orc.set_state(xHI=xHI)
    gm.set_state(xHI=xHI)
    orc.transport_pass(nthreads=4)
    gm.ionizing_field_and_rates()
    to, tg = orc.tallies(), gm.tallies()
    assert to["seg_scattered"] > 0
    assert abs(tg["seg_scattered"] - to["seg_scattered"]) <= 1e-3 * to["seg_scattered"]
    for key in ("hhe", "jt"):
        assert abs(tg[key].sum() / to[key].sum() - 1) < 1e-6
    np.testing.assert_allclose(gm.energy_ledger(), to["ledger"], rtol=1e-12)


@pytest.mark.parametrize("prec", [MC_FP64, MC_MIXED])
def test_car_face_exit_walk_matches_oracle(prec):
    """par%car_walk = 'amr' on a car grid: the face-exit walk over raster
    cells (src/define.f90:665-672)."""
    n = 16
    params, nH = stromgren(n=n, nphotons=20000, precision=prec)
    params.update(car_walk_amr=1)
    orc, gm = oracle_model(params, nH), gpu_model(params, nH)
    orc.transport_pass(nthreads=4)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    tol = 1e-9 if prec == MC_FP64 else 1e-4
    ok, err = rel_close(gm.rates()["gHI"], orc.rates()["gHI"], tol, floor_frac=1e-4)
    assert ok, err
