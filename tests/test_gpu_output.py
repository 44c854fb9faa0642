"""The rates file of a converged model (gas_rates_write,
src/gas_rates_mod.f90:216-366, with species_write): every HDU read back by an
independent FITS reader equals the device model's arrays -- rates, state,
J_nu (jt_ion over 4 pi V dnu d2cm^2), band centres and widths, leaf
geometry, the metal stage blocks at the written state -- and the run
keywords carry the ledger and convergence record."""
import numpy as np
import pytest

import mochii_amd as mc
from mochii_amd.output import gas_rates_write
from mochii_amd.workloads import HI_COOLING_FIT, elements_of, hii_octree, hii_thermal
from tests.cases import gpu_model, read_fits_images

pytestmark = pytest.mark.gpu


def _check(path, m, run, els):
    hdus = {h["EXTNAME"]: (h, a) for h, a in read_fits_images(path)}
    h0 = read_fits_images(path)[0][0]
    assert h0["EXTNAME"] == "Gamma_HI" and h0["CONVERGD"] == run["converged"]
    assert h0["NITERDON"] == run["iterations"]
    np.testing.assert_array_equal([h0["L_EMIT"], h0["L_ABS"], h0["L_ESC"]], m.energy_ledger())
    r, s = m.rates(), m.state()
    for key, name in (("gHI", "Gamma_HI"), ("gHeI", "Gamma_HeI"), ("gHeII", "Gamma_HeII"),
                      ("hHI", "Heat_HI"), ("hHeI", "Heat_HeI"), ("hHeII", "Heat_HeII")):
        np.testing.assert_array_equal(hdus[name][1], r[key])
    for key, name in (("xHI", "x_HI"), ("xHeI", "x_HeI"), ("xHeII", "x_HeII"), ("ne", "n_e"),
                      ("Te", "T_e")):
        np.testing.assert_array_equal(hdus[name][1], s[key])
    e = m.band_edges()
    np.testing.assert_allclose(hdus["E_bin"][1], np.sqrt(e[:-1] * e[1:]), rtol=1e-15)
    np.testing.assert_allclose(hdus["dE_bin"][1], np.diff(e), rtol=1e-15)
    size = hdus["LeafSize"][1]
    xyz = hdus["LeafXYZ"][1]
    assert xyz.shape == (3, m.nleaf) and size.shape == (m.nleaf,)
    box = 2.0 * m.params.xmax
    assert np.sum(size ** 3) == pytest.approx(box ** 3, rel=1e-12)     # leaves tile the box
    assert np.all(np.abs(xyz) + 0.5 * size[None, :] <= 0.5 * box * (1 + 1e-12))
    jt = m.tallies()["jt"]
    dnu = np.diff(e) * 1.602176634e-12 / 6.62607015e-27
    want = jt / (4 * np.pi * size[:, None] ** 3 * dnu[None, :] * m.params.distance2cm ** 2)
    np.testing.assert_allclose(hdus["J_nu"][1], want, rtol=1e-14, atol=0)
    for ie, el in enumerate(els):
        fr = hdus[f"x_{el['name']}_stages"]
        assert fr[0]["ABUND"] == el["abund"] and fr[1].shape == (m.nleaf, el["nstage"])
        np.testing.assert_allclose(fr[1].sum(axis=1), 1.0, rtol=1e-12)
        g, h = m.species_rates(ie)
        np.testing.assert_array_equal(hdus[f"Gamma_{el['name']}_stages"][1], g)
        np.testing.assert_array_equal(hdus[f"Heat_{el['name']}_stages"][1], h)
    return hdus


def test_rates_file_of_a_converged_car_model(tmp_path):
    params, nH = hii_thermal(n=8, nphotons=20000, niter=6,
                             cooling_model=mc.lib.MC_COOL_LOW_DENSITY)
    m = gpu_model(params, nH)
    m.cooling_setup(HI_COOLING_FIT)
    run = mc.run_model(m)
    path = gas_rates_write(m, str(tmp_path), "hii.fits.gz", run=run)
    hdus = _check(path, m, run, [])
    # raster order: leaf 1 is the (-x, -y, -z) corner, leaf 2 its +x neighbour
    xyz = hdus["LeafXYZ"][1]
    dc = 2 * params["xmax"] / 8
    np.testing.assert_allclose(xyz[:, 0], [-params["xmax"] + 0.5 * dc] * 3)
    assert xyz[0, 1] - xyz[0, 0] == pytest.approx(dc)


def test_rates_file_of_an_octree_model_with_metals(tmp_path):
    params, nH, octree = hii_octree(n_field=16, lmax=4, nphotons=20000, niter=3, diffuse=False,
                                    cooling_model=mc.lib.MC_COOL_LOW_DENSITY)
    els = elements_of(params)
    assert els
    m = gpu_model(params, nH, octree=octree)
    m.cooling_setup(HI_COOLING_FIT)
    run = mc.run_model(m)
    path = gas_rates_write(m, str(tmp_path), "oct", run=run)
    _check(path, m, run, els)
