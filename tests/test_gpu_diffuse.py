"""HIP diffuse ground-recombination field against the FP64 oracle, through
the C-ABI: same emission build from the same state, same stream-2 Sobol
launch, same Milne spectra, same walk."""
import numpy as np
import pytest

from mochii_amd import MC_FP64, MC_MIXED
from mochii_amd.workloads import hii_thermal, lognormal_field, octree_from_field
from tests.cases import gpu_model, oracle_model, rel_close

pytestmark = pytest.mark.gpu


def _ionised_state(n, nl, seed=4, octree_centres=None):
    """A partly ionised H + He state with a warm ionised core."""
    rng = np.random.default_rng(seed)
    if octree_centres is None:
        c = (np.arange(n) + 0.5) / n * 2 - 1
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        r = np.sqrt(X**2 + Y**2 + Z**2).transpose(2, 1, 0).ravel()
    else:
        r = np.linalg.norm(octree_centres, axis=1)
    core = r < 0.55
    xHI = np.where(core, 10 ** rng.uniform(-4, -2.5, nl), 1.0)
    xHeI = np.where(core, 10 ** rng.uniform(-3, -1, nl), 1.0)
    xHeII = np.where(core, rng.uniform(0.3, 0.9, nl), 0.0)
    Te = np.where(core, rng.uniform(8e3, 2e4, nl), 1e3)
    ne = 100.0 * ((1 - xHI) + 0.1 * (xHeII + 2 * np.maximum(0, 1 - xHeI - xHeII)))
    return dict(xHI=xHI, xHeI=xHeI, xHeII=xHeII, Te=Te, ne=ne)


def _run_pair(params, nH, state, octree=None, nthreads=8):
    orc = oracle_model(params, nH, octree=octree)
    gm = gpu_model(params, nH, octree=octree)
    orc.set_state(**state)
    gm.set_state(**state)
    orc.transport_pass(nthreads=nthreads)
    orc.rates_compute()
    gm.ionizing_field_and_rates()
    return orc, gm


@pytest.mark.parametrize("prec", [MC_FP64, MC_MIXED])
def test_diffuse_pass_car_matches_oracle(prec):
    n = 16
    params, nH = hii_thermal(n=n, nphotons=20000, precision=prec, diffuse=True)
    orc, gm = _run_pair(params, nH, _ionised_state(n, n ** 3))
    h = orc._head()
    L, nd = gm.diffuse_info()
    assert nd == h.diffuse_nphot and nd > 5000
    assert abs(L / h.diffuse_lum - 1.0) < 1e-12
    to, tg = orc.tallies(), gm.tallies()
    tol = 1e-9 if prec == MC_FP64 else 1e-4
    for k in ("hhe", "jt"):
        ok, err = rel_close(tg[k], to[k], tol, floor_frac=1e-5 if prec == MC_MIXED else 1e-6)
        assert ok, (k, err)
    if prec == MC_FP64:
        assert tg["seg_direct"] == to["seg_direct"]
    led = gm.energy_ledger()
    np.testing.assert_allclose(led, to["ledger"], rtol=1e-12 if prec == MC_FP64 else 1e-5)
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]


def test_diffuse_pass_octree_matches_oracle():
    field = lognormal_field(16, mean=100.0, sigma_ln=0.5, seed=2)
    xl, yl, zl, lev, nHl = octree_from_field(field, half=8.0, lmax=5, thresh_dex=0.3, lmin=2)
    oct_ = (xl, yl, zl, lev, 8.0)
    params, _ = hii_thermal(n=16, nphotons=20000, diffuse=True, octree=True)
    cen = np.stack([xl, yl, zl], axis=1) / 8.0
    state = _ionised_state(0, xl.size, octree_centres=cen)
    state["ne"] = nHl / 100.0 * state["ne"]
    orc, gm = _run_pair(params, nHl, state, octree=oct_)
    h = orc._head()
    L, nd = gm.diffuse_info()
    assert nd == h.diffuse_nphot and nd > 1000
    to, tg = orc.tallies(), gm.tallies()
    for k in ("hhe", "jt"):
        ok, err = rel_close(tg[k], to[k], 1e-9)
        assert ok, (k, err)
    led = gm.energy_ledger()
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]


def test_no_diffuse_packets_from_neutral_gas():
    params, nH = hii_thermal(n=8, nphotons=2000, diffuse=True)
    gm = gpu_model(params, nH)
    gm.ionizing_field_and_rates()
    assert gm.diffuse_info() == (0.0, 0)


@pytest.mark.parametrize("prec", [MC_FP64, MC_MIXED])
def test_hei_diffuse_pass_car_matches_oracle(prec):
    """par%hei_diffuse: the He I excited channel (branch lines and the
    two-photon pair) in the emission build and the launch, against the
    oracle; the two-photon tables equal the oracle's bit for bit."""
    from mochii_amd.atomic import load_hei_two_photon
    from oracle import pyoracle as po
    n = 16
    params, nH = hii_thermal(n=n, nphotons=20000, precision=prec, diffuse=True)
    params["hei_diffuse"] = 1
    state = _ionised_state(n, n ** 3)
    orc, gm = _run_pair(params, nH, state)
    h = po.hei_two_photon(*load_hei_two_photon(), eion_max=params["eion_max"])
    nper, eper, gy, cdf = gm.diffuse_hei_info()
    assert (nper, eper) == (h.n_per_decay, h.e_per_decay)
    assert np.array_equal(gy, np.asarray(h.grid_y)) and np.array_equal(cdf, np.asarray(h.cdf_e))
    hd = orc._head()
    L, nd = gm.diffuse_info()
    assert nd == hd.diffuse_nphot and nd > 5000
    assert abs(L / hd.diffuse_lum - 1.0) < 1e-12
    ch, _ = orc.diffuse_tables()
    assert ch[:, 3].sum() > 0.05 * L          # the channel carries a real share
    to, tg = orc.tallies(), gm.tallies()
    tol = 1e-9 if prec == MC_FP64 else 1e-4
    for k in ("hhe", "jt"):
        ok, err = rel_close(tg[k], to[k], tol, floor_frac=1e-5 if prec == MC_MIXED else 1e-6)
        assert ok, (k, err)
    if prec == MC_FP64:
        assert tg["seg_direct"] == to["seg_direct"]
    led = gm.energy_ledger()
    np.testing.assert_allclose(led, to["ledger"], rtol=1e-12 if prec == MC_FP64 else 1e-5)
    assert abs(led[1] + led[2] - led[0]) <= 5e-13 * led[0]


def test_hei_diffuse_pass_octree_matches_oracle():
    field = lognormal_field(16, mean=100.0, sigma_ln=0.5, seed=2)
    xl, yl, zl, lev, nHl = octree_from_field(field, half=8.0, lmax=5, thresh_dex=0.3, lmin=2)
    oct_ = (xl, yl, zl, lev, 8.0)
    params, _ = hii_thermal(n=16, nphotons=20000, diffuse=True, octree=True)
    params["hei_diffuse"] = 1
    cen = np.stack([xl, yl, zl], axis=1) / 8.0
    state = _ionised_state(0, xl.size, octree_centres=cen)
    state["ne"] = nHl / 100.0 * state["ne"]
    orc, gm = _run_pair(params, nHl, state, octree=oct_)
    h = orc._head()
    L, nd = gm.diffuse_info()
    assert nd == h.diffuse_nphot and nd > 1000
    assert abs(L / h.diffuse_lum - 1.0) < 1e-12
    to, tg = orc.tallies(), gm.tallies()
    for k in ("hhe", "jt"):
        ok, err = rel_close(tg[k], to[k], 1e-9)
        assert ok, (k, err)


def test_hei_diffuse_without_table_is_an_error():
    """A pass with hei_diffuse on and no two-photon table fails before any
    packet is launched (the reference aborts in hei_2ph_setup)."""
    import ctypes as C
    from mochii_amd.lib import Params, library
    params, nH = hii_thermal(n=8, nphotons=2000, diffuse=True)
    params["hei_diffuse"] = 1
    L = library()
    h = C.c_void_p()
    assert L.mc_init(C.byref(Params.make(**params)), C.byref(h)) == 0
    try:
        nH = np.ascontiguousarray(nH)
        assert L.mc_gas_state_setup(h, nH.ctypes.data_as(C.POINTER(C.c_double)), None) == 0
        assert L.mc_ionizing_field_and_rates(h, 0) != 0
        assert b"mc_diffuse_hei_setup" in L.mc_last_error()
    finally:
        L.mc_finalize(h)
