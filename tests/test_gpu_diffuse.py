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
