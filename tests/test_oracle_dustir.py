"""Pins of the oracle's dust thermal-infrared band (oracle/orc_dustir.c and
the infrared pass in oracle/orc_transport.c).

The reference's modules for this band (src/dust_temp_mod.f90,
src/dust_emis_transport_mod.f90, src/dust_ir_band_mod.f90) need MPI and the
grain model, so they are not compiled here, and the reference's tests hold
no golden values for them.  The restatement is pinned by what the
mathematics fixes instead:
  - femit(T) of a grey table is 4 s_abs sigma_SB T^4 (Stefan-Boltzmann), so a
    dropped 4 pi, a reversed frequency interval or a wrong Planck constant
    fails; femit_of is exact at the knots and log-linear between them;
  - dust_temp_compute of a grey table inverts that closed form;
  - the emission grid and the band cut follow from the band floor and the
    active gas thresholds;
  - sample_linear_spectrum_bin inverts its own linear CDF, with the three
    limiting forms its comment names;
  - the stratified packet counts sum to N and round N L_il / L_tot;
  - the launched wavelengths follow the leaf's piecewise-linear density
    (chi-square over bins, half-bin fractions within bins), positions stay in
    the leaf, directions are isotropic;
  - the transport conserves energy (closure <= 5e-13, the reference's abort
    threshold, src/dust_emis_transport_mod.f90:320-324), the path estimator
    equals the extinction ledger without scattering, the escaped spectrum sums
    to the escaped ledger, and the reabsorbed fraction of a uniform grey cube
    matches an independent Monte Carlo of 1 - <exp(-kappa l)>.
"""
import numpy as np
import pytest

from mochii_amd.atomic import registry, metal_thresholds
from mochii_amd.workloads import stromgren
from oracle import pyoracle as po
from tests.cases import oracle_model

SIGMA_SB = 5.670374419e-5
ALB = 0.3


def grey_table(n=4000, cext=1.0e-21, alb=ALB, g=0.5):
    lam = np.geomspace(1.0e-3, 1.0e5, n)
    return lam, np.full(n, alb), np.full(n, g), np.full(n, cext)


def _model(n=8, table=None, scatter=0, eion_min=13.598, elements=(), rk=0.1, heat=None,
           nphotons=1000):
    params, nH = stromgren(n=n, nphotons=nphotons)
    params.update(ion_add_dust=1, ion_dust_scatter=scatter, lambda_ref=0.55,
                  eion_min=eion_min)
    if elements:
        params["metal_eth"] = metal_thresholds(elements)
    rhokap = np.full(nH.size, rk)                   # per code length (pc)
    m = oracle_model(params, nH, rhokap=rhokap,
                     dust_table=grey_table() if table is None else table,
                     elements=list(elements))
    return m


def _vol(m):
    h = m._head()
    return (h.dcell * m.params["distance2cm"]) ** 3


def test_femit_grey_is_stefan_boltzmann():
    m = _model()
    m.dust_temp_setup()
    a = m.ir_arrays()
    T = a["tgrid"]
    np.testing.assert_allclose(T, 10.0 ** np.linspace(0.5, 3.5, 61), rtol=1e-14)
    np.testing.assert_allclose(a["femit"], 4.0 * (1.0 - ALB) * SIGMA_SB * T ** 4, rtol=5e-5)


def test_femit_of_interpolates_log_log():
    m = _model()
    m.dust_temp_setup()
    a = m.ir_arrays()
    T, f = a["tgrid"], a["femit"]
    for i in range(0, 61, 7):
        assert m.femit_of(T[i]) == pytest.approx(f[i], rel=1e-13)
    for i in range(0, 60, 6):
        tm = np.sqrt(T[i] * T[i + 1])
        assert m.femit_of(tm) == pytest.approx(np.sqrt(f[i] * f[i + 1]), rel=1e-12)
    assert m.femit_of(1.0) == pytest.approx(f[0], rel=1e-13)
    assert m.femit_of(1.0e5) == pytest.approx(f[-1], rel=1e-13)


def test_dust_temp_inverts_grey_closed_form():
    m = _model(n=8)
    m.dust_temp_setup()
    rng = np.random.default_rng(3)
    kd = 0.1 / m.params["distance2cm"]
    Ttrue = 10.0 ** rng.uniform(0.7, 3.3, 512)
    heat = 4.0 * (1.0 - ALB) * SIGMA_SB * Ttrue ** 4 * kd
    heat[:5] = 0.0
    heat[5:8] = -1.0
    heat[8] = 1.0e30                           # above femit(tgrid(NT)) kd
    td = m.dust_temp_compute(heat)
    np.testing.assert_array_equal(td[:8], 0.0)
    assert td[8] == pytest.approx(10.0 ** 3.5, rel=1e-14)
    np.testing.assert_allclose(td[9:], Ttrue[9:], rtol=2e-5)


def test_emission_grid_edges_and_band_cut():
    m = _model()
    m.dust_emis_transport_setup()
    a = m.ir_arrays()
    lam = a["lam"]
    assert lam.size == 300 and a["nzero"] == 0
    assert lam[0] == pytest.approx(0.1, rel=1e-15) and lam[-1] == pytest.approx(3000.0, rel=1e-13)
    np.testing.assert_allclose(lam[1:] / lam[:-1], (3.0e4) ** (1 / 299), rtol=1e-12)
    np.testing.assert_allclose(a["sabs"], 1.0 - ALB, rtol=1e-14)
    e = a["edge"]
    np.testing.assert_allclose(e[1:-1], np.sqrt(lam[1:] * lam[:-1]), rtol=1e-15)
    assert e[0] < lam[0] < e[1] and e[-2] < lam[-1] < e[-1]
    # a band floor below 12.4 eV moves the grid start to it
    m6 = _model(eion_min=6.0)
    m6.dust_emis_transport_setup()
    assert m6.ir_arrays()["lam"][0] == pytest.approx(1.23984 / 6.0, rel=1e-15)


@pytest.mark.parametrize("abund", [{}, {"s": 0.0}])
def test_band_cut_clears_the_metal_thresholds(abund):
    els = registry(abund)
    m = _model(elements=els)
    m.dust_emis_transport_setup()
    a = m.ir_arrays()
    emin = min(metal_thresholds(els))           # S I (10.36 eV) or C I (11.26 eV)
    k = a["nzero"]                              # nodes 1..kpass zeroed
    lam = a["lam"]
    assert k > 1
    assert 1.23984 / lam[k - 1] < emin <= 1.23984 / lam[k - 2]
    assert a["lam_hardest"] == lam[k - 1]


def test_sample_linear_bin_inverts_its_cdf():
    rng = np.random.default_rng(7)
    for _ in range(300):
        p0, p1 = rng.uniform(0, 2, 2)
        u = rng.uniform()
        lam = po.sample_linear_bin(2.0, 3.0, p0, p1, u)
        t = lam - 2.0
        F = (p0 * t + (p1 - p0) * t * t / 2) / ((p0 + p1) / 2)
        assert abs(F - u) < 1e-12
    # u = 0 with p0 = 0 is 0/0 in the reference's root; the launch never
    # passes it (u > cdf(k) strictly in the drawn bin, so ubin > 0)
    for u in (1e-9, 0.1, 0.5, 0.97, 1.0):
        assert po.sample_linear_bin(1.0, 2.0, 0.7, 0.7, u) == pytest.approx(1.0 + u, abs=1e-15)
        assert po.sample_linear_bin(1.0, 2.0, 0.0, 0.4, u) == pytest.approx(1.0 + np.sqrt(u), abs=1e-15)
        assert po.sample_linear_bin(1.0, 2.0, 0.4, 0.0, u) == pytest.approx(2.0 - np.sqrt(1 - u), abs=1e-15)
        assert po.sample_linear_bin(1.0, 2.0, 0.0, 0.0, u) == pytest.approx(1.0 + u, abs=1e-15)
    assert po.sample_linear_bin(1.0, 2.0, 0.3, 0.5, -0.2) == 1.0
    assert po.sample_linear_bin(1.0, 2.0, 0.3, 0.5, 1.5) == 2.0


def test_stratified_counts_sum_to_n():
    m = _model(n=8)
    m.dust_temp_setup()
    m.dust_emis_transport_setup()
    rng = np.random.default_rng(11)
    heat = rng.lognormal(0.0, 2.0, 512) * 1e-20
    heat[rng.choice(512, 60, replace=False)] = 0.0
    for N in (1000, 123457):
        L, xi, khi = m.ir_counts(1, N, heat)
        assert 1e-12 <= xi <= 1 - 1e-12
        cnt = np.diff(np.concatenate([[0], khi]))
        assert cnt.sum() == N and khi[-1] == N
        expect = N * heat * _vol(m) / L
        assert np.all(np.abs(cnt - expect) < 1.0)
        assert np.all(cnt[heat == 0.0] == 0)
        assert L == pytest.approx(heat.sum() * _vol(m), rel=1e-13)


def test_launch_positions_directions_and_wavelength_density():
    m = _model(n=8, table=po_table_powerlaw())
    m.dust_temp_setup()
    heat = np.full(512, 3.0e-19)
    m.dust_temp_compute(heat)
    m.dust_emis_transport_setup()
    il = 100
    pdf, cdf, esum, _ = m.ir_leaf_spectrum(il)
    assert esum == cdf[-1] and esum > 0
    h = m._head()
    ic = il
    ix, iy, iz = ic % 8, (ic // 8) % 8, ic // 64
    c0 = np.array([h.xmin + (ix + 0.5) * h.dcell, h.ymin + (iy + 0.5) * h.dcell,
                   h.zmin + (iz + 0.5) * h.dcell])
    M = 40000
    out = np.array([m.ir_launch(il, pdf, cdf, (1 << 40) | k) for k in range(1, M + 1)])
    assert np.all(np.abs(out[:, :3] - c0) <= 0.5 * h.dcell * (1 + 1e-15))
    k = out[:, 3:6]
    np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, rtol=1e-14)
    assert np.all(np.abs(k.mean(0)) < 4 / np.sqrt(3 * M))
    assert abs((k[:, 2] ** 2).mean() - 1 / 3) < 4 * np.sqrt(4 / 45 / M)
    lam = m.ir_arrays()["lam"]
    ln = out[:, 6]
    assert np.all((ln >= lam[0]) & (ln <= lam[-1]))
    bins = np.searchsorted(lam, ln, side="right") - 1
    p = np.diff(cdf) / cdf[-1]
    obs = np.bincount(bins, minlength=299)[:299]
    big = p * M >= 20
    chi2 = ((obs[big] - M * p[big]) ** 2 / (M * p[big])).sum()
    dof = big.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
    assert obs[~big].sum() < 20 + M * p[~big].sum() * 2
    # inside the bins: the fraction below each bin's midpoint follows the
    # linear density F(1/2) = (p0/2 + (p1 - p0)/8) / ((p0 + p1)/2)
    t = (ln - lam[bins]) / (lam[bins + 1] - lam[bins])
    p0, p1 = pdf[bins], pdf[bins + 1]
    F = (0.5 * p0 + (p1 - p0) / 8) / ((p0 + p1) / 2)
    below = (t < 0.5).astype(float)
    assert abs(below.sum() - F.sum()) < 5 * np.sqrt((F * (1 - F)).sum())


def po_table_powerlaw():
    lam = np.geomspace(1.0e-3, 1.0e5, 800)
    return lam, np.full(800, 0.2), np.full(800, 0.5), 1e-21 * (lam / 0.55) ** -1.5


def _run(m, heat, N=50000, niter=2, tol=1e-12):
    m.dust_temp_setup()
    m.dust_temp_compute(heat)
    m.dust_emis_transport_setup()
    return m.transport_dust_emission(heat, npacket=N, niter=niter, tol=tol, nthreads=4)


@pytest.mark.parametrize("scatter", [0, 1])
def test_ir_pass_conserves_energy_and_estimators_agree(scatter):
    m = _model(n=8, table=po_table_powerlaw(), scatter=scatter, rk=0.3)
    rng = np.random.default_rng(5)
    heat = rng.lognormal(0.0, 1.0, 512) * 1e-19
    out, rep = _run(m, heat, N=40000, niter=1)
    a = m.ir_arrays()
    assert rep["passes"] == 1
    assert rep["closure"] <= 5e-13
    assert rep["emitted"] == pytest.approx(rep["L_emit"], rel=1e-12)
    assert rep["L_emit"] == pytest.approx(heat.sum() * _vol(m), rel=1e-13)
    assert a["lesc"].sum() == pytest.approx(rep["escaped"], rel=1e-12)
    np.testing.assert_allclose(out, heat + a["heat_ir"], rtol=1e-15)
    np.testing.assert_array_equal(a["tdust"], m.dust_temp_compute(out))
    if scatter:
        assert a["seg_scattered"] > 0
        assert rep["L_abs_path"] < rep["extinguished"]
    else:
        assert a["seg_scattered"] == 0
        assert rep["L_abs_path"] == pytest.approx(rep["extinguished"], rel=1e-12)
    assert 0.0 < rep["extinguished"] < rep["emitted"]


def test_reabsorbed_fraction_of_a_uniform_grey_cube():
    """Uniform emission in a uniform grey cube without scattering: the
    extinguished fraction is 1 - <exp(-kappa_abs l)> over uniform launch
    points and isotropic directions, l the distance to the box face."""
    rk = 0.15
    m = _model(n=8, rk=rk)
    heat = np.full(512, 2.0e-19)
    N = 100000
    out, rep = _run(m, heat, N=N, niter=1)
    frac = rep["extinguished"] / rep["emitted"]
    kap = rk * (1.0 - ALB)
    h = m._head()
    lo, hi = h.xmin, h.xmaxb
    rng = np.random.default_rng(17)
    M = 2_000_000
    x = rng.uniform(lo, hi, (M, 3))
    mu = rng.uniform(-1, 1, M)
    ph = rng.uniform(0, 2 * np.pi, M)
    s = np.sqrt(1 - mu * mu)
    k = np.stack([s * np.cos(ph), s * np.sin(ph), mu], 1)
    with np.errstate(divide="ignore"):
        d = np.where(k > 0, (hi - x) / k, np.where(k < 0, (lo - x) / k, np.inf))
    ell = d.min(1)
    f = 1.0 - np.exp(-kap * ell)
    sig = f.std() * np.sqrt(1.0 / N + 1.0 / M)
    assert abs(frac - f.mean()) < 4 * sig, (frac, f.mean(), sig)


def test_self_heating_iteration_converges():
    m = _model(n=8, table=po_table_powerlaw(), rk=0.5)
    heat = np.full(512, 1.0e-19)
    _, r1 = _run(m, heat.copy(), N=30000, niter=1)
    m2 = _model(n=8, table=po_table_powerlaw(), rk=0.5)
    out, r3 = _run(m2, heat.copy(), N=30000, niter=4, tol=1e-12)
    assert r1["passes"] == 1 and r3["passes"] == 4 and r3["converged"] == 0
    # pass 1's change is the reabsorbed fraction itself; later ones shrink
    assert r1["drel"] == pytest.approx(r1["L_abs_path"] / (r1["L_emit"] + r1["L_abs_path"]),
                                       rel=1e-10)
    assert r3["drel"] < 0.3 * r1["drel"]
    m3 = _model(n=8, table=po_table_powerlaw(), rk=0.5)
    _, rc = _run(m3, heat.copy(), N=30000, niter=5, tol=0.5)
    assert rc["passes"] == 1 and rc["converged"] == 1
