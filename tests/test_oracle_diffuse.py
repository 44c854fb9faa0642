"""Oracle pins for the explicit diffuse recombination field (Milne ground
continua, per-leaf emission, diffuse launch).

Independent checks: the reference's milne_recomb_spectrum_mod compiled from
source (bit for bit), the gates of the reference's tests/milne/check_milne.f90
redone against scipy quadrature of the Milne integrand, the Milne relation
for alpha_1 against alpha_A - alpha_B, and the photon-rate identity that the
equal-energy packets represent exactly n_e n_ion alpha_1 V recombinations."""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle import pyoracle as po

KB, EV = 1.380649e-16, 1.602176634e-12
ETH = (13.598, 24.587, 54.416)
EMAX = 100.0


def _ref():
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    return ref


def _sigma(ich, E):
    L = po.lib()
    return (L.orc_sigma_HI, L.orc_sigma_HeI, L.orc_sigma_HeII)[ich](E)


def test_milne_tables_equal_reference_module(capfd):
    ref = _ref()
    ref.ref_milne_setup(EMAX, 1e3, 5e4)
    capfd.readouterr()
    m = po.Milne(EMAX)
    rng = np.random.default_rng(5)
    for ich in range(3):
        for T in np.concatenate([np.geomspace(50.0, 2e6, 60), [1e2, 1e4, 1e6]]):
            assert m.energy_per_recomb(ich, T) == ref.ref_milne_eper(ich + 1, T), (ich, T)
            for u in np.concatenate([rng.random(20), [0.0, 1e-12, 0.5, 1.0 - 1e-12, 1.0]]):
                assert m.sample(ich, T, u) == ref.ref_milne_sample(ich + 1, T, u), (ich, T, u)


def _quad_moments(ich, T):
    """Photon-number integrals of n(E) = E^2 sigma(E) exp(-(E-E_th)/kT) over the
    band and beyond: (in-band N, in-band N<E>, in-band N<E^2>, total N)."""
    kT = KB * T / EV
    e0 = ETH[ich]

    def n(E):
        return E * E * _sigma(ich, E) * np.exp(-(E - e0) / kT)

    top = min(EMAX, e0 + 60 * kT)
    pts = [e0 + kT * k for k in (0.5, 2, 5, 10, 20) if e0 + kT * k < top]
    kw = dict(limit=500, epsrel=1e-11, points=pts or None)
    N = quad(n, e0, top, **kw)[0]
    NE = quad(lambda E: E * n(E), e0, top, **kw)[0]
    NE2 = quad(lambda E: E * E * n(E), e0, top, **kw)[0]
    Ntot = N + (quad(n, top, e0 + 60 * kT, limit=500)[0] if top < e0 + 60 * kT else 0.0)
    return N, NE, NE2, Ntot


@pytest.mark.parametrize("ich", [0, 1, 2])
@pytest.mark.parametrize("T", [3.0e3, 8.0e3, 1.5e4, 3.0e4])
def test_milne_sampler_gates(ich, T):
    """tests/milne/check_milne.f90 of the reference: range, energy-weighted
    mean, <1/E> <E>_n = 1, median; plus the in-band energy per recombination
    against quadrature."""
    m = po.Milne(EMAX)
    N, NE, NE2, Ntot = _quad_moments(ich, T)
    mean_n = NE / N                      # photon-number mean energy
    mean_e = NE2 / NE                    # energy-luminosity-weighted mean
    ns = 20000
    E = np.array([m.sample(ich, T, (i + 0.5) / ns) for i in range(ns)])
    assert E.min() >= ETH[ich] - 1e-9 and E.max() <= EMAX
    assert abs(E.mean() / mean_e - 1.0) < 2e-4
    assert abs(np.mean(1.0 / E) * mean_n - 1.0) < 2e-4
    # median of the energy-weighted distribution from the quadrature CDF
    kT = KB * T / EV
    e0 = ETH[ich]
    med = np.median(E)
    below = quad(lambda x: x * x * x * _sigma(ich, x) * np.exp(-(x - e0) / kT), e0, med,
                 limit=500, epsrel=1e-11)[0]
    assert abs(below / NE - 0.5) < 5e-4
    # in-band energy per recombination = fraction in band x <E>_n
    assert abs(m.energy_per_recomb(ich, T) / (N / Ntot * mean_n) - 1.0) < 1e-4


def test_milne_cdf_continuous_across_table_rows():
    """The CDF is interpolated in log T before inversion: the sampled energy
    (in units of kT above threshold) moves smoothly across a table node."""
    m = po.Milne(EMAX)
    lg = np.log10(1e2) + 4.0 * np.arange(192) / 191
    Tn = 10 ** lg[100]
    for ich in range(3):
        for u in (0.1, 0.5, 0.9):
            xs = [(m.sample(ich, T, u) - ETH[ich]) / (KB * T / EV)
                  for T in (Tn * (1 - 1e-9), Tn, Tn * (1 + 1e-9))]
            assert abs(xs[0] - xs[2]) < 1e-6


@pytest.mark.parametrize("ich,gfac", [(0, 1.0), (1, 0.25), (2, 1.0)])
def test_milne_alpha1_matches_recombination_rates(ich, gfac):
    """alpha_1(T) = (g_n/2g_+) (2/sqrt(pi)) (kT)^-3/2 sqrt(2/m_e)/(m_e c^2)
    Int E^2 sigma(E) e^-(E-E_th)/kT dE against the alpha_A - alpha_B that
    sets the packet count (src/milne_recomb_spectrum_mod.f90:257-276)."""
    L = po.lib()
    me, c = 9.1093837015e-28, 2.99792458e10
    fA = (L.orc_alphaA_HII, L.orc_alphaA_HeII, L.orc_alphaA_HeIII)[ich]
    fB = (L.orc_alphaB_HII, L.orc_alphaB_HeII, L.orc_alphaB_HeIII)[ich]
    for T in (5e3, 1e4, 2e4):
        kT = KB * T
        e0 = ETH[ich]
        integ = quad(lambda E: E * E * _sigma(ich, E) * np.exp(-(E - e0) * EV / kT), e0,
                     e0 + 60 * kT / EV, limit=500, epsrel=1e-11)[0] * EV ** 3
        a1 = gfac * (2 / np.sqrt(np.pi)) * kT ** -1.5 * np.sqrt(2 / me) / (me * c * c) * integ
        assert abs(a1 / (fA(T) - fB(T)) - 1.0) < 0.02, (T, a1, fA(T) - fB(T))


def _ionised_box(n=4, nH=100.0, T=1.2e4, diffuse_model=2, **kw):
    params = dict(nphotons=20000, qmc_seed=7, eion_min=13.598, eion_max=EMAX,
                  source_cdf_tol=1e-8, align_edges=1, nnu_ion=16, nsource=1, tstar=[4.5e4],
                  lum=[3e38], sx=[0.0], sy=[0.0], sz=[0.0], nx=n, ny=n, nz=n, xmax=2.0,
                  ymax=2.0, zmax=2.0, distance2cm=3.0856776e18, He_abund=0.1, te_fixed=T,
                  ion_relax=1.0, te_min=1e3, te_max=5e4, caseA=1, diffuse_field=1,
                  diffuse_model=diffuse_model)
    params.update(kw)
    m = po.Model(params)
    rng = np.random.default_rng(3)
    nl = m.nleaf
    m.state_init(np.full(nl, nH))
    xHI = 10 ** rng.uniform(-4, -2, nl)
    xHeI = 10 ** rng.uniform(-3, -1, nl)
    xHeII = rng.uniform(0.3, 0.9, nl)
    Te = T * rng.uniform(0.8, 1.2, nl)
    ne = nH * ((1 - xHI) + 0.1 * (xHeII + 2 * np.maximum(0, 1 - xHeI - xHeII)))
    m.set_state(xHI=xHI, xHeI=xHeI, xHeII=xHeII, Te=Te, ne=ne)
    return m


def test_diffuse_build_channel_luminosities_and_packet_count():
    """L_ch = n_e n_ion alpha_1(T) E_per_recomb(T) V per leaf; leaf CDF ends at
    1; packet count = nint(L / L_packet)."""
    m = _ionised_box()
    L, npk = m.diffuse_build()
    ch, cdf = m.diffuse_tables()
    st = m.state()
    Lib = po.lib()
    mil = po.Milne(EMAX)
    V = (1.0 * 3.0856776e18) ** 3                          # 1 pc cells
    x3 = np.maximum(0.0, 1 - st["xHeI"] - st["xHeII"])
    ions = (st["nH"] * (1 - st["xHI"]), 0.1 * st["nH"] * st["xHeII"], 0.1 * st["nH"] * x3)
    fa = (Lib.orc_alphaA_HII, Lib.orc_alphaA_HeII, Lib.orc_alphaA_HeIII)
    fb = (Lib.orc_alphaB_HII, Lib.orc_alphaB_HeII, Lib.orc_alphaB_HeIII)
    for k in range(3):
        want = np.array([st["ne"][i] * ions[k][i] * (fa[k](T) - fb[k](T))
                         * mil.energy_per_recomb(k, T) * EV * V for i, T in enumerate(st["Te"])])
        np.testing.assert_allclose(ch[:, k], want, rtol=1e-12)
    assert cdf[-1] == 1.0 and np.all(np.diff(cdf) >= 0)
    assert abs(L / ch.sum() - 1.0) < 1e-13
    Lpk = m.band()["Ltot"] / 20000
    assert npk == int(np.floor(L / Lpk + 0.5))


def test_diffuse_packets_represent_the_recombination_rate():
    """Equal-energy packets drawn from E n(E) represent exactly the
    ground-level recombination rate: sum_i L_packet / E_i = sum n_e n_ion
    alpha_1 V (to the Sobol noise); leaves are chosen in proportion to their
    luminosity and every packet starts inside its leaf."""
    m = _ionised_box(n=4, lum=[3e37])
    L, npk = m.diffuse_build()
    ch, cdf = m.diffuse_tables()
    assert npk > 20000
    Lpk = m.band()["Ltot"] / 20000
    out = np.array([m.launch_diffuse(i) for i in range(1, npk + 1)])
    E, leaf = out[:, 6], out[:, 8].astype(int)
    st = m.state()
    Lib = po.lib()
    x3 = np.maximum(0.0, 1 - st["xHeI"] - st["xHeII"])
    ions = (st["nH"] * (1 - st["xHI"]), 0.1 * st["nH"] * st["xHeII"], 0.1 * st["nH"] * x3)
    fa = (Lib.orc_alphaA_HII, Lib.orc_alphaA_HeII, Lib.orc_alphaA_HeIII)
    fb = (Lib.orc_alphaB_HII, Lib.orc_alphaB_HeII, Lib.orc_alphaB_HeIII)
    V = (1.0 * 3.0856776e18) ** 3
    R = sum(float(np.sum(st["ne"] * ions[k] * np.array([fa[k](T) - fb[k](T) for T in st["Te"]])))
            for k in range(3)) * V
    # in-band photons only (a tiny fraction of He II continuum lies above 100 eV)
    rate = np.sum(Lpk / (E * EV))
    assert abs(rate / R - 1.0) < 5e-3, (rate, R)
    counts = np.bincount(leaf - 1, minlength=m.nleaf)
    expect = ch.sum(axis=1) / L * npk
    assert np.max(np.abs(counts - expect)) <= 2.0 + 1e-3 * npk
    # positions inside the leaf (1 pc cells of a 4 pc box centred on 0)
    ix = np.floor((out[:, 0] + 2.0)).astype(int)
    iy = np.floor((out[:, 1] + 2.0)).astype(int)
    iz = np.floor((out[:, 2] + 2.0)).astype(int)
    assert np.array_equal(1 + ix + 4 * (iy + 4 * iz), leaf)
    np.testing.assert_allclose(np.linalg.norm(out[:, 3:6], axis=1), 1.0, atol=1e-14)


def test_pass_with_diffuse_closes_energy():
    """A pass with stellar + diffuse packets closes its energy ledger, and the
    emitted energy is (N_stellar + N_diffuse) L_packet."""
    m = _ionised_box(n=6, lum=[3e37], nphotons=5000)
    m.transport_pass(nthreads=4)
    t = m.tallies()
    h = m._head()
    emit, ab, esc = t["ledger"]
    assert h.diffuse_nphot > 100
    Lpk = m.band()["Ltot"] / 5000
    assert abs(emit / ((5000 + h.diffuse_nphot) * Lpk) - 1.0) < 1e-12
    assert abs(ab + esc - emit) <= 5e-13 * emit
