"""Oracle pins for the He I excited-level channel of the diffuse field
(par%hei_diffuse: src/diffuse_mod.f90:95-105, 254-274, 308-324,
src/hei_cascade_mod.f90, src/hei_twophoton_mod.f90).

Independent checks: the reference's hei_cascade_mod compiled from source
(branch fractions and the 584 A closure, bit for bit); the two-photon moments
against the exact integrals of the piecewise-linear shape (closed form per
table interval) and the reference's stated 0.5564 photons per decay; the
sampler's energy-weighted mean, median and <1/E> = N/E against those exact
integrals; and the photon-rate identity that the channel's equal-energy
packets represent n_e n_HeII alpha_B V (f_3S + f_1P p584 + (f_1S + f_1P (1 -
p584)) N_2ph) photons per second."""
import numpy as np
import pytest

from mochii_amd.atomic import load_hei_two_photon
from oracle import pyoracle as po

E_TOT, E3S, E1P, ETH_HI = 20.62, 19.82, 21.22, 13.598
EV = 1.602176634e-12


def _ref():
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    return ref


def test_branch_fractions_and_584_closure_equal_reference_module():
    ref = _ref()
    f = np.zeros(3)
    dp = po._dp
    for T in np.concatenate([np.geomspace(100.0, 1e6, 80), [2999.0, 3e3, 1e4, 2e4, 2.0001e4]]):
        ref.ref_hei_branch_fractions(T, f[0:].ctypes.data_as(dp), f[1:].ctypes.data_as(dp),
                                     f[2:].ctypes.data_as(dp))
        got = po.hei_branch_fractions(T)
        assert got == tuple(f), T
        assert abs(sum(got) - 1.0) < 1e-15
    rng = np.random.default_rng(11)
    for _ in range(400):
        T = 10 ** rng.uniform(2, 6)
        xHI, xHeI = 10 ** rng.uniform(-8, 0), 10 ** rng.uniform(-8, 0)
        assert po.hei_584_ionizes_H(T, xHI, xHeI) == ref.ref_hei_584_ionizes_H(T, xHI, xHeI)
    for args in ((1e4, 0.0, 0.1), (1e4, 0.1, 0.0), (0.0, 0.5, 0.5), (1e4, 1.0, -1.0)):
        assert po.hei_584_ionizes_H(*args) == ref.ref_hei_584_ionizes_H(*args), args


def test_branch_fractions_closed_form():
    """The AGN3 power laws at 10^4 K are the coefficients themselves; the
    fit is clamped to [3e3, 2e4] K."""
    f = po.hei_branch_fractions(1e4)
    s = 2.10e-13 + 2.06e-14 + 4.17e-14
    np.testing.assert_allclose(f, (2.10e-13 / s, 2.06e-14 / s, 4.17e-14 / s), rtol=1e-15)
    assert po.hei_branch_fractions(1e3) == po.hei_branch_fractions(3e3)
    assert po.hei_branch_fractions(1e5) == po.hei_branch_fractions(2e4)
    assert po.hei_branch_fractions(8e3) != po.hei_branch_fractions(1.2e4)


def _exact(y, A, lo, hi, k):
    """Int_lo^hi y^k A(y) dy for A linear between the rows, exactly (Simpson
    is exact for the cubic y^k A at k <= 2 on each clipped interval)."""
    tot = 0.0
    for i in range(len(y) - 1):
        a, b = max(y[i], lo), min(y[i + 1], hi)
        if b <= a:
            continue

        def f(v, i=i):
            return v ** k * (A[i] + (A[i + 1] - A[i]) * (v - y[i]) / (y[i + 1] - y[i]))
        tot += (b - a) / 6.0 * (f(a) + 4.0 * f(0.5 * (a + b)) + f(b))
    return tot


def test_two_photon_moments_against_exact_integrals():
    y, A = load_hei_two_photon()
    assert y.size == 41 and np.all(np.diff(y) > 0)
    h = po.hei_two_photon(y, A, eion_max=100.0)
    norm = _exact(y, A, y[0], y[-1], 0)
    lo, hi = ETH_HI / E_TOT, 1.0
    n_per = 2.0 * _exact(y, A, lo, hi, 0) / norm
    e_per = 2.0 * _exact(y, A, lo, hi, 1) / norm * E_TOT
    assert abs(h.n_per_decay / n_per - 1.0) < 1e-6, (h.n_per_decay, n_per)
    assert abs(h.e_per_decay / e_per - 1.0) < 1e-6, (h.e_per_decay, e_per)
    # the reference's stated in-band figure (src/hei_twophoton_mod.f90 header)
    assert abs(h.n_per_decay - 0.5564) < 1e-4
    # the pair spectrum is symmetric about y = 1/2: two photons per decay
    assert abs(2.0 * _exact(y, A, y[0], 0.5, 0) / norm - 1.0) < 1e-3
    # a band edge below 20.62 eV truncates the window; one below 13.598 closes it
    h2 = po.hei_two_photon(y, A, eion_max=18.0)
    assert abs(h2.e_per_decay / (2.0 * _exact(y, A, lo, 18.0 / E_TOT, 1) / norm * E_TOT) - 1) < 1e-6
    h3 = po.hei_two_photon(y, A, eion_max=13.0)
    assert h3.n_per_decay == 0.0 and h3.e_per_decay == 0.0


def test_two_photon_sampler_against_exact_integrals():
    y, A = load_hei_two_photon()
    h = po.hei_two_photon(y, A, eion_max=100.0)
    lo = ETH_HI / E_TOT
    ns = 100000
    E = np.array([po.hei2ph_sample_energy(h, (i + 0.5) / ns) for i in range(ns)])
    assert E.min() >= ETH_HI * (1 - 1e-12) and E.max() <= E_TOT * (1 + 1e-12)
    I1, I2 = _exact(y, A, lo, 1.0, 1), _exact(y, A, lo, 1.0, 2)
    # energy-weighted draws: <E> = E_TOT Int y^2 A / Int y A
    assert abs(E.mean() / (E_TOT * I2 / I1) - 1.0) < 1e-5
    # <1/E> over energy-weighted draws = photons / energy per decay
    assert abs(np.mean(1.0 / E) / (h.n_per_decay / h.e_per_decay) - 1.0) < 1e-5
    med = np.median(E) / E_TOT
    assert abs(_exact(y, A, lo, med, 1) / I1 - 0.5) < 1e-4
    assert po.hei2ph_sample_energy(h, 0.0) == pytest.approx(ETH_HI, rel=1e-12)
    assert po.hei2ph_sample_energy(h, 1.0) == pytest.approx(E_TOT, rel=1e-12)


def test_branch_energies_limits():
    y, A = load_hei_two_photon()
    h = po.hei_two_photon(y, A)
    T = 9000.0
    f3s, f1s, f1p = po.hei_branch_fractions(T)
    # no neutral He: every 584 A photon ionizes H
    e3s, e1p, e1s = po.hei_branch_energies(h, T, 1e-3, 0.0)
    assert (e3s, e1p, e1s) == (f3s * E3S, f1p * E1P, f1s * h.e_per_decay)
    # no neutral H: every 584 A photon converts to 2^1S
    e3s, e1p, e1s = po.hei_branch_energies(h, T, 0.0, 0.5)
    assert e1p < 1e-200 and abs(e1s / ((f1s + f1p) * h.e_per_decay) - 1) < 1e-14
    # the closure probability at a textbook point: 77 T^-1/2 x_HeI / x_HI = 1
    xHI = 77.0 / np.sqrt(T) * 0.3
    assert abs(po.hei_584_ionizes_H(T, xHI, 0.3) - 0.5) < 1e-15


def _hei_box(n=4, nH=100.0, T=1.1e4, hei=1, **kw):
    """Neutral-H (x_HI = 1) and He I/He II only: channels 1 and 3 vanish, so
    every packet below 22 eV comes from the He I excited channel and every
    other from the He I ground continuum (>= 24.587 eV)."""
    params = dict(nphotons=20000, qmc_seed=7, eion_min=13.598, eion_max=100.0,
                  source_cdf_tol=1e-8, align_edges=1, nnu_ion=16, nsource=1, tstar=[4.5e4],
                  lum=[3e37], sx=[0.0], sy=[0.0], sz=[0.0], nx=n, ny=n, nz=n, xmax=2.0,
                  ymax=2.0, zmax=2.0, distance2cm=3.0856776e18, He_abund=0.1, te_fixed=T,
                  ion_relax=1.0, te_min=1e3, te_max=5e4, caseA=1, diffuse_field=1,
                  diffuse_model=2, hei_diffuse=hei)
    params.update(kw)
    m = po.Model(params)
    rng = np.random.default_rng(9)
    nl = m.nleaf
    m.state_init(np.full(nl, nH))
    if hei:
        m.diffuse_hei_setup(*load_hei_two_photon())
    xHeII = rng.uniform(0.3, 0.9, nl)
    xHeI = 1.0 - xHeII
    Te = T * rng.uniform(0.8, 1.2, nl)
    ne = nH * 0.1 * xHeII
    m.set_state(xHI=np.ones(nl), xHeI=xHeI, xHeII=xHeII, Te=Te, ne=ne)
    return m


def test_diffuse_build_with_hei_channel():
    on, off = _hei_box(hei=1), _hei_box(hei=0)
    L1, n1 = on.diffuse_build()
    L0, n0 = off.diffuse_build()
    ch1, _ = on.diffuse_tables()
    ch0, _ = off.diffuse_tables()
    assert np.array_equal(ch1[:, :3], ch0[:, :3]) and np.all(ch0[:, 3] == 0.0)
    # x_HeIII = (1 - x_HeI) - x_HeII is a rounding residue here
    assert np.all(ch1[:, 0] == 0.0) and np.all(ch1[:, 2] < 1e-12 * ch1[:, 1])
    assert np.all(ch1[:, 3] > 0.0) and n1 > n0
    assert abs(L1 / ch1.sum() - 1.0) < 1e-13 and abs((L1 - L0) / ch1[:, 3].sum() - 1) < 1e-12


def test_hei_packets_represent_the_excited_recombination_rate():
    """Sum over channel-4 packets of L_packet / E = sum n_e n_HeII alpha_B V
    (f_3S + f_1P p584 + (f_1S + f_1P (1 - p584)) N_2ph) to the Sobol noise;
    each line sits at its energy with the weight of its branch."""
    m = _hei_box(lum=[1e36])
    L, npk = m.diffuse_build()
    assert npk > 20000
    Lpk = m.band()["Ltot"] / 20000
    out = np.array([m.launch_diffuse(i) for i in range(1, npk + 1)])
    E = out[:, 6]
    hei = E < 22.0
    assert np.all(E[~hei] >= 24.587 - 1e-9)
    st = m.state()
    y, A = load_hei_two_photon()
    h = po.hei_two_photon(y, A)
    aB = po.lib().orc_alphaB_HeII
    V = 3.0856776e18 ** 3
    rate = lines = 0.0
    for i in range(m.nleaf):
        T = st["Te"][i]
        f3s, f1s, f1p = po.hei_branch_fractions(T)
        p = po.hei_584_ionizes_H(T, st["xHI"][i], st["xHeI"][i])
        r = st["ne"][i] * 0.1 * st["nH"][i] * st["xHeII"][i] * aB(T) * V
        rate += r * (f3s + f1p * p + (f1s + f1p * (1 - p)) * h.n_per_decay)
        lines += r * (f3s + f1p * p)
    got = np.sum(Lpk / (E[hei] * EV))
    assert abs(got / rate - 1.0) < 5e-3, (got, rate)
    is_line = np.isin(E[hei], (E3S, E1P))
    got_lines = np.sum(Lpk / (E[hei][is_line] * EV))
    assert abs(got_lines / lines - 1.0) < 1e-2, (got_lines, lines)
    pair = E[hei][~is_line]
    assert pair.min() >= ETH_HI * (1 - 1e-12) and pair.max() <= E_TOT * (1 + 1e-12)
