"""Pins of the oracle's n-level restatement (oracle/orc_nlevel.c): the
statistical-equilibrium line cooling and line emissivities of
src/nlevel_mod.f90, the (T, n_e) cooling tables of
src/nlevel_cooling_mod.f90 (cooling_model = 'local_ne') and the
recombination-line interpolation of src/sh95_mod.f90.

Against the reference modules compiled from source (oracle/_ref): every atom
of the data snapshot, bit for bit.  Against physics: a two-level atom's
closed-form populations, the low-density limit (every collisional excitation
out of the ground level radiated), the high-density limit (Boltzmann
populations), and the interpolation's exactness on power-law tables."""
import ctypes as C

import numpy as np
import pytest

from mochii_amd.atomic import load_nlevel
from oracle import pyoracle as po

ATOMS = load_nlevel()
REFDIR = b"/root/reference/data/atomic"
KB, HC = 1.380649e-16, 1.9864458571489287e-16


def _ref():
    L = po.ref_lib()
    if L is None:
        pytest.skip("reference modules not built (oracle/_ref)")
    return L


@pytest.mark.parametrize("key", sorted(ATOMS))
def test_cooling_and_emissivities_match_reference(key):
    R = _ref()
    f = REFDIR + b"/nlevel_" + key.encode() + b".txt"
    nl = R.ref_nlevel_load(f, len(f))
    if nl < 0:
        pytest.skip(f"{f} not present")
    a = po.NlAtom.make(ATOMS[key])
    assert nl == a.nrad
    e = np.zeros(600)
    for T in (500.0, 3e3, 8e3, 1e4, 2e4, 6e4):
        for ne in (1e-6, 1e-2, 10.0, 1e3, 1e5, 1e8):
            assert R.ref_nlevel_cooling_at(T, ne) == a.cooling_at(T, ne), (T, ne)
            R.ref_nlevel_emissivities(T, ne, e.ctypes.data_as(po._dp))
            assert np.array_equal(e[:nl], a.emis(T, ne)), (T, ne)
    assert np.array_equal([R.ref_nlevel_line_ident(k + 1) for k in range(nl)], a.line_ident())


# the ions cooling_model = 'local_ne' tabulates for the H II region workloads:
# H I and every stage of C N O Ne S Ar with a cooling fit and an n-level atom
TABLE_IONS = ["h_1", "c_1", "c_2", "c_3", "n_1", "n_2", "n_3", "o_1", "o_2", "o_3", "ne_2",
              "ne_3", "s_2", "s_3", "s_4", "ar_3", "ar_4"]


@pytest.mark.parametrize("key", TABLE_IONS)
def test_cooling_tables_match_reference(key):
    R = _ref()
    el, st = key.split("_")
    idx = R.ref_nlevel_cooling_add(REFDIR, len(REFDIR), el.encode(), len(el), int(st))
    if idx <= 0:
        pytest.skip(f"no table for {key}")
    tab = po.NlCool.build(po.NlAtom.make(ATOMS[key]))
    rng = np.random.default_rng(7)
    T = np.concatenate([10 ** rng.uniform(1.8, 6.2, 1500), 10 ** (2 + 0.01 * np.arange(401))])
    ne = np.concatenate([10 ** rng.uniform(-7, 11, 1500), 10 ** rng.uniform(-6, 10, 401)])
    lf = 10 ** rng.uniform(-25, -19, T.size)
    for t, n, l in zip(T, ne, lf):
        assert R.ref_nlevel_cooling_lam_se(idx, t, n) == tab.coeff(t, n), (t, n)
        assert R.ref_nlevel_cooling_apply(idx, l, t, n) == tab.apply(l, t, n), (t, n, l)


def _two_level(ups, A, g1=2.0, g2=4.0, e2=1000.0):
    """A two-level atom with a constant collision strength (type 2, one
    Chebyshev term)."""
    return po.NlAtom.make({"g": [g1, g2], "e_cm1": [0.0, e2], "rad": [[1, 2, A]],
                           "ups": [{"l": 1, "u": 2, "type": 2, "C": 1.0, "dE": 0.01,
                                    "st_lo": 0.0, "st_hi": 1.0, "logf": 0, "coef": [ups]}]})


def test_two_level_atom_closed_form():
    """n2/n1 = n_e q12 / (A21 + n_e q21), q21 = 8.629e-6 Ups / (g2 sqrt T),
    q12 = (g2/g1) q21 e^-dE/kT; cooling per ion and electron
    n2 A dE / n_e; emissivity n2 A dE."""
    ups, A, e2 = 2.5, 3e-3, 1000.0
    a = _two_level(ups, A, e2=e2)
    dE = e2 * HC
    for T in (1e3, 1e4, 3e4):
        for ne in (1e-4, 1.0, 1e2, 1e4, 1e7):
            q21 = 8.629e-6 * ups / (4.0 * np.sqrt(T))
            q12 = 2.0 * q21 * np.exp(-dE / (KB * T))
            r = ne * q12 / (A + ne * q21)
            n2 = r / (1.0 + r)
            assert abs(a.emis(T, ne)[0] / (n2 * A * dE) - 1) < 1e-13
            assert abs(a.cooling_at(T, ne) / (n2 * A * dE / ne) - 1) < 1e-13


@pytest.mark.parametrize("key", ["o_3", "n_2", "s_2", "c_2", "ne_3", "ar_3"])
def test_low_density_limit_radiates_every_excitation(key):
    """At n_e -> 0 every collisional excitation out of the ground level
    cascades back to it radiatively: Lambda_SE = sum_u q_1u (E_u - E_1)."""
    atom = ATOMS[key]
    a = po.NlAtom.make(atom)
    e = np.array(atom["e_cm1"])
    g = np.array(atom["g"])
    ground = int(np.argmin(e)) + 1
    for T in (5e3, 1e4, 2e4):
        want = 0.0
        for k, r in enumerate(atom["ups"]):
            lo, hi = (r["l"], r["u"]) if e[r["u"] - 1] >= e[r["l"] - 1] else (r["u"], r["l"])
            if lo != ground:
                continue
            dE = (e[hi - 1] - e[lo - 1]) * HC
            q = 8.629e-6 * a.ups(k, T) / (g[lo - 1] * np.sqrt(T)) * np.exp(-dE / (KB * T))
            want += q * dE
        got = a.cooling_at(T, 1e-8)
        assert abs(got / want - 1) < 1e-5, (T, got, want)


@pytest.mark.parametrize("key", ["o_3", "n_2", "s_3"])
def test_high_density_limit_is_boltzmann(key):
    """At n_e -> infinity the populations are Boltzmann: n_u / n_1 =
    g_u/g_1 e^-(E_u - E_1)/kT, so n_e Lambda_SE = sum_k n_u A_k dE_k."""
    atom = ATOMS[key]
    a = po.NlAtom.make(atom)
    e = np.array(atom["e_cm1"]) * HC
    g = np.array(atom["g"])
    for T in (5e3, 1e4, 2e4):
        pop = g * np.exp(-(e - e.min()) / (KB * T))
        pop /= pop.sum()
        want = 0.0
        for l, u, A in atom["rad"]:
            lo, hi = (l, u) if e[u - 1] >= e[l - 1] else (u, l)
            want += pop[hi - 1] * A * (e[hi - 1] - e[lo - 1])
        ne = 1e22
        assert abs(a.cooling_at(T, ne) * ne / want - 1) < 1e-6


def test_table_interpolation_exact_on_single_line_and_power_law():
    """The table's ln-linear-in-1/T rows and log-linear n_e columns follow a
    single line (exp(-Tx/T) times a smooth factor): a two-level atom is
    reproduced between the nodes to the interpolation tolerance the
    reference gates (0.5%), and the suppression S -> 1 at the n_e floor."""
    a = _two_level(2.0, 1.0, e2=2000.0)
    tab = po.NlCool.build(a)
    rng = np.random.default_rng(3)
    for T in 10 ** rng.uniform(2.5, 5.5, 200):
        for ne in (1e-6, 1e-3, 1e2):
            want = a.cooling_at(T, ne)
            assert abs(tab.coeff(T, ne) / want - 1) < 5e-3, (T, ne)
        assert tab.apply(tab.low_density(T), T, 1e-7) == pytest.approx(tab.low_density(T),
                                                                        rel=1e-14)
    # the fit covers more than the n-level model (f_cov < 1): the uncovered
    # part stays at its low-density value
    T, ne = 1e4, 1e8
    lam_ref = tab.low_density(T)
    S = tab.coeff(T, ne) / lam_ref
    assert tab.apply(2 * lam_ref, T, ne) == pytest.approx(2 * lam_ref * (0.5 * S + 0.5), rel=1e-14)


def test_recombination_line_interpolation_is_log_bilinear():
    """sh95_emis interpolates log e bilinearly in (log T, log n_e): on a
    table of e = c T^a n_e^b it is exact between the nodes; outside the grid
    it holds the edge value."""
    tab = po.lib()

    class Sh(C.Structure):          # orc_sh95
        _fields_ = [("nT", C.c_int), ("nD", C.c_int), ("nlin", C.c_int),
                    ("tgrid", C.c_double * 32), ("dgrid", C.c_double * 32),
                    ("wls", C.c_double * 12), ("etab", C.c_double * (12 * 32 * 32))]
    sh = Sh()
    tg = np.array([3e2, 1e3, 5e3, 1e4, 2e4, 3e4])
    dg = np.array([1e2, 1e4, 1e6, 1e8])
    sh.nT, sh.nD, sh.nlin = tg.size, dg.size, 2
    sh.tgrid[:tg.size] = tg
    sh.dgrid[:dg.size] = dg
    for k, (c, a, b) in enumerate([(1e-25, -0.9, 0.05), (3e-26, -1.1, -0.02)]):
        for i, T in enumerate(tg):
            for j, n in enumerate(dg):
                sh.etab[(k * tg.size + i) * dg.size + j] = c * T ** a * n ** b
    rng = np.random.default_rng(5)
    for T, n in zip(10 ** rng.uniform(2.5, 4.47, 300), 10 ** rng.uniform(2, 8, 300)):
        for k, (c, a, b) in enumerate([(1e-25, -0.9, 0.05), (3e-26, -1.1, -0.02)]):
            got = tab.orc_sh95_emis(C.byref(sh), k, T, n)
            assert abs(got / (c * T ** a * n ** b) - 1) < 1e-12
    # clamped at the grid edges (T above 3e4 K, n_e beyond the table)
    e_edge = tab.orc_sh95_emis(C.byref(sh), 0, 3e4, 1e8)
    assert tab.orc_sh95_emis(C.byref(sh), 0, 5e4, 1e9) == e_edge
