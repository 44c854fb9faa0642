"""Pins of the oracle's trace-metal restatement (oracle/orc_metals.c).

Against the reference compiled from source (oracle/_ref, when built here):
sigma_vfky96 / sigma_vy95 for every registry transition, the Voronov form
through the reference's ci_HI / ci_HeI / ci_HeII, the Badnell RR + DR forms
through alphaA_HII / alphaA_HeII / alphaA_HeIII (the He II rate exercises the
C exp(-T2/T) term and a three-term DR sum), ci_dere_ratio.  Against physics
the formulas do not restate: charge-exchange detailed balance of the first
transitions (the reference's tests/charge_exchange/check_detailed_balance.f90
gate), the stage fractions against a dense linear solve of the stage
balance, the metal sums against the fractions, the charge-exchange solve
against the hydrogen residual it zeroes and its own fixed point
(tests/charge_exchange/check_cell_convergence.f90), the two-level H-impact
cooling's low/high density limits, metal_cooling one element and one
cooling stage at a time against the data files' fits times that stage's
fraction, metal_freefree against the Gaunt-factor vectors of the reference's
Python mirror, and the aligned band edges with metal thresholds against the
merge / drop rules.
"""
import ctypes as C
import os

import numpy as np
import pytest

from mochii_amd.atomic import Element, load_snapshot, metal_thresholds, registry, to_struct
from oracle import pyoracle as po
from tests.cases import oracle_params
from mochii_amd.workloads import hii_thermal, HI_COOLING_FIT

EV2ERG, KB = 1.602176634e-12, 1.380649e-16
DATA = load_snapshot()
ALL = registry({k: 1e-5 for k in DATA}, data=DATA)      # every registry element


def _ref():
    L = po.ref_lib()
    if L is None:
        pytest.skip("reference modules not built (oracle/_ref)")
    return L


def _ptr(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _el(d):
    return to_struct([d])


# ---- cross sections ----------------------------------------------------------
def test_cross_sections_match_reference():
    R, L = _ref(), po.lib()
    n = 0
    for el in ALL:
        for it, tr in enumerate(el["transitions"]):
            p = np.array(tr["photo"], dtype=np.float64)
            for E in np.concatenate([[tr["eth"], tr["eth"] * (1 + 1e-9)],
                                     np.geomspace(tr["eth"], 200.0, 40)]):
                if tr["photo_form"] == 2:
                    ref = R.ref_vy95(E, _ptr(p[:7]))
                else:
                    ref = R.ref_vfky96(E, _ptr(p))
                got = L.orc_species_sigma(C.byref(_el(el)[0]), it, E)
                assert got == pytest.approx(ref, rel=1e-13, abs=1e-40), (el["name"], it, E)
                n += 1
            # zero below threshold
            assert L.orc_species_sigma(C.byref(_el(el)[0]), it, tr["eth"] * 0.999) == 0.0
    assert n > 1000


def test_cross_section_threshold_magnitudes():
    """Threshold cross sections are the physical few-Mb values (a units slip
    in the Mb -> cm^2 factor or a dropped term fails this)."""
    L = po.lib()
    for el in registry({"ar": 2.5e-6}, data=DATA):
        for it, tr in enumerate(el["transitions"]):
            s = L.orc_species_sigma(C.byref(_el(el)[0]), it, tr["eth"] * 1.0001)
            assert 5e-20 < s < 1e-16, (el["name"], it, s)


# ---- rate coefficients -------------------------------------------------------
HHE_VORONOV = {6: (13.6, 0.0, 2.91e-8, 0.232, 0.39), 7: (24.6, 0.0, 1.75e-8, 0.180, 0.35),
               8: (54.4, 1.0, 2.05e-9, 0.265, 0.25)}      # ref_rate ids of ci_HI/HeI/HeII


@pytest.mark.parametrize("which", [6, 7, 8])
def test_voronov_form_matches_reference(which):
    R, L = _ref(), po.lib()
    c = np.array(HHE_VORONOV[which])
    for T in np.geomspace(1e3, 1e5, 23):
        assert L.orc_voronov_ci(_ptr(c), T) == pytest.approx(R.ref_rate(which, T), rel=1e-13)


def _rec_element(rr, dr_c=(), dr_e=()):
    d = dict(DATA["c"])
    d = dict(d, abund=1e-4, nstage=2)
    tr = dict(d["transitions"][0], rr=list(rr), ndr=len(dr_c), dr_c=list(dr_c), dr_e=list(dr_e),
              dr2=None, rr2=None)
    d["transitions"] = [tr]
    return d


@pytest.mark.parametrize("which,rr,dr", [
    (0, (8.318e-11, 0.7472, 2.965, 7.001e5, 0.0, 0.0), ((), ())),
    (2, (5.235e-11, 0.6988, 7.301, 4.475e6, 0.0829, 1.682e5),
     ((1.417e-3, 2.235e-4, -2.185e-5), (4.633e5, 5.532e5, 8.887e5))),
    (4, (1.818e-10, 0.7492, 1.017e1, 2.786e6, 0.0, 0.0), ((), ())),
])
def test_alpha_rec_form_matches_reference(which, rr, dr):
    """alpha_rec's Badnell RR + DR sum fed the reference's own H II / He II /
    He III coefficients equals alphaA_* of the compiled recomb_mod."""
    R, L = _ref(), po.lib()
    e = _el(_rec_element(rr, *dr))
    for T in np.geomspace(1e3, 1e5, 23):
        assert L.orc_alpha_rec(C.byref(e[0]), 0, T) == pytest.approx(R.ref_rate(which, T),
                                                                      rel=2e-14)


def test_alpha_rec_power_law_and_dr2_branches():
    L = po.lib()
    fe = DATA["fe"]                       # RR2 power laws
    e = _el(dict(fe, abund=1e-6))
    for it, tr in enumerate(fe["transitions"]):
        A, eta = tr["rr2"]
        # DR terms vanish at low T: the power law alone, and its slope is -eta
        a1, a2 = (L.orc_alpha_rec(C.byref(e[0]), it, T) for T in (1.0e3, 2.0e3))
        dr = sum(c * (T ** -1.5) * np.exp(-ee / T) for c, ee in zip(tr["dr_c"], tr["dr_e"])
                 for T in (1.0e3,))
        if dr < 1e-6 * a1:
            assert a1 == pytest.approx(A * 0.1 ** (-eta), rel=1e-12)
            assert np.log(a2 / a1) / np.log(2.0) == pytest.approx(-eta, rel=1e-6)
    ar = DATA["ar"]                       # DR2 (Shull & Van Steenberg) on Ar I, Ar II
    e = _el(dict(ar, abund=1e-6))
    for it, tr in enumerate(ar["transitions"]):
        if tr["dr2"] is None:
            continue
        A, B, T0, T1 = tr["dr2"]
        T = 1e9                           # exponentials -> 1: DR2 -> A T^-1.5 (1 + B)
        rr = L.orc_alpha_rec(C.byref(e[0]), it, T)
        assert rr >= A * T ** -1.5 * (1 + B) * 0.999


def test_ci_dere_ratio_matches_reference():
    R, L = _ref(), po.lib()
    for Z in (1, 2, 6, 7, 8, 10, 12, 14, 16, 17, 18, 20, 26, 3):
        for st in range(0, 7):
            assert L.orc_ci_dere_ratio(1, Z, st) == pytest.approx(R.ref_ci_dere_ratio(Z, st),
                                                                   rel=1e-7)
            assert L.orc_ci_dere_ratio(0, Z, st) == 1.0


# (tests/charge_exchange/check_detailed_balance.f90) ground-term weights
GW = {"c": (9, 6), "n": (4, 9), "o": (9, 4), "ne": (1, 6), "mg": (1, 2), "si": (9, 6),
      "s": (9, 4), "ar": (1, 6), "fe": (25, 30), "cl": (6, 9), "ca": (1, 2)}


def test_dere_hybrid_h_he_rates_match_reference():
    """ci_model 'dere_hybrid' scales ci_HI / ci_HeI / ci_HeII by the Dere
    ratio (src/recomb_mod.f90:263-279): the oracle under the switch against
    the reference module with par%ci_model = 'dere_hybrid', and the switch
    restores the Voronov rates."""
    R, L = _ref(), po.lib()
    T = np.geomspace(3e3, 1e6, 25)
    plain = [[L.orc_ci_HI(t), L.orc_ci_HeI(t), L.orc_ci_HeII(t)] for t in T]
    prev = L.orc_ci_model_set(1)
    try:
        for t in T:
            for which, f in ((6, L.orc_ci_HI), (7, L.orc_ci_HeI), (8, L.orc_ci_HeII)):
                assert f(t) == pytest.approx(R.ref_rate_dere(which, t), rel=1e-13, abs=1e-300)
    finally:
        L.orc_ci_model_set(prev)
    again = [[L.orc_ci_HI(t), L.orc_ci_HeI(t), L.orc_ci_HeII(t)] for t in T]
    assert again == plain
    np.testing.assert_allclose(R.ref_rate_dere(6, 2e4) / R.ref_rate(6, 2e4), 0.9063, rtol=1e-14)


def test_charge_exchange_detailed_balance():
    L = po.lib()
    tested = 0
    for el in ALL:
        tr = el["transitions"][0]
        if tr["cxi_form"] <= 0 or tr["cxr_form"] <= 0:
            continue
        ci, cr = np.array(tr["cxi"]), np.array(tr["cxr"])
        k = lambda T: (L.orc_cx_rate(tr["cxi_form"], _ptr(ci), T),
                       L.orc_cx_rate(tr["cxr_form"], _ptr(cr), T))
        ki, kr = k(1.0e4)
        if ki <= 0 or kr <= 0 or min(ki, kr) < 1e-15:
            continue
        g0, g1 = GW[el["name"]]
        dE = (tr["eth"] - 13.5984) * EV2ERG / KB
        for T in (5e3, 8e3, 1e4, 2e4):
            ki, kr = k(T)
            req = (g1 * 2.0) / (g0 * 1.0) * np.exp(-dE / T)
            assert abs(ki / kr / req - 1.0) <= 0.35, (el["name"], T, ki / kr / req)
            tested += 1
    assert tested >= 4            # O, N and the other near-resonant pairs


# ---- model-level pieces on a tiny grid --------------------------------------
def _metal_model(n=4, cx_two_way=1, metal_ne=1, seed=3):
    els = registry({"ar": 2.5e-6}, data=DATA)
    params, nH = hii_thermal(n=n, nphotons=1000)
    op = oracle_params(params)
    m = po.Model({k: v for k, v in op.items() if k not in ("xHI_init", "xHeI_init", "xHeII_init")}
                 | dict(cx_two_way=cx_two_way, metal_ne=metal_ne), elements=els)
    m.params.update(op)
    m.state_init(nH)
    A = np.array([a for a, _ in HI_COOLING_FIT])
    Ti = np.array([t for _, t in HI_COOLING_FIT])
    m.cooling_setup(A, Ti)
    rng = np.random.default_rng(seed)
    for ie, el in enumerate(els):
        nt = el["nstage"] - 1
        g = 10.0 ** rng.uniform(-12, -7, size=(m.nleaf, nt))
        m.set_species_rates(ie, g, 10.0 ** rng.uniform(-24, -19, size=(m.nleaf, nt)))
    return m, els


def _stage_rates(el, T, ne, nHI, nHII, gam):
    L = po.lib()
    e = _el(el)
    rion, rrec = [], []
    for it, tr in enumerate(el["transitions"]):
        ci = L.orc_voronov_ci(_ptr(tr["ci"]), T)
        a = L.orc_alpha_rec(C.byref(e[0]), it, T)
        cxi = L.orc_cx_rate(tr["cxi_form"], _ptr(tr["cxi"]), T) if tr["cxi_form"] > 0 else 0.0
        cxr = L.orc_cx_rate(tr["cxr_form"], _ptr(tr["cxr"]), T) if tr["cxr_form"] > 0 else 0.0
        rion.append(gam[it] + ne * ci + nHII * cxi)
        rrec.append(ne * a + nHI * cxr)
    return np.array(rion), np.array(rrec)


def test_stage_fractions_solve_the_balance_equations():
    m, els = _metal_model()
    rng = np.random.default_rng(7)
    for _ in range(20):
        il = int(rng.integers(m.nleaf))
        T, ne, x = 10 ** rng.uniform(3, 4.6), 10 ** rng.uniform(-1, 2.5), rng.uniform(0, 1)
        nH = m.arr("nH")[il]
        for ie, el in enumerate(els):
            f = m.species_fractions_at(ie, il, T, ne, x)
            gam, _ = m.species_rates(ie)
            rion, rrec = _stage_rates(el, T, ne, nH * x, nH * (1 - x), gam[il])
            ns = el["nstage"]
            # dn/dt = 0 for every stage of the chain, plus sum f = 1
            A = np.zeros((ns, ns))
            for i in range(ns - 1):
                A[i, i] -= rion[i]; A[i + 1, i] += rion[i]
                A[i + 1, i + 1] -= rrec[i]; A[i, i + 1] += rrec[i]
            A[-1, :] = 1.0
            b = np.zeros(ns); b[-1] = 1.0
            ref = np.linalg.solve(A, b)
            assert np.allclose(f, ref, rtol=1e-9, atol=1e-14), (el["name"], f, ref)


def test_metal_sums_follow_the_fractions():
    m, els = _metal_model()
    rng = np.random.default_rng(11)
    L = po.lib()
    for _ in range(10):
        il = int(rng.integers(m.nleaf))
        T, ne, x = 10 ** rng.uniform(3.3, 4.5), 10 ** rng.uniform(0, 2), rng.uniform(0.01, 0.99)
        nH = m.arr("nH")[il]
        nem, rin, rout = m.metal_sums(il, T, ne, x)
        e_nem = e_in = e_out = 0.0
        for ie, el in enumerate(els):
            f = m.species_fractions_at(ie, il, T, ne, x)
            e_nem += el["abund"] * nH * sum(f[i] * i for i in range(el["nstage"]))
            for it, tr in enumerate(el["transitions"]):
                if tr["cxi_form"] > 0:
                    e_in += L.orc_cx_rate(tr["cxi_form"], _ptr(tr["cxi"]), T) * el["abund"] * nH * f[it]
                if tr["cxr_form"] > 0:
                    e_out += (L.orc_cx_rate(tr["cxr_form"], _ptr(tr["cxr"]), T) * el["abund"] * nH *
                              f[it + 1])
        assert nem == pytest.approx(e_nem, rel=1e-12)
        assert rin == pytest.approx(e_in, rel=1e-12, abs=1e-300)
        assert rout == pytest.approx(e_out, rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("two_way", [1, 0])
def test_cell_solve_reaches_its_fixed_point(two_way):
    """One solve_ion_cell call from the state of another lands on the same
    state (check_cell_convergence's gate, 1e-9), and with two-way charge
    exchange the hydrogen balance including both exchange terms holds."""
    m, els = _metal_model(cx_two_way=two_way)
    rng = np.random.default_rng(5)
    for _ in range(8):
        il = int(rng.integers(m.nleaf))
        T = 10 ** rng.uniform(3.6, 4.3)
        gam = (10 ** rng.uniform(-11, -7), 10 ** rng.uniform(-12, -8), 10 ** rng.uniform(-14, -10))
        a = m.solve_cell(il, T, x=(1.0, 1.0, 0.0), gam=gam)
        b = m.solve_cell(il, T, x=a[:3], gam=gam)
        for u, v in zip(a, b):
            assert abs(u - v) <= 1e-9 * max(abs(u), 1e-12) + 1e-15
        xHI, ne = a[0], a[3]
        nH = m.arr("nH")[il]
        if two_way:
            nem, rin, rout = m.metal_sums(il, T, ne, xHI)
            aH, cH = po.lib().orc_alphaA_HII(T), po.lib().orc_ci_HI(T)
            loss = nH * xHI * (gam[0] + cH * ne + rout)
            gain = (aH * ne + rin) * nH * (1 - xHI)
            assert abs(loss - gain) <= 1e-6 * loss + 1e-300


def _element_model(els, seed=13):
    """A 4^3 oracle model with the given registry and random metal
    photoionisation rates; cooling_model 'low_density' (no n-level tables)."""
    params, nH = hii_thermal(n=4, nphotons=1000)
    op = oracle_params(params)
    m = po.Model({k: v for k, v in op.items() if k not in ("xHI_init", "xHeI_init", "xHeII_init")},
                 elements=els)
    m.params.update(op)
    m.state_init(nH)
    m.cooling_setup(np.zeros(0), np.zeros(0))
    rng = np.random.default_rng(seed)
    for ie, el in enumerate(els):
        nt = el["nstage"] - 1
        g = 10.0 ** rng.uniform(-11, -8, size=(m.nleaf, nt))
        m.set_species_rates(ie, g, np.zeros((m.nleaf, nt)))
    return m


def _lam_fit(fit, T):
    """cooling_eval (src/cooling_mod.f90:125-134): T^-1/2 sum A_i exp(-T_i/T)."""
    return sum(a * np.exp(-t / T) for a, t in fit) / np.sqrt(T)


@pytest.mark.parametrize("name", sorted(DATA))
def test_metal_cooling_one_stage_in_isolation(name):
    """metal_cooling (src/species_mod.f90:825-846) with one element of the
    registry and its cooling fit on one stage only: n_e n_el f_s
    Lambda_fit(T) with f_s that stage's fraction (pinned by the dense solve
    above) and Lambda_fit the data file's fit; every draw also checks that
    the neighbouring stage's fraction would give another value, so a stage
    index, abundance or n_e factor slip fails."""
    L = po.lib()
    rng = np.random.default_rng(len(name))
    for st, fit in sorted(DATA[name]["cooling"].items()):
        el = dict(DATA[name], abund=3.0e-5, cooling={st: fit})
        m = _element_model([el])
        nH = m.arr("nH")
        discriminating = 0
        for _ in range(6):
            il = int(rng.integers(m.nleaf))
            T, ne, x = 10 ** rng.uniform(3.4, 4.4), 10 ** rng.uniform(0, 2), rng.uniform(0.05, 0.95)
            f = m.species_fractions_at(0, il, T, ne, x)
            want = ne * (3.0e-5 * nH[il]) * f[st - 1] * _lam_fit(fit, T)
            got = L.orc_metal_cooling(m.c, il, T, nH[il], ne, nH[il] * x, nH[il] * (1 - x))
            assert got == pytest.approx(want, rel=1e-13, abs=1e-300), (name, st, T, ne)
            other = [f[j] for j in (st - 2, st) if 0 <= j < len(f)]
            discriminating += all(abs(o - f[st - 1]) > 1e-3 * f[st - 1] for o in other)
        assert discriminating >= 3, (name, st)


def test_metal_freefree_against_gaunt_vectors():
    """metal_freefree (src/species_mod.f90:852-868) for every registry
    element at the temperatures of the reference's Gaunt-factor vectors
    (tests/golden/rates/cooling_rates.txt, columns Z = 1..4): 1.42554e-27
    T^1/2 n_e n_H sum over elements of abund sum_{j>=2} f_j (j-1)^2
    gbar_ff(T, j-1), with the stage fractions of the oracle."""
    tab = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "rates",
                                  "cooling_rates.txt"))
    rows = tab[(tab[:, 0] >= 2e3) & (tab[:, 0] <= 5e4)]
    L = po.lib()
    els = registry({k: 0.0 for k in DATA} | {n: 1e-5 * (1 + i) for i, n in enumerate(sorted(DATA))},
                   data=DATA)
    m = _element_model(els)
    nH = m.arr("nH")
    rng = np.random.default_rng(17)
    for row in rows:
        T = row[0]
        il = int(rng.integers(m.nleaf))
        ne, x = 10 ** rng.uniform(0, 2), rng.uniform(0.05, 0.95)
        s = 0.0
        for ie, el in enumerate(els):
            f = m.species_fractions_at(ie, il, T, ne, x)
            for j in range(2, el["nstage"] + 1):
                s += el["abund"] * f[j - 1] * (j - 1) ** 2 * row[6 + (j - 1)]
        want = 1.42554e-27 * np.sqrt(T) * ne * nH[il] * s
        got = L.orc_metal_freefree(m.c, il, T, nH[il], ne, nH[il] * x, nH[il] * (1 - x))
        # the Gaunt nodes agree with the vectors to a few 1e-8 (single-precision fit)
        assert got == pytest.approx(want, rel=2e-7), T


def test_h_impact_cooling_two_level_limits():
    """metal_cooling_H's [C II] / [O I] terms: linear in n_HI far below the
    H critical density, saturating at n_ion A q_lu/(q_ul+q_lu) dE above it."""
    m, els = _metal_model()
    il, T, ne, x = 0, 120.0, 1e-3, 0.999
    nH = m.arr("nH")[il]
    L = po.lib()
    ic = [e["name"] for e in els].index("c")
    io = [e["name"] for e in els].index("o")
    fC = m.species_fractions_at(ic, il, T, ne, x)
    fO = m.species_fractions_at(io, il, T, ne, x)

    def cool(nHI):
        return L.orc_metal_cooling_H(m.c, il, T, nH, ne, nHI, nH * (1 - x))

    qc = 7.6e-10 * (T / 100) ** 0.14
    qo = 9.2e-11 * (T / 100) ** 0.67
    lc = 2 * qc * np.exp(-91.25 / T)
    lo = 0.6 * qo * np.exp(-227.7 / T)
    # the stage fractions depend on n_HI only through charge exchange, which
    # is weak here; compare per unit n_ion with the fractions at each n_HI
    low = 1e-6
    exp_low = (els[ic]["abund"] * nH * fC[1] * low * lc * 91.25 * KB +
               els[io]["abund"] * nH * fO[0] * low * lo * 227.7 * KB)
    assert cool(low) == pytest.approx(exp_low, rel=1e-3)
    high = 1e12
    fC = m.species_fractions_at(ic, il, T, ne, high / nH) if high / nH <= 1 else fC
    exp_high = (els[ic]["abund"] * nH * fC[1] * 2.321e-6 * lc / (qc + lc) * 91.25 * KB +
                els[io]["abund"] * nH * fO[0] * 8.865e-5 * lo / (qo + lo) * 227.7 * KB)
    assert cool(high) == pytest.approx(exp_high, rel=1e-3)


def test_aligned_edges_with_metal_thresholds():
    els = registry({"ar": 2.5e-6}, data=DATA)
    meth = metal_thresholds(els)
    base = dict(nphotons=10, qmc_seed=1, eion_min=13.598, eion_max=100.0, source_cdf_tol=1e-8,
                nnu_ion=16, align_edges=1, nsource=1, tstar=[4.5e4], lum=[1e38], grid_type=0,
                nx=2, ny=2, nz=2, xmax=1, ymax=1, zmax=1, distance2cm=3.0856776e18,
                He_abund=0.1, te_fixed=1e4, ion_relax=1, te_min=1e3, te_max=5e4, caseA=1,
                lambda_ref=0.55)
    e16 = po.Model(dict(base, metal_eth=meth)).band()["eedge"]
    assert e16.size == 17
    # H I sits at eion_min, so it is not a candidate and O I (13.618 eV, 0.15%
    # above it) is an edge of its own, as build_aligned_edges has it
    inside = sorted(t for t in meth if 13.598 < t < 100.0)
    for t in inside + [24.587, 54.416]:
        near = np.min(np.abs(np.log(e16 / t)))
        if t == 24.383:                  # C II merged onto He I (within 1%)
            assert near < 0.01 and not np.any(np.isclose(e16, t))
        elif t in (35.121, 40.963):      # merged onto S III / Ar III (first kept)
            assert near < 0.01
        else:
            assert np.any(np.isclose(e16, t, rtol=1e-13, atol=0)), t
    # too few bins: metal edges are dropped, the H/He edges always stay
    e6 = po.Model(dict(base, nnu_ion=6, metal_eth=meth)).band()["eedge"]
    assert e6.size == 7
    for t in (24.587, 54.416):
        assert np.any(np.isclose(e6, t, rtol=1e-13, atol=0))
    assert np.all(np.diff(e6) > 0)
    # no metals: the H/He-only band
    e0 = po.Model(dict(base)).band()["eedge"]
    assert not np.array_equal(e0, e16)
