"""The oracle's recombination cooling and free-free Gaunt factor against
vectors of the reference's own Python mirror of src/cooling_mod.f90
(tools/python/mochii_rates.py, held to the Fortran at 1e-10 by the
reference's tests/rates/check_python_rates.py), written by
tests/golden/make_rates_fixtures.py.  Each term is isolated from the
oracle's cooling_total by the species it multiplies."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

TAB = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "rates", "cooling_rates.txt"))
T = TAB[:, 0]
KB = 1.380649e-16
EV2ERG = 1.602176634e-12
NEB = (T >= 1e3) & (T <= 1e5)        # the fits' validity range


def _recomb(T, caseA, xHeI, xHeII, yHe=1.0):
    """Cooling of ionised hydrogen (nH = ne = 1) plus helium at n_He = yHe in
    the given stages; free-free and line terms off (no Gaunt table, no line
    fit), no neutral hydrogen."""
    cl = po.Cooling.make(gaunt=False)
    return cl.total(T, 1.0, 1.0, 0.0, xHeI, xHeII, yHe, caseA)


@pytest.mark.parametrize("caseA,col", [(1, 1), (0, 2)])
def test_hydrogen_recombination_cooling(caseA, col):
    got = np.array([_recomb(t, caseA, 0.0, 0.0, yHe=0.0) for t in T[NEB]])
    np.testing.assert_allclose(got, TAB[NEB, col], rtol=1e-13, atol=0)


@pytest.mark.parametrize("caseA,col", [(1, 3), (0, 4)])
def test_helium_ii_recombination_cooling(caseA, col):
    """H II + He II minus the H II term and the He II collisional-ionisation
    cooling (54.416 eV per ionisation) leaves k T alpha_HeII."""
    L = po.lib()
    got = np.array([_recomb(t, caseA, 0.0, 1.0) - _recomb(t, caseA, 0.0, 0.0, yHe=0.0)
                    - L.orc_ci_HeII(t) * 54.416 * EV2ERG for t in T[NEB]])
    np.testing.assert_allclose(got, TAB[NEB, col], rtol=1e-11, atol=0)


@pytest.mark.parametrize("caseA,col", [(1, 5), (0, 6)])
def test_helium_iii_recombination_cooling(caseA, col):
    got = np.array([_recomb(t, caseA, 0.0, 0.0) - _recomb(t, caseA, 0.0, 0.0, yHe=0.0)
                    for t in T[NEB]])
    np.testing.assert_allclose(got, TAB[NEB, col], rtol=1e-12, atol=0)


def test_free_free_gaunt_table():
    """gbar_ff(T, Z) for Z = 1..4 over the table's 1e2-1e6 K: the table of the
    Maxwellian mean of the Hummer fit (built by the oracle) interpolated
    linearly in log10 T, against the nodes the reference dumped from its
    Fortran table.  The fit is single precision: the 64 quadrature inputs
    log10(h nu / Ryd) are rounded to float, where a last-place difference of
    the host's log10 / exp moves a node by a few 1e-8 (the node differences
    scatter in sign, +-3e-8, with no trend in T or Z); a wrong Z, node grid,
    u range or normalisation is off by >= 1e-4."""
    lt, gf = np.zeros(41), np.zeros(4 * 41)
    L = po.lib()
    L.orc_gff_table(po.ptr(lt), po.ptr(gf))
    for z in range(1, 5):
        got = np.array([L.orc_gbar_ff(po.ptr(lt), po.ptr(gf), t, z) for t in T])
        np.testing.assert_allclose(got, TAB[:, 6 + z], rtol=1e-7, atol=2e-7, err_msg=f"Z={z}")
