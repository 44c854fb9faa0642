"""The line table writer (lines_write's '<base>_lines.txt' layout,
src/lines_mod.f90:111-121, format a4,i4,f12.2,2es14.5,2x,a) and the oracle's
line sum rule: the n-level rows of an ion sum to n_e times its line-cooling
coefficient (nlevel_emissivities against nlevel_cooling_at, two functions of
the reference computing the same populations)."""
import numpy as np

from mochii_amd.atomic import load_line_tables, load_nlevel
from mochii_amd.lines import format_row
from oracle import pyoracle as po


def test_row_layout_matches_fortran_format():
    row = format_row("o", 3, 5008.24, 1.234567e37, 3.21, "")
    assert row == "   o   3     5008.24   1.23457E+37   3.21000E+00"
    row = format_row("h", 1, 4862.7, 2.0e36, 1.0, "Hbeta")
    assert row.endswith("  Hbeta") and row[:8] == "   h   1"


def test_emissivities_sum_to_cooling():
    atoms = load_nlevel()
    for key in ("o_3", "n_2", "s_2", "ne_3", "h_1"):
        a = po.NlAtom.make(atoms[key])
        for T in (5e3, 1e4, 2e4):
            for ne in (1.0, 1e3, 1e6):
                s = a.emis(T, ne).sum()
                assert abs(s / (ne * a.cooling_at(T, ne)) - 1) < 1e-12, (key, T, ne)


def test_recombination_tables_reproduce_their_nodes():
    """At a table node the log-bilinear interpolation returns the tabulated
    emissivity (to rounding of exp(log e))."""
    for key, tab in load_line_tables().items():
        t = po.Sh95.make(tab)
        for k in range(len(tab["lines"])):
            for i in (0, len(tab["T"]) // 2, len(tab["T"]) - 2):
                for j in (0, len(tab["NE"]) // 2, len(tab["NE"]) - 2):
                    want = tab["lines"][k]["e"][i][j]
                    got = t.emis(k, tab["T"][i], tab["NE"][j])
                    assert abs(got / want - 1) < 1e-14, (key, k, i, j)
