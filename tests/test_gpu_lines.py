"""Line luminosities on the GPU (mc_lines_compute: lines_write,
src/lines_mod.f90:37-352) against the oracle on the same state: the H I /
He II / He I recombination lines from their (T, n_e) tables and every C N O
Ne S Ar ion's n-level rows; and the written table."""
import numpy as np
import pytest

from mochii_amd import MC_FP64
from mochii_amd.atomic import load_line_tables, load_nlevel
from mochii_amd.lines import lines_write
from mochii_amd.workloads import HI_COOLING_FIT, elements_of, hii_octree, hii_thermal
from tests.cases import gpu_model, oracle_model

pytestmark = pytest.mark.gpu


def _converge(gm, n=3):
    gm.cooling_setup(HI_COOLING_FIT)
    for _ in range(n):
        gm.ionizing_field_and_rates()
        gm.gas_thermal_update()
    gm.ionizing_field_and_rates()


def _oracle_lines(params, nH, gm, octree=None, case="B"):
    orc = oracle_model(params, nH, octree=octree)
    st = gm.state()
    orc.set_state(**{k: st[k] for k in ("xHI", "xHeI", "xHeII", "ne", "Te")})
    for ie in range(len(elements_of(params))):
        g, h = gm.species_rates(ie)
        orc.set_species_rates(ie, g, h)
    tabs = load_line_tables()
    return orc.lines((tabs[f"sh95_hi_case{case}"], tabs[f"sh95_heii_case{case}"],
                      tabs["hei_porter_caseB"]), load_nlevel())


@pytest.mark.parametrize("case", ["A", "B"])
def test_lines_match_oracle_car(case, tmp_path):
    params, nH = hii_thermal(n=12, nphotons=40000, metals=True, diffuse=True)
    gm = gpu_model(params, nH)
    _converge(gm)
    gm.lines_setup(line_case=case)
    got = gm.lines_compute()
    want = _oracle_lines(params, nH, gm, case=case)
    assert got["L"].size == want.size > 100
    big = want > 1e-12 * want.max()
    rel = np.abs(got["L"][big] / want[big] - 1.0)
    print("lines:", want.size, "nonzero", big.sum(), "max rel", rel.max())
    assert rel.max() < 1e-10
    assert np.all(got["L"][~big] <= 1e-12 * want.max())
    # H beta: case B's SH95 emissivity at the leaves' (T, n_e)
    hb = got["labels"].index("Hbeta")
    assert got["wl"][hb] == pytest.approx(4862.7) and got["ion"][hb] == 101
    # the [O III] 5007 line is among the rows, brighter than 4959
    o3 = np.where(got["ion"] == 803)[0]
    wl = got["wl"][o3]
    k5007, k4959 = o3[np.argmin(abs(wl - 5008.2))], o3[np.argmin(abs(wl - 4960.3))]
    assert 2.5 < got["L"][k5007] / got["L"][k4959] < 3.5
    path = tmp_path / "model_lines.txt"
    lines_write(gm, str(path), got)
    rows = [ln for ln in open(path) if not ln.startswith("#")]
    assert len(rows) == got["L"].size
    assert rows[hb].split()[:2] == ["h", "1"] and float(rows[hb].split()[4]) == pytest.approx(1.0)


def test_lines_match_oracle_octree():
    params, nH, octree = hii_octree(n_field=32, nphotons=20000, lmax=5)
    gm = gpu_model(params, nH, octree=octree)
    _converge(gm, 2)
    gm.lines_setup()
    got = gm.lines_compute()
    want = _oracle_lines(params, nH, gm, octree=octree)
    big = want > 1e-12 * want.max()
    rel = np.abs(got["L"][big] / want[big] - 1.0)
    print("octree lines: max rel", rel.max())
    assert rel.max() < 1e-10
