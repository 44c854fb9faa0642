"""The launch on the device, through the C-ABI: the Sobol/Owen uniforms of
both streams bit-exact against the reference's golden sequences
(tests/qmc/q0_scr_*.txt, q2_scr_*.txt, q2_dif_*.txt of the reference;
src/qmc_mod.f90:73-260), and the stellar launch records (position,
direction, energy, band bin, start leaf; src/ion_band_mod.f90:1880-1935,
2232-2385) against the FP64 oracle for point sources and both external-field
geometries."""
import os

import numpy as np
import pytest

from mochii_amd import MC_GRID_AMR
from mochii_amd.workloads import lognormal_field, octree_from_field, stromgren
from tests.cases import gpu_model, oracle_model

pytestmark = pytest.mark.gpu

QDIR = os.path.join(os.path.dirname(__file__), "golden", "qmc")


def _codes(u):
    """32-bit code of u = (code + 0.5) 2^-32 (16 printed digits resolve it)."""
    return np.rint(u * 2.0**32 - 0.5).astype(np.int64)


@pytest.mark.parametrize("fname,seed,ndim,stream", [
    ("q0_scr_12345.txt", 12345, 3, 1), ("q0_scr_777.txt", 777, 3, 1),
    ("q0_scr_2024.txt", 2024, 3, 1),
    ("q2_scr_12345.txt", 12345, 12, 1), ("q2_scr_777.txt", 777, 12, 1),
    ("q2_scr_999983.txt", 999983, 12, 1),
    ("q2_dif_12345.txt", 12345, 12, 2), ("q2_dif_777.txt", 777, 12, 2),
    ("q2_dif_999983.txt", 999983, 12, 2)])
def test_device_sobol_matches_reference_sequences(fname, seed, ndim, stream):
    tab = np.loadtxt(os.path.join(QDIR, fname))
    idx = tab[:, 0].astype(np.int64)
    assert np.array_equal(idx, np.arange(idx[0], idx[0] + idx.size))
    params, nH = stromgren(n=4, nphotons=1000, seed=seed)
    gm = gpu_model(params, nH)
    u = gm.qmc_uniforms(stream, int(idx[0]), idx.size, ndim)
    want = _codes(tab[:, 1:1 + ndim])
    got = _codes(u)
    assert np.array_equal(got, want), np.argwhere(got != want)[:4]
    # the device uniform is exactly (code + 0.5) 2^-32
    assert np.array_equal(u, (want + 0.5) / 2.0**32)


def test_device_sobol_high_indices_match_oracle():
    """Indices beyond the tabulated rows (2^20 .. and 2^32 + k, whose bits
    above 32 carry no direction number) against the oracle's sequence."""
    from oracle import pyoracle as po
    params, nH = stromgren(n=4, nphotons=1000, seed=4242)
    gm = gpu_model(params, nH)
    q = po.Qmc(4242)
    for first in (2**20 - 3, 987654, 2**31 - 2):
        for stream in (1, 2):
            u = gm.qmc_uniforms(stream, first, 64, 12)
            want = np.array([q.uniforms(first + i, 12, stream) for i in range(64)])
            assert np.array_equal(u, want), (first, stream)


def _compare_records(orc, gm, n):
    rec, bl = gm.launch_records(1, n)
    want = np.array([orc.launch(i) for i in range(1, n + 1)])
    # positions and directions: sin / cos / sqrt of the device maths library
    # against the host's, a few ulp of the unit vector apart at most
    for k in range(6):
        d = np.abs(rec[:, k] - want[:, k])
        tol = 1e-15 if k >= 3 else 4.5e-15 * np.maximum(np.abs(want[:, k]), 1.0)
        assert np.all(d <= tol), (k, d.max())
        print("component", k, "bit-equal fraction", np.mean(d == 0.0), "max diff", d.max())
    # energies: the same CDF inverse in the same arithmetic order
    assert np.array_equal(rec[:, 6], want[:, 6]), np.max(np.abs(rec[:, 6] / want[:, 6] - 1))
    assert np.array_equal(bl[:, 0], want[:, 7].astype(np.int32) - 1)
    # the start leaf: equal, except where a position lies within rounding of
    # a leaf face
    same = bl[:, 1] == want[:, 8].astype(np.int32) - 1
    assert np.mean(same) >= 0.999, np.mean(same)
    return rec, bl


def test_launch_records_point_sources_match_oracle():
    params, nH = stromgren(n=16, nphotons=4096)
    params.update(nsource=3, tstar=[4.0e4, 3.0e4, 5.5e4], src_lum=[1e38, 5e37, 2e38],
                  src_x=[0.0, 1.3, -2.1], src_y=[0.0, -0.7, 2.5], src_z=[0.0, 0.4, -1.9])
    orc, gm = oracle_model(params, nH), gpu_model(params, nH)
    rec, bl = _compare_records(orc, gm, 4096)
    k = np.linalg.norm(rec[:, 3:6], axis=1)
    assert np.allclose(k, 1.0, atol=1e-14)


@pytest.mark.parametrize("geometry", [1, 2])
def test_launch_records_external_field_on_octree_match_oracle(geometry):
    """A point source plus an external field entering through the bounding
    sphere (1) or the box faces (2), on a refined octree (the start leaf is
    found from the LDS-staged top levels)."""
    f = lognormal_field(16, 100.0, 0.8, seed=3)
    x, y, z, lev, nh = octree_from_field(f, 4.0, lmax=4, thresh_dex=0.3)
    octree = (x, y, z, lev, 4.0)
    params, _ = stromgren(n=2, nphotons=4096)
    params.update(grid_type=MC_GRID_AMR, ext_on=1, ext_geometry=geometry, ext_tstar=3.0e4,
                  ext_intensity=1e-3, rmax=3.9 if geometry == 1 else -999.0)
    orc = oracle_model(params, nh, octree=octree)
    gm = gpu_model(params, nh, octree=octree)
    rec, bl = _compare_records(orc, gm, 4096)
    assert np.all(np.abs(rec[:, :3]) <= 4.0)
