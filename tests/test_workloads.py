"""Seeded synthetic inputs: the lognormal field and the octree leaf lists."""
import numpy as np
import pytest

from mochii_amd.workloads import lognormal_field, octree_from_field, synthetic_kext


def test_lognormal_field_moments_and_seed():
    f = lognormal_field(32, mean=100.0, sigma_ln=0.5, seed=7)
    assert abs(f.mean() / 100.0 - 1) < 1e-12
    assert abs(np.log(f).std() - 0.5) < 1e-12
    assert np.array_equal(f, lognormal_field(32, 100.0, 0.5, seed=7))
    assert not np.array_equal(f, lognormal_field(32, 100.0, 0.5, seed=8))


def test_octree_leaf_list_tiles_the_box_and_keeps_mass():
    f = lognormal_field(32, 100.0, 0.8, seed=3)
    x, y, z, lev, nh = octree_from_field(f, 4.0, lmax=5, thresh_dex=0.3)
    w = 8.0 / 2.0 ** lev
    assert abs((w**3).sum() - 512.0) < 1e-9                    # volumes tile the box
    assert abs((nh * w**3).sum() / 512.0 - f.mean()) < 1e-9    # mass conserved
    # leaves do not overlap: every fine cell belongs to exactly one leaf
    n = 32
    cover = np.zeros((n, n, n), dtype=int)
    for xi, yi, zi, li in zip(x, y, z, lev):
        s = n // 2**li
        i0 = int(round((xi + 4.0) / 8.0 * n - s / 2))
        j0 = int(round((yi + 4.0) / 8.0 * n - s / 2))
        k0 = int(round((zi + 4.0) / 8.0 * n - s / 2))
        cover[k0:k0 + s, j0:j0 + s, i0:i0 + s] += 1
    assert np.all(cover == 1)
    # a coarse leaf (level < 5) spans less than thresh_dex in log density
    lg = np.log10(f)
    for xi, yi, zi, li in zip(x, y, z, lev):
        if li < 5:
            s = n // 2**li
            i0 = int(round((xi + 4.0) / 8.0 * n - s / 2))
            j0 = int(round((yi + 4.0) / 8.0 * n - s / 2))
            k0 = int(round((zi + 4.0) / 8.0 * n - s / 2))
            blk = lg[k0:k0 + s, j0:j0 + s, i0:i0 + s]
            assert blk.max() - blk.min() <= 0.3 or li < 1


def test_synthetic_kext_table():
    lam, alb, g, cext = synthetic_kext()
    assert lam.size == 200 and np.all(np.diff(lam) > 0)
    assert lam[0] < 1.23984 / 100.0 and lam[-1] > 0.55
    assert np.allclose(alb, 0.4) and np.allclose(g, 0.6)
    assert np.allclose(np.diff(np.log(cext)) / np.diff(np.log(lam)), -1.2)


def test_multifractal_octree_small_tiles_box_and_refines_on_mass():
    """The cascade octree at a coarse threshold: leaves tile the box, the
    volume-weighted mean is the given mean, denser leaves sit deeper, and
    the field is seeded."""
    from mochii_amd.workloads import multifractal_octree
    x, y, z, lev, nh = multifractal_octree(lmax=6, lmin=2, half=8.0, mean=10.0, seed=3,
                                           mass_thr=1e-4, target=1e6)
    w = 16.0 / 2.0 ** lev
    assert abs((w ** 3).sum() - 16.0 ** 3) < 1e-6
    assert abs((nh * w ** 3).sum() / 16.0 ** 3 - 10.0) < 1e-9
    assert lev.min() >= 2 and lev.max() == 6 and np.unique(lev).size >= 3
    x2 = multifractal_octree(lmax=6, lmin=2, half=8.0, mean=10.0, seed=3, mass_thr=1e-4,
                             target=1e6)[0]
    assert np.array_equal(x, x2)
    x3 = multifractal_octree(lmax=6, lmin=2, half=8.0, mean=10.0, seed=4, mass_thr=1e-4,
                             target=1e6)[0]
    assert x3.size != x.size or not np.array_equal(x, x3)
    # denser leaves sit deeper: the median density grows with level
    med = [np.median(nh[lev == l]) for l in range(3, 7) if np.any(lev == l)]
    assert all(b > a for a, b in zip(med, med[1:]))


def test_large_workload_parameters():
    """The turbulent cloud's field moments and sources at a small grid size
    (the full-size cascade octree is built by the GPU tests and the bench)."""
    from mochii_amd.workloads import turbulent_cloud
    p, nH = turbulent_cloud(n=32, nphotons=1000)
    assert nH.size == 32 ** 3 and abs(nH.mean() / 50.0 - 1) < 1e-12
    assert abs(np.log(nH).std() - 1.0) < 1e-12
    assert p["nsource"] == 4 and all(3.5e4 <= t <= 5.0e4 for t in p["tstar"])
    assert len(p["metal_eth"]) == 15


@pytest.mark.parametrize("model,q_h", [("hii20", 1.004e49), ("hii40", 4.265e49)])
def test_lexington_inputs(model, q_h):
    """The band luminosities of tests/g2_hii/hii{20,40}_bench.in carry the
    benchmark's ionizing photon rate (the inputs' own comments: Q_H = 1.004e49
    and 4.265e49 s^-1): Q = L_band Int B_nu / h nu / Int B_nu over
    13.598-100 eV; the hollow shell sits between 3e18 cm and R_out."""
    from scipy.integrate import quad
    from mochii_amd.workloads import LEXINGTON, PC2CM, lexington
    params, nH, (x, y, z, lev, half), els = lexington(model)
    T, L = params["tstar"][0], params["src_lum"][0]
    kT = 1.380649e-16 * T / 1.602176634e-12

    def planck(E, k):
        return E ** (3 - k) / np.expm1(E / kT)
    Q = L * quad(planck, 13.598, 100.0, args=(1,))[0] / quad(planck, 13.598, 100.0, args=(0,))[0]
    Q /= 1.602176634e-12
    assert abs(Q / q_h - 1.0) < 2e-3, Q
    r = np.sqrt(x * x + y * y + z * z) * PC2CM
    gas = nH > 0
    assert half * PC2CM == pytest.approx(LEXINGTON[model][2], rel=1e-15)
    assert r[gas].min() >= 3.0e18 and r[gas].max() < LEXINGTON[model][2]
    assert np.all(nH[gas] == 100.0) and lev.min() == lev.max() == 6
    assert params["hei_diffuse"] == 1 and params["diffuse_field"] == 1
