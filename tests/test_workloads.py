"""Seeded synthetic inputs: the lognormal field and the octree leaf lists."""
import numpy as np

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
