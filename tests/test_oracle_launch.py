"""Launch geometry of the oracle (gen_ion_photon_qmc, emit_point_qmc,
emit_external_rec_qmc, emit_external_sph_qmc of src/ion_band_mod.f90):
entry points on the bounding surface, Lambert inward directions
(<mu> = 2/3, <mu^2> = 1/2), area-weighted faces, uniform sphere, and the
luminosity split between components (ion_setup)."""
import numpy as np

from oracle import pyoracle as po

PC = 3.0856776e18


def _model(**kw):
    p = dict(nphotons=4096, qmc_seed=11, eion_min=13.598, eion_max=100.0, source_cdf_tol=1e-8,
             align_edges=1, nnu_ion=8, nsource=0, nx=8, ny=8, nz=8, xmax=2.0, ymax=2.0,
             zmax=2.0, distance2cm=PC, He_abund=0.1, te_fixed=1e4, ion_relax=1.0, te_min=1e3,
             te_max=5e4, ext_on=1, ext_geometry=2, ext_tstar=3.0e4, ext_intensity=1e-3,
             rmax=1.5)
    p.update(kw)
    return po.Model(p)


def _launch(m, n):
    return np.array([m.launch(i) for i in range(1, n + 1)])


def test_external_box_entry_points_and_lambert_directions():
    m = _model()
    out = _launch(m, 4096)
    pos, k = out[:, :3], out[:, 3:6]
    nudge = 1e-6 * 4.0
    face = np.argmax(np.abs(pos), axis=1)
    side = np.sign(pos[np.arange(len(pos)), face])
    # on a face (nudged inside by 1e-6 of the span) and inside the box
    assert np.allclose(np.abs(pos[np.arange(len(pos)), face]), 2.0 - nudge, atol=1e-12)
    assert np.all(np.abs(pos) <= 2.0)
    mu = -side * k[np.arange(len(pos)), face]           # cosine to the inward normal
    assert np.all(mu > 0)
    assert abs(mu.mean() - 2 / 3) < 3e-3 and abs((mu ** 2).mean() - 0.5) < 3e-3
    frac = np.bincount(face * 2 + (side > 0), minlength=6) / len(pos)
    assert np.max(np.abs(frac - 1 / 6)) < 2e-3          # equal faces on a cube
    np.testing.assert_allclose(np.linalg.norm(k, axis=1), 1.0, atol=1e-14)
    # band luminosity = pi I A for the legacy-shape field
    A = 6 * (4.0 * PC) ** 2
    assert abs(m.band()["Ltot"] / (np.pi * 1e-3 * A) - 1) < 1e-13


def test_external_sphere_entry_points_and_lambert_directions():
    m = _model(ext_geometry=1)
    out = _launch(m, 4096)
    pos, k = out[:, :3], out[:, 3:6]
    r = np.linalg.norm(pos, axis=1)
    assert np.allclose(r, 1.5 - 1e-6 * 4.0, atol=1e-12)
    rhat = pos / r[:, None]
    mu = -np.sum(rhat * k, axis=1)
    assert np.all(mu > 0)
    assert abs(mu.mean() - 2 / 3) < 3e-3 and abs((mu ** 2).mean() - 0.5) < 3e-3
    assert np.all(np.abs(rhat.mean(axis=0)) < 5e-3)      # uniform on the sphere
    assert abs(m.band()["Ltot"] / (np.pi * 1e-3 * 4 * np.pi * (1.5 * PC) ** 2) - 1) < 1e-13


def test_components_split_by_band_luminosity():
    """Two point sources + the external box field: the Sobol component draw
    gives each component its share of ion_Ltot; point packets start at
    their source and leave isotropically."""
    src = dict(nsource=2, tstar=[4.0e4, 5.0e4], lum=[2e37, 6e37], sx=[-0.5, 0.7],
               sy=[0.1, -0.3], sz=[0.2, 0.0])
    m = _model(ext_intensity=2e-2, **src)
    n = 8192
    out = _launch(m, n)
    pos, k = out[:, :3], out[:, 3:6]
    Lext = np.pi * 2e-2 * 6 * (4.0 * PC) ** 2
    L = np.array([2e37, 6e37, Lext])
    assert abs(m.band()["Ltot"] / L.sum() - 1) < 1e-13
    at = [np.all(np.isclose(pos, [sx, sy, sz], atol=0), axis=1)
          for sx, sy, sz in zip(src["sx"], src["sy"], src["sz"])]
    ext = ~(at[0] | at[1])
    frac = np.array([at[0].mean(), at[1].mean(), ext.mean()])
    assert np.max(np.abs(frac - L / L.sum())) < 2e-3
    for a in at:                                        # isotropic point emission
        assert np.all(np.abs(k[a].mean(axis=0)) < 0.03)
        assert abs((k[a, 2] ** 2).mean() - 1 / 3) < 0.02
    assert m.band()["qmc_ndim"] == 7
