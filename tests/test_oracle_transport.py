"""The oracle's walk, estimator and cell solve against what the mathematics
fixes: analytic single-ray sums, brute-force cell enumeration, the octree
walk on single-level trees, closed-form equilibria."""
import numpy as np
import pytest

from oracle import pyoracle as po
from mochii_amd.workloads import stromgren
from tests.cases import oracle_model


def _brute_cells(n, half, pos, k, nsamp=200000):
    """Cells a ray crosses, by dense sampling of the path (raster 1-based)."""
    d = 2 * half / n
    pos, k = np.asarray(pos), np.asarray(k)
    # exit distance of the box
    t_exit = np.min([((half if kk > 0 else -half) - p) / kk for p, kk in zip(pos, k) if kk != 0])
    t = (np.arange(nsamp) + 0.5) / nsamp * t_exit
    pts = pos[None, :] + t[:, None] * k[None, :]
    idx = np.clip(((pts + half) / d).astype(int), 0, n - 1)
    il = 1 + idx[:, 0] + n * (idx[:, 1] + n * idx[:, 2])
    keep = np.concatenate([[True], il[1:] != il[:-1]])
    return il[keep]


def test_dda_cells_equal_brute_force():
    params, nH = stromgren(n=8, nphotons=10)
    m = oracle_model(params, nH)
    rng = np.random.default_rng(5)
    for _ in range(50):
        pos = rng.uniform(-3.5, 3.5, 3)
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        cells, segs = m.trace_ray(*pos, *k)
        assert np.array_equal(cells, _brute_cells(8, 4.0, pos, k))
        # segments sum to the path length to the box edge
        t_exit = np.min([((4.0 if kk > 0 else -4.0) - p) / kk for p, kk in zip(pos, k)])
        assert abs(segs.sum() - t_exit) < 1e-12


def test_octree_single_level_equals_car_walk():
    """A single-level octree walked face to face visits the same leaves with
    the same segment lengths as the car walk (src/define.f90:665-672)."""
    n, half = 8, 4.0
    params, nH = stromgren(n=n, nphotons=10)
    car = oracle_model(params, nH)
    params_amr = dict(params, car_walk_amr=1)
    car_amr = oracle_model(params_amr, nH)
    # leaf list in raster order, level log2(n)
    c = -half + (np.arange(n) + 0.5) * (2 * half / n)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    oct_model = oracle_model(dict(params, grid_type=1), nH,
                             octree=(X.ravel(), Y.ravel(), Z.ravel(),
                                     np.full(n**3, 3, dtype=np.int32), half))
    rng = np.random.default_rng(9)
    for _ in range(30):
        pos = rng.uniform(-3.9, 3.9, 3)
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        c1, s1 = car_amr.trace_ray(*pos, *k)
        c2, s2 = oct_model.trace_ray(*pos, *k)
        c3, _ = car.trace_ray(*pos, *k)
        assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
        assert np.array_equal(c1, c3)


def test_octree_refined_walk_matches_brute_force():
    """Two-level tree (one octant refined): the face-neighbour walk visits the
    leaves a dense point sampling of the ray finds."""
    half = 4.0
    xs, ys, zs, lv = [], [], [], []
    for o in range(8):
        cx, cy, cz = [(-half / 2 if (o >> b) & 1 == 0 else half / 2) for b in range(3)]
        if o == 7:
            for q in range(8):
                xs.append(cx + (-1 if q & 1 == 0 else 1) * half / 4)
                ys.append(cy + (-1 if q & 2 == 0 else 1) * half / 4)
                zs.append(cz + (-1 if q & 4 == 0 else 1) * half / 4)
                lv.append(2)
        else:
            xs.append(cx); ys.append(cy); zs.append(cz); lv.append(1)
    params, _ = stromgren(n=2, nphotons=10)
    nl = len(xs)
    m = oracle_model(dict(params, grid_type=1), np.full(nl, 1.0),
                     octree=(np.array(xs), np.array(ys), np.array(zs), np.array(lv), half))
    xs, ys, zs = map(np.asarray, (xs, ys, zs))
    h = np.where(np.asarray(lv) == 1, half / 2, half / 4)
    rng = np.random.default_rng(3)
    for _ in range(40):
        pos = rng.uniform(-3.9, 3.9, 3)
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        cells, segs = m.trace_ray(*pos, *k)
        t_exit = np.min([((half if kk > 0 else -half) - p) / kk for p, kk in zip(pos, k)])
        t = (np.arange(100000) + 0.5) / 100000 * t_exit
        pts = pos[None, :] + t[:, None] * k[None, :]
        inside = ((np.abs(pts[:, None, 0] - xs[None]) <= h[None]) &
                  (np.abs(pts[:, None, 1] - ys[None]) <= h[None]) &
                  (np.abs(pts[:, None, 2] - zs[None]) <= h[None]))
        il = inside.argmax(axis=1) + 1
        keep = np.concatenate([[True], il[1:] != il[:-1]])
        assert np.array_equal(cells, il[keep])
        assert abs(segs.sum() - t_exit) < 1e-12


def test_uniform_medium_analytic_tally():
    """A neutral uniform medium: the direct walk's path estimator summed over
    a packet's leaves is w (1 - e^-tau_edge) / kappa (the telescoping sum of
    w (e^-tau_in - e^-tau_out)/kappa), and the energy ledger closes."""
    n = 16
    params, nH = stromgren(n=n, nphotons=4000)
    m = oracle_model(params, nH)
    m.set_state(xHI=np.full(n**3, 1e-3))     # optically thin-ish everywhere
    m.transport_pass(nthreads=2)
    T = m.tallies()
    e, a, s = T["ledger"]
    assert abs(a + s - e) <= 5e-13 * e
    # total jt over leaves = sum over packets L (1 - e^-tau_edge)/kappa is not
    # directly available; instead check a single-packet case on its own
    m1 = oracle_model(dict(params, nphotons=1), nH)
    m1.set_state(xHI=np.full(n**3, 1e-3))
    m1.transport_pass(nthreads=1)
    T1 = m1.tallies()
    launch = m1.launch(1)
    E = launch[6]
    sig = po.lib().orc_sigma_HI(E)
    kap = 100.0 * 1e-3 * sig * 3.0856776e18          # per pc
    L = m1.band()["Ltot"]
    x, y, z, kx, ky, kz = launch[:6]
    t_exit = np.min([((4.0 if kk > 0 else -4.0) - p) / kk for p, kk in zip((x, y, z), (kx, ky, kz))])
    want = L * (1 - np.exp(-kap * t_exit)) / kap
    assert abs(T1["jt"].sum() / want - 1) < 1e-12
    assert abs(T1["ledger"][1] / (L * (1 - np.exp(-kap * t_exit))) - 1) < 1e-12


def test_pure_h_equilibrium_closed_form():
    """Pure H (y_He = 0): the fixed point x = a ne/(G + (C + a) ne), ne = nH(1-x)
    is the root of a quadratic; one solve_ion_cell call lands on it."""
    L = po.lib()
    T = 1e4
    a, cH = L.orc_alphaB_HII(T), L.orc_ci_HI(T)
    for G in (1e-12, 1e-9, 1e-7):
        for nH in (1.0, 100.0, 1e4):
            x, _, _, ne = po.solve_ion_cell(G, 0.0, 0.0, nH, T, False, 0.0, 1.0, 1.0, 0.0)
            # (C + a) nH (1-x) x + G x - a nH (1-x) = 0
            A = -(cH + a) * nH
            B = (cH + a) * nH + G + a * nH
            Cq = -a * nH
            roots = np.roots([A, B, Cq])
            xr = roots[(roots >= 0) & (roots <= 1)].real[0]
            assert abs(x - xr) <= 1e-8 * max(xr, 1e-30) + 1e-14
            assert abs(ne - nH * (1 - x)) <= 1e-9 * ne


def test_he_equilibrium_balances():
    """With He: the converged fractions satisfy each stage's balance at the
    returned n_e to the solver tolerance."""
    L = po.lib()
    T = 1e4
    gH, g1, g2, nH, y = 1e-8, 3e-9, 1e-10, 100.0, 0.1
    x, x1, x2, ne = po.solve_ion_cell(gH, g1, g2, nH, T, False, y, 1.0, 1.0, 0.0)
    aH, a2, a3 = L.orc_alphaB_HII(T), L.orc_alphaB_HeII(T), L.orc_alphaB_HeIII(T)
    c1, c2 = L.orc_ci_HeI(T), L.orc_ci_HeII(T)
    x3 = 1 - x1 - x2
    assert abs(x1 * (g1 + c1 * ne) - a2 * ne * x2) < 1e-9 * x1 * (g1 + c1 * ne)
    assert abs(x2 * (g2 + c2 * ne) - a3 * ne * x3) < 1e-9 * x2 * (g2 + c2 * ne) + 1e-30
    ne_closure = nH * ((1 - x) + y * (x2 + 2 * x3))
    assert abs(ne - ne_closure) < 1e-9 * ne


def test_counter_rng_uniform_and_independent():
    u = np.array([po.rng_uniform(12345, 1, i, k) for i in range(2000) for k in range(8)])
    assert 0 < u.min() and u.max() < 1
    h, _ = np.histogram(u, bins=16, range=(0, 1))
    chi2 = ((h - u.size / 16) ** 2 / (u.size / 16)).sum()
    assert chi2 < 45      # 15 dof, p ~ 1e-4
    v = u.reshape(2000, 8)
    assert abs(np.corrcoef(v[:, 0], v[:, 1])[0, 1]) < 0.1
    assert po.rng_uniform(1, 1, 5, 0) != po.rng_uniform(2, 1, 5, 0)


def test_hg_moments():
    """The HG inversion used on scattering has <cos theta> = g and
    <cos^2 theta> = (1 + 2 g^2)/3."""
    for g in (0.3, 0.6, -0.4):
        u = (np.arange(200000) + 0.5) / 200000
        q = (1 - g * g) / (1 + g * (2 * u - 1))
        c = (1 + g * g - q * q) / (2 * g)
        assert abs(c.mean() - g) < 1e-4
        assert abs((c * c).mean() - (1 + 2 * g * g) / 3) < 1e-4


@pytest.mark.slow
def test_stromgren_radius_converges_near_analytic():
    """End to end on a 32^3 grid: the ionised volume's equivalent radius is
    within a cell of the 1D reference radius (3.0243 pc, the value the
    reference's dusty gate quotes for this configuration)."""
    params, nH = stromgren(n=32, nphotons=200000)
    m = oracle_model(params, nH)
    for it in range(60):
        m.transport_pass(nthreads=8)
        m.rates_compute()
        mx, dv = m.equilibrium_update()
        if dv < 2e-3:
            break
    x = m.state()["xHI"]
    vol = (8.0 / 32) ** 3
    R = (3 * (1 - x).sum() * vol / (4 * np.pi)) ** (1 / 3)
    assert abs(R - 3.0243) < 8.0 / 32
