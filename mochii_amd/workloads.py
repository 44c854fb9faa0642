"""Synthetic model set-ups of the benchmark configurations (parameters and
seeded inputs).  Units as in the reference: code length = parsec
(par%distance_unit = 'pc'), densities in cm^-3, luminosities in erg/s."""
import numpy as np

from .lib import MC_COOL_LOCAL_NE, MC_FP64, MC_GRID_CAR

PC2CM = 3.0856776e18

# H I collisional line-cooling fit Lambda(T) = T^-1/2 sum A_i exp(-T_i/T)
# [erg cm^3 s^-1 K^1/2, K], the four rows of the reference's
# data/atomic/cooling_h_1.txt (CHIANTI v11.0.2 low-density fit over
# 1e3-1e5 K, max fit error 0.52%); atomic data, carried as data.
HI_COOLING_FIT = ((4.42616622e-17, 1.184356e+05), (2.62917725e-17, 1.269656e+05),
                  (1.03385873e-16, 1.528382e+05), (2.26605668e-16, 2.577561e+05))
# Q(H) = 1e49 s^-1 for a 40000 K blackbody over 13.598-100 eV
# (tests/g1_stromgren/g1_stromgren.in:17-22 of the reference)
L_Q49_T40K = 3.177837e38


def stromgren(n=64, nphotons=1_000_000, precision=MC_FP64, nH=100.0, half=4.0,
              niter=100, tol=2e-3, seed=12345):
    """Pure-H Stromgren sphere on a uniform n^3 Cartesian grid of side 2*half pc,
    uniform nH, one 40000 K blackbody at the box centre with Q(H) = 1e49 s^-1,
    case B, T = 1e4 K, Sobol launch, convergence on the volume criterion."""
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=16, ion_align_edges=1, nsource=1,
                  tstar=[4.0e4], src_lum=[L_Q49_T40K], src_x=[0.0], src_y=[0.0], src_z=[0.0],
                  grid_type=MC_GRID_CAR, nx=n, ny=n, nz=n, xmax=half, ymax=half, zmax=half,
                  distance2cm=PC2CM, He_abund=0.0, te_fixed=1.0e4, ion_relax=1.0,
                  gas_tol=tol, conv_vol=1, gas_niter=niter, case_A=0,
                  xHI_init=1.0, xHeI_init=1.0, xHeII_init=0.0)
    nHarr = np.full(n ** 3, nH)
    return params, nHarr


# Trace metals of the H II region workloads: C N O Ne S at the reference's
# defaults (src/define.f90:847-851) and Ar, which the reference leaves at 0,
# at the solar Ar/H = 2.5e-6 (Asplund et al. 2009), giving the 15 transitions
# of C N O Ne S Ar.
METAL_ABUND = {"ar": 2.5e-6}


def elements_of(params):
    """The trace-metal registry a workload's params were built for ([] when
    the metal path is off)."""
    from .atomic import registry
    return registry(METAL_ABUND) if params.get("metal_eth") else []


def hii_thermal(n=32, nphotons=200_000, precision=MC_FP64, seed=12345, niter=100,
                tol=2e-3, tol_te=2e-3, half=8.0, octree=None, nH=None, diffuse=False,
                metals=False, cooling_model=MC_COOL_LOCAL_NE):
    """H + He (He/H = 0.1) H II region with thermal balance: one 45000 K
    blackbody at the box centre with Q(H) = 3e49 s^-1, case A, T_e solved in
    [1e3, 5e4] K, volume convergence criterion; with diffuse=True explicit
    diffuse ground-recombination packets (Milne spectra).  Uniform car grid
    unless a leaf list (octree) with per-leaf nH is given.  Line cooling:
    the reference default 'local_ne' (the fits suppressed by the n-level
    (T, n_e) tables; the model needs nlevel_setup) or 'low_density'."""
    lum = 3.0 * L_Q49_T40K * _mean_energy_ratio(4.5e4, 4.0e4)
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=16, ion_align_edges=1, nsource=1,
                  tstar=[4.5e4], src_lum=[lum], src_x=[0.0], src_y=[0.0], src_z=[0.0],
                  grid_type=MC_GRID_CAR if octree is None else 1, nx=n, ny=n, nz=n,
                  xmax=half, ymax=half, zmax=half, distance2cm=PC2CM, He_abund=0.1,
                  te_fixed=1.0e4, ion_relax=1.0, te_min=1.0e3, te_max=5.0e4, gas_tol=tol,
                  gas_tol_te=tol_te, conv_vol=1, gas_niter=niter, case_A=1, solve_te=1,
                  xHI_init=1.0, xHeI_init=1.0, xHeII_init=0.0, diffuse_field=int(diffuse),
                  diffuse_model=2, cooling_model=int(cooling_model))
    if metals:
        from .atomic import metal_thresholds, registry
        params["metal_eth"] = metal_thresholds(registry(METAL_ABUND))
    if nH is None:
        nH = np.full(n ** 3, 100.0)
    return params, nH


def synthetic_kext(nlam=200, lam_min=0.0912 / 8, lam_max=10.0, slope=-1.2, albedo=0.4,
                   g=0.6, cext_v=4.868e-22):
    """Smooth power-law dust optics table: C_ext/H proportional to
    lambda^slope normalised to cext_v at 0.55 um, constant albedo and <cos>,
    on a log grid covering the ionizing band and lambda_ref."""
    lam = np.geomspace(lam_min, lam_max, nlam)
    cext = cext_v * (lam / 0.55) ** slope
    return lam, np.full(nlam, albedo), np.full(nlam, g), cext


def dusty_nebula(n=128, nphotons=30_000_000, precision=MC_FP64, nH=300.0, half=4.0,
                 dgr=0.01, seed=12345, niter=100, tol=2e-3):
    """Uniform n^3 grid at nH, pure H, dust-to-gas mass ratio dgr with the
    synthetic table (albedo 0.4, HG g = 0.6, scattering on), two blackbody
    sources (40000 K and 50000 K, Q(H) = 1e49 s^-1 each) placed from a seed."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * half, 0.5 * half, size=(2, 3))
    # Q(H)=1e49 for each: L = Q <E> with <E> the band-mean photon energy
    lum = [L_Q49_T40K, L_Q49_T40K * _mean_energy_ratio(5.0e4, 4.0e4)]
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=16, ion_align_edges=1, nsource=2,
                  tstar=[4.0e4, 5.0e4], src_lum=lum, src_x=list(pos[:, 0]),
                  src_y=list(pos[:, 1]), src_z=list(pos[:, 2]),
                  grid_type=MC_GRID_CAR, nx=n, ny=n, nz=n, xmax=half, ymax=half, zmax=half,
                  distance2cm=PC2CM, He_abund=0.0, te_fixed=1.0e4, ion_relax=1.0,
                  gas_tol=tol, conv_vol=1, gas_niter=niter, case_A=0,
                  xHI_init=1.0, xHeI_init=1.0, xHeII_init=0.0,
                  ion_add_dust=1, ion_dust_scatter=1, lambda_ref=0.55)
    nHarr = np.full(n ** 3, nH)
    # 'global_dgr' dust law (src/grid_mod_amr.f90:274-275): rhokap = nH cext DGR d2cm
    # with cext the per-H extinction at lambda_ref for a dust-to-gas ratio of 1
    rhokap = nHarr * 4.868e-22 * (dgr / 0.01) * PC2CM
    return params, nHarr, rhokap


def _mean_energy_ratio(T1, T2, emin=13.598, emax=100.0):
    """<E>_photon(T1) / <E>_photon(T2) over the band, for scaling L to equal Q."""
    e = np.linspace(emin, emax, 20001)
    kb = 8.617333262145e-5

    def mean(T):
        f = e ** 3 / np.expm1(e / (kb * T))
        return np.trapezoid(f, e) / np.trapezoid(f / e, e)

    return mean(T1) / mean(T2)


def lognormal_field(n, mean=100.0, sigma_ln=0.5, seed=1, slope=-11.0 / 3.0):
    """Seeded lognormal density on an n^3 periodic grid: a Gaussian random
    field with power spectrum P(k) ~ k^slope, scaled to unit variance,
    exponentiated with log-standard deviation sigma_ln and normalised to the
    given mean.  Array index order [iz, iy, ix]."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n) * n
    kz, ky, kx = np.meshgrid(k, k, np.fft.rfftfreq(n) * n, indexing="ij")
    kk = np.sqrt(kx**2 + ky**2 + kz**2)
    amp = np.zeros_like(kk)
    amp[kk > 0] = kk[kk > 0] ** (slope / 2.0)
    noise = rng.normal(size=kk.shape) + 1j * rng.normal(size=kk.shape)
    g = np.fft.irfftn(noise * amp, s=(n, n, n), axes=(0, 1, 2))
    g = (g - g.mean()) / g.std()
    rho = np.exp(sigma_ln * g)
    return rho * (mean / rho.mean())


def octree_from_field(field, half, lmax=None, thresh_dex=0.3, lmin=1):
    """Leaf list of an octree over the box [-half, half]^3 refined where the
    log10 density range inside a cell exceeds thresh_dex (down to the field's
    own resolution, level log2(n)), never coarser than lmin.  Returns
    (x, y, z, level, nH) of the leaves, nH the mean over the covered cells."""
    n = field.shape[0]
    L = int(round(np.log2(n)))
    lmax = L if lmax is None else min(lmax, L)
    lg = np.log10(field)
    mins, maxs, sums = {L: lg}, {L: lg}, {L: field.astype(np.float64)}
    for l in range(L - 1, -1, -1):
        m = 2 ** l
        mins[l] = mins[l + 1].reshape(m, 2, m, 2, m, 2).min(axis=(1, 3, 5))
        maxs[l] = maxs[l + 1].reshape(m, 2, m, 2, m, 2).max(axis=(1, 3, 5))
        sums[l] = sums[l + 1].reshape(m, 2, m, 2, m, 2).sum(axis=(1, 3, 5))
    xs, ys, zs, lv, nh = [], [], [], [], []
    cand = np.zeros((1, 3), dtype=np.int64)          # (iz, iy, ix) at level 0
    for l in range(0, lmax + 1):
        iz, iy, ix = cand[:, 0], cand[:, 1], cand[:, 2]
        rng_dex = maxs[l][iz, iy, ix] - mins[l][iz, iy, ix]
        refine = (l < lmin) | ((rng_dex > thresh_dex) & (l < lmax))
        leaf = ~refine
        if leaf.any():
            w = 2.0 * half / 2 ** l
            xs.append(-half + (ix[leaf] + 0.5) * w)
            ys.append(-half + (iy[leaf] + 0.5) * w)
            zs.append(-half + (iz[leaf] + 0.5) * w)
            lv.append(np.full(leaf.sum(), l, dtype=np.int32))
            nh.append(sums[l][iz[leaf], iy[leaf], ix[leaf]] / (2 ** (L - l)) ** 3)
        r = cand[refine]
        if r.size == 0:
            break
        off = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        cand = (2 * r[:, None, :] + off[None, :, :]).reshape(-1, 3)
    return (np.concatenate(xs), np.concatenate(ys), np.concatenate(zs), np.concatenate(lv),
            np.concatenate(nh))


def hii_octree(n_field=128, nphotons=10_000_000, precision=MC_FP64, seed=12345, field_seed=1,
               half=8.0, lmax=7, thresh_dex=0.3, lmin=1, niter=100, tol=2e-3, tol_te=2e-3,
               diffuse=True, metals=True, cooling_model=MC_COOL_LOCAL_NE):
    """H + He H II region with thermal balance on an adaptive octree: n_H
    lognormal (mean 100 cm^-3, sigma_ln 0.5) from a seeded Gaussian random
    field with P(k) ~ k^-11/3 on n_field^3, leaves refined to at most lmax
    levels where the density varies by more than thresh_dex across a cell,
    He/H = 0.1, trace metals C N O Ne S Ar, one 45000 K blackbody with
    Q(H) = 3e49 s^-1 at the centre, explicit diffuse recombination packets
    (case A), nphotons stellar packets per iteration.  Returns (params, nH per leaf, octree leaf list
    (x, y, z, level, half))."""
    field = lognormal_field(n_field, mean=100.0, sigma_ln=0.5, seed=field_seed)
    xl, yl, zl, lev, nHl = octree_from_field(field, half=half, lmax=lmax, thresh_dex=thresh_dex,
                                             lmin=lmin)
    params, _ = hii_thermal(n=n_field, nphotons=nphotons, precision=precision, seed=seed,
                            niter=niter, tol=tol, tol_te=tol_te, half=half, octree=True,
                            nH=nHl, diffuse=diffuse, metals=metals,
                            cooling_model=cooling_model)
    return params, nHl, (xl, yl, zl, lev, half)


# the two Lexington benchmark inputs (tests/g2_hii/hii20_bench.in and
# hii40_bench.in): blackbody T [K], band luminosity 13.598-100 eV [erg/s],
# outer radius of the shell [cm]
LEXINGTON = {"hii20": (2.0e4, 2.5326e38, 9.5e18), "hii40": (4.0e4, 1.2734e39, 1.46e19)}


def lexington(model="hii40", nphotons=40_000_000, precision=MC_FP64, seed=1234, level=6):
    """The reference's Lexington benchmark inputs (tests/g2_hii/hii20_bench.in,
    hii40_bench.in; grid of tests/g2_hii/make_hii_grid.py): a uniform
    level-`level` octree over [-R_out, R_out]^3 pc with n_H = 100 cm^-3 in the
    hollow shell R_in = 3e18 cm <= r < R_out and 0 elsewhere, a blackbody at
    the centre (HII20: 20000 K, 2.5326e38 erg/s in 13.598-100 eV, R_out =
    9.5e18 cm; HII40: 40000 K, 1.2734e39 erg/s, R_out = 1.46e19 cm), 32
    threshold-aligned band bins, He/H = 0.1 and C N O Ne S at 2.2e-4 / 4e-5 /
    3.3e-4 / 5e-5 / 9e-6 (the reference defaults) without the metal electrons
    and photoheating, thermal balance in [3e3, 3e4] K with 'local_ne'
    cooling, the explicit diffuse field (case A, Milne spectra) with the He I
    excited-level channel (par%hei_diffuse), ionised start, 'vol' convergence
    at gas_tol 5e-3 / gas_tol_te 1e-2.  Returns (params, nH, octree, elements)."""
    from .atomic import metal_thresholds, registry
    tstar, lum, rout_cm = LEXINGTON[model]
    rin, rout = 3.0e18 / PC2CM, rout_cm / PC2CM
    n = 2 ** level
    c = -rout + (np.arange(n) + 0.5) * (2.0 * rout / n)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    r = np.sqrt(x * x + y * y + z * z)
    nH = np.where((r >= rin) & (r < rout), 100.0, 0.0)
    lev = np.full(x.size, level, dtype=np.int32)
    els = registry({})
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=32, ion_align_edges=1, nsource=1, tstar=[tstar],
                  src_lum=[lum], src_x=[0.0], src_y=[0.0], src_z=[0.0], grid_type=1,
                  nx=1, ny=1, nz=1, xmax=rout, ymax=rout, zmax=rout, distance2cm=PC2CM,
                  He_abund=0.1, te_fixed=1.0e4, ion_relax=1.0, te_min=3.0e3, te_max=3.0e4,
                  gas_tol=5.0e-3, gas_tol_te=1.0e-2, conv_vol=1, gas_niter=100, case_A=1,
                  solve_te=1, xHI_init=1.0e-3, xHeI_init=0.5, xHeII_init=0.5, diffuse_field=1,
                  diffuse_model=2, hei_diffuse=1, cooling_model=int(MC_COOL_LOCAL_NE),
                  metal_eth=metal_thresholds(els), metal_ne=0, metal_heat=0)
    return params, nH, (x, y, z, lev, rout), els


def lexington_hii40(nphotons=40_000_000, precision=MC_FP64, seed=1234, level=6):
    """lexington("hii40", ...)."""
    return lexington("hii40", nphotons=nphotons, precision=precision, seed=seed, level=level)


def ionised_weighted_te(state, nH):
    """T_e weighted by n_HII n_e over the gas leaves (equal volumes): the
    reference's <T[NpNe]> of its Lexington gate (tests/g2_hii/check_g2_hii.py)."""
    w = nH * (1.0 - state["xHI"]) * state["ne"]
    return float(np.sum(state["Te"] * w) / np.sum(w))


# ---- large workloads ---------------------------------------------------------
def _hash_normal(seed, level, ix, iy, iz):
    """Seeded standard normals keyed by (level, cell index): two splitmix64
    finaliser rounds over the packed key, Box-Muller on the two halves."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9) & M
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB) & M
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        k = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(level) * np.uint64(0x632BE59BD9B4E019)) & M
        k = mix(k ^ ix.astype(np.uint64))
        k = mix(k ^ (iy.astype(np.uint64) << np.uint64(21)))
        k = mix(k ^ (iz.astype(np.uint64) << np.uint64(42)))
    u1 = ((k >> np.uint64(40)).astype(np.float64) + 0.5) / 2.0**24
    u2 = (((k >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float64) + 0.5) / 2.0**24
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# the cascade levels and per-level log-widths of the multifractal density
# (three lognormal multipliers, on cells of levels 3, 6 and 9): the leaf
# densities of the default field span 1.2e5; MF_MASS_THR refines it to
# ~2.0e7 leaves on levels 5-10
CASCADE_LEVELS = (3, 6, 9)
CASCADE_SIGMA = (0.85, 0.65, 0.5)
MF_MASS_THR = 1.65e-7


def _cascade_at(ix, iy, iz, level, seed, mean):
    """Point density [cm^-3] at the centres of cells (ix, iy, iz) of a level
    of the 3-level lognormal cascade: n_H = mean x prod_k exp(s_k g_k -
    s_k^2 / 2) with g_k the seeded standard normal of the level-
    CASCADE_LEVELS[k] cell holding the point (E[n_H] = mean)."""
    rho = np.full(ix.shape, float(mean))
    for lc, s in zip(CASCADE_LEVELS, CASCADE_SIGMA):
        d = lc - level
        if d <= 0:
            jx, jy, jz = ix >> -d, iy >> -d, iz >> -d
        else:          # a coarser cell: the cascade cell at its centre (x >= c convention)
            h = 1 << (d - 1)
            jx, jy, jz = (ix << d) + h, (iy << d) + h, (iz << d) + h
        rho *= np.exp(s * _hash_normal(seed, lc, jx, jy, jz) - 0.5 * s * s)
    return rho


def _spread3(v):
    """Bits of v (< 2^21) spread to every third bit."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order(x, y, z, level, half, lmax):
    """Permutation that lists leaves in depth-first (Morton) order of their
    lower corners on the level-lmax grid: leaves close in space are close in
    the list."""
    n = 2 ** lmax
    w = 2.0 * half / n
    ix = np.floor((x + half) / w - 0.5 * 2.0 ** (lmax - level) + 0.5).astype(np.int64)
    iy = np.floor((y + half) / w - 0.5 * 2.0 ** (lmax - level) + 0.5).astype(np.int64)
    iz = np.floor((z + half) / w - 0.5 * 2.0 ** (lmax - level) + 0.5).astype(np.int64)
    key = _spread3(ix) | (_spread3(iy) << np.uint64(1)) | (_spread3(iz) << np.uint64(2))
    return np.argsort(key, kind="stable")


def multifractal_octree(lmax=10, lmin=3, half=50.0, mean=10.0, seed=5, mass_thr=None,
                        target=2.0e7, order="morton"):
    """Leaf list of an octree over [-half, half]^3 pc refined on the
    multifractal density: a cell is split while its point density times its
    volume (in units of the root volume) exceeds mass_thr, down to lmax and
    never coarser than lmin; leaf n_H is the density at the leaf centre.
    mass_thr defaults to the value that gives ~target leaves for the default
    field.  order 'morton' lists the leaves depth-first (spatial neighbours
    near each other in every per-leaf array), 'level' coarse to fine.
    Returns (x, y, z, level, nH)."""
    thr = MF_MASS_THR if mass_thr is None else mass_thr
    xs, ys, zs, lv, nh = [], [], [], [], []
    cand = np.zeros((1, 3), dtype=np.int64)
    off = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.int64)
    for l in range(0, lmax + 1):
        ix, iy, iz = cand[:, 0], cand[:, 1], cand[:, 2]
        rho = _cascade_at(ix, iy, iz, l, seed, mean)
        refine = (l < lmin) | ((rho * 8.0 ** (-l) > thr * mean) & (l < lmax))
        leaf = ~refine
        if leaf.any():
            w = 2.0 * half / 2 ** l
            xs.append(-half + (ix[leaf] + 0.5) * w)
            ys.append(-half + (iy[leaf] + 0.5) * w)
            zs.append(-half + (iz[leaf] + 0.5) * w)
            lv.append(np.full(int(leaf.sum()), l, dtype=np.int32))
            nh.append(rho[leaf])
        r = cand[refine]
        if r.size == 0:
            break
        if 8 * r.shape[0] > 8 * target:
            raise ValueError(f"multifractal_octree: {8 * r.shape[0]} cells at level {l + 1}; "
                             f"raise mass_thr")
        cand = (2 * r[:, None, :] + off[None, :, :]).reshape(-1, 3)
    x, y, z, lev, nh = (np.concatenate(xs), np.concatenate(ys), np.concatenate(zs),
                        np.concatenate(lv), np.concatenate(nh))
    # the volume-weighted mean density is the given mean
    nh *= mean / (np.sum(nh * 8.0 ** (-lev.astype(np.float64))))
    if order == "morton":
        p = morton_order(x, y, z, lev, half, lmax)
        x, y, z, lev, nh = x[p], y[p], z[p], lev[p], nh[p]
    return x, y, z, lev, nh


def multifractal_cloud(nphotons=1_000_000_000, precision=MC_FP64, seed=12345, field_seed=5,
                       half=50.0, lmax=10, nsource=16, dust=True, ir_packets=100_000_000,
                       niter=100, tol=2e-3, tol_te=2e-3, mass_thr=None,
                       cooling_model=MC_COOL_LOCAL_NE, order="morton"):
    """Large AMR model: a 10-level octree (~2e7 leaves) over a (2 half pc)^3
    box refined on the seeded multifractal density (3-level lognormal
    cascade, volume-weighted mean n_H 10 cm^-3, dynamic range ~1e5),
    nsource blackbody point sources at seeded positions in the central 80%
    of the box with T in [30000, 55000] K and Q(H) log-uniform in
    [3e48, 3e49] s^-1, plus an isotropic 30000 K external field entering
    through the inscribed sphere (10% of the stellar band luminosity); H + He
    (He/H = 0.1), thermal balance, case B; dust on (dust-to-gas 0.01, the
    synthetic power-law table to 1e4 um, HG scattering) with the dust
    thermal-infrared band (ir_packets per pass); nphotons packets per
    iteration.  Returns (params, nH, octree, rhokap, dust table)."""
    x, y, z, lev, nH = multifractal_octree(lmax=lmax, half=half, seed=field_seed,
                                           mass_thr=mass_thr, order=order)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.8 * half, 0.8 * half, size=(nsource, 3))
    T = rng.uniform(3.0e4, 5.5e4, nsource)
    Q = 10 ** rng.uniform(np.log10(3e48), np.log10(3e49), nsource)
    lum = [float(q / 1e49 * L_Q49_T40K * _mean_energy_ratio(t, 4.0e4)) for q, t in zip(Q, T)]
    rmax = 0.98 * half
    r_cm = rmax * PC2CM
    ext_I = 0.1 * sum(lum) / (np.pi * 4.0 * np.pi * r_cm ** 2)
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=16, ion_align_edges=1, nsource=nsource,
                  tstar=list(T), src_lum=lum, src_x=list(pos[:, 0]), src_y=list(pos[:, 1]),
                  src_z=list(pos[:, 2]), ext_on=1, ext_geometry=1, ext_tstar=3.0e4,
                  ext_intensity=float(ext_I), rmax=rmax, grid_type=1, nx=1, ny=1, nz=1,
                  xmax=half, ymax=half, zmax=half, distance2cm=PC2CM, He_abund=0.1,
                  te_fixed=1.0e4, ion_relax=1.0, te_min=1.0e3, te_max=5.0e4, gas_tol=tol,
                  gas_tol_te=tol_te, conv_vol=1, gas_niter=niter, case_A=0, solve_te=1,
                  xHI_init=1.0, xHeI_init=1.0, xHeII_init=0.0, cooling_model=int(cooling_model),
                  ion_add_dust=int(dust), ion_dust_scatter=int(dust), lambda_ref=0.55,
                  dust_no_photons=float(ir_packets), dust_niter=2, dust_tol=1e-3)
    rhokap = nH * 4.868e-22 * PC2CM if dust else None
    return params, nH, (x, y, z, lev, half), rhokap, synthetic_kext(lam_max=1.0e4)


def turbulent_cloud(n=256, nphotons=100_000_000, precision=MC_FP64, seed=12345, field_seed=4,
                    half=16.0, nsource=4, metals=True, niter=100, tol=2e-3, tol_te=2e-3,
                    cooling_model=MC_COOL_LOCAL_NE):
    """Turbulent cloud on a uniform n^3 Cartesian grid over a (2 half pc)^3
    box: lognormal n_H (mean 50 cm^-3, sigma_ln 1.0) from a seeded Gaussian
    random field with P(k) ~ k^-11/3, nsource blackbody sources at seeded
    positions in the central half of the box with T in [35000, 50000] K and
    Q(H) log-uniform in [1e48, 1e49] s^-1, H/He (He/H = 0.1) with the trace
    metals C N O Ne S Ar, thermal balance, case B; nphotons packets per
    iteration.  Returns (params, nH)."""
    rho = lognormal_field(n, mean=50.0, sigma_ln=1.0, seed=field_seed)
    nH = np.ascontiguousarray(rho.ravel())        # [iz, iy, ix]: the car leaf order
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * half, 0.5 * half, size=(nsource, 3))
    T = rng.uniform(3.5e4, 5.0e4, nsource)
    Q = 10 ** rng.uniform(48.0, 49.0, nsource)
    lum = [float(q / 1e49 * L_Q49_T40K * _mean_energy_ratio(t, 4.0e4)) for q, t in zip(Q, T)]
    params = dict(nphotons=nphotons, qmc_seed=seed, precision=precision, eion_min=13.598,
                  eion_max=100.0, nnu_ion=16, ion_align_edges=1, nsource=nsource,
                  tstar=list(T), src_lum=lum, src_x=list(pos[:, 0]), src_y=list(pos[:, 1]),
                  src_z=list(pos[:, 2]), grid_type=MC_GRID_CAR, nx=n, ny=n, nz=n,
                  xmax=half, ymax=half, zmax=half, distance2cm=PC2CM, He_abund=0.1,
                  te_fixed=1.0e4, ion_relax=1.0, te_min=1.0e3, te_max=5.0e4, gas_tol=tol,
                  gas_tol_te=tol_te, conv_vol=1, gas_niter=niter, case_A=0, solve_te=1,
                  xHI_init=1.0, xHeI_init=1.0, xHeII_init=0.0, cooling_model=int(cooling_model))
    if metals:
        from .atomic import metal_thresholds, registry
        params["metal_eth"] = metal_thresholds(registry(METAL_ABUND))
    return params, nH
