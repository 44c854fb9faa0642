"""Synthetic model set-ups of the benchmark configurations (parameters and
seeded inputs).  Units as in the reference: code length = parsec
(par%distance_unit = 'pc'), densities in cm^-3, luminosities in erg/s."""
import numpy as np

from .lib import MC_FP64, MC_GRID_CAR

PC2CM = 3.0856776e18
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
