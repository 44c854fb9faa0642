"""ctypes binding to the FP64 CPU oracle (test infrastructure only).

Loaded by the tests, the smoke check and the CPU-baseline leg of bench.py.
The product package never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmochii_ref.so")

MAXSRC = 16
GRID_CAR, GRID_AMR = 0, 1
GEOM_POINT, GEOM_EXT_SPH, GEOM_EXT_REC = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [
        ("nphotons", C.c_int64), ("qmc_seed", C.c_int),
        ("eion_min", C.c_double), ("eion_max", C.c_double), ("source_cdf_tol", C.c_double),
        ("nnu_ion", C.c_int), ("align_edges", C.c_int),
        ("nsource", C.c_int),
        ("tstar", C.c_double * MAXSRC), ("lum", C.c_double * MAXSRC),
        ("sx", C.c_double * MAXSRC), ("sy", C.c_double * MAXSRC), ("sz", C.c_double * MAXSRC),
        ("ext_on", C.c_int), ("ext_geometry", C.c_int),
        ("ext_tstar", C.c_double), ("ext_intensity", C.c_double), ("rmax", C.c_double),
        ("grid_type", C.c_int), ("car_walk_amr", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("distance2cm", C.c_double),
        ("He_abund", C.c_double), ("te_fixed", C.c_double), ("ion_relax", C.c_double),
        ("te_min", C.c_double), ("te_max", C.c_double),
        ("caseA", C.c_int),
        ("add_dust", C.c_int), ("dust_scatter", C.c_int),
        ("lambda_ref", C.c_double),
        ("diffuse_field", C.c_int), ("diffuse_model", C.c_int),
        ("nmetal", C.c_int), ("metal_eth", C.c_double * 64),
        ("ion_metal_abs", C.c_int), ("metal_ne", C.c_int), ("metal_heat", C.c_int),
        ("metal_freefree", C.c_int), ("cx_two_way", C.c_int), ("ci_dere", C.c_int),
        ("hei_diffuse", C.c_int),
    ]

    # metal switches at the reference defaults (src/define.f90:593-652, 801)
    METAL_DEFAULTS = dict(ion_metal_abs=1, metal_ne=1, metal_heat=1, metal_freefree=1,
                          cx_two_way=1, ci_dere=0)


class Cooling(C.Structure):
    """Mirror of orc_cooling (H I line fit + free-free Gaunt table)."""
    _fields_ = [("nfit", C.c_int), ("A", C.c_double * 16), ("Ti", C.c_double * 16),
                ("logT", C.c_double * 41), ("gff", C.c_double * (4 * 41)),
                ("hi_tab", C.c_void_p)]

    @classmethod
    def make(cls, fit=(), gaunt=True):
        """fit: (A_i, T_i) pairs; gaunt=False leaves the Gaunt table at zero,
        a number sets it to that constant."""
        c = cls()
        c.nfit = len(fit)
        for i, (a, t) in enumerate(fit):
            c.A[i], c.Ti[i] = a, t
        lt, gf = np.zeros(41), np.zeros(4 * 41)
        lib().orc_gff_table(ptr(lt), ptr(gf))
        c.logT[:] = lt
        if gaunt is True:
            c.gff[:] = gf
        elif gaunt is not False:
            c.gff[:] = np.full(4 * 41, float(gaunt))
        return c

    def total(self, T, nH, ne, xHI, xHeI, xHeII, yHe, caseA):
        return lib().orc_cooling_total(C.byref(self), T, nH, ne, xHI, xHeI, xHeII, yHe,
                                       int(caseA))


def gaunt_ff(z, log10Te, xlf):
    """Hummer (1988) fit of the free-free Gaunt factor, single precision."""
    x = np.ascontiguousarray(xlf, dtype=np.float32)
    g = np.zeros(x.size, dtype=np.float32)
    lib().orc_gaunt_ff(z, log10Te, ptr(x, C.c_float), ptr(g, C.c_float), x.size)
    return g


def gaunt_ff_mean(T, Z):
    return lib().orc_gaunt_ff_mean(T, float(Z))


class OrcCtx(C.Structure):
    """Leading fields of orc_ctx that Python reads directly."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    lib.orc_create.restype = C.c_void_p
    lib.orc_create.argtypes = [C.POINTER(OrcParams)]
    lib.orc_destroy.argtypes = [C.c_void_p]
    for name in ("orc_setup_band", "orc_grid_car"):
        getattr(lib, name).argtypes = [C.c_void_p]
        getattr(lib, name).restype = C.c_int
    lib.orc_grid_octree.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _ip, C.c_double]
    lib.orc_state_init.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double]
    lib.orc_set_dust_table.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
    lib.orc_dust_at.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
    lib.orc_bin_of.argtypes = [C.c_void_p, C.c_double]
    lib.orc_bin_of.restype = C.c_int
    lib.orc_transport_pass.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int]
    lib.orc_rates_compute.argtypes = [C.c_void_p]
    lib.orc_equilibrium_update.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_residual_grid.argtypes = [C.c_void_p, _dp, _dp]
    lib.orc_solve_ion_cell.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    lib.orc_trace_ray.argtypes = [C.c_void_p] + [C.c_double] * 6 + [_ip, _dp, C.c_int]
    lib.orc_trace_ray.restype = C.c_int
    lib.orc_launch.argtypes = [C.c_void_p, C.c_int64, _dp]
    lib.orc_qmc_uniforms.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, _dp]
    lib.orc_qmc_raw.argtypes = [C.c_void_p, C.c_uint64, C.c_int, _dp]
    lib.orc_qmc_setup.argtypes = [C.c_void_p, C.c_int]
    lib.orc_sobol_code.argtypes = [C.c_void_p, C.c_uint64, C.c_int]
    lib.orc_sobol_code.restype = C.c_uint32
    lib.orc_rng_uniform.argtypes = [C.c_uint32, C.c_int, C.c_uint64, C.c_uint32]
    lib.orc_rng_uniform.restype = C.c_double
    for name in ("orc_sigma_HI", "orc_sigma_HeI", "orc_sigma_HeII", "orc_alphaA_HII",
                 "orc_alphaB_HII", "orc_alphaA_HeII", "orc_alphaB_HeII", "orc_alphaA_HeIII",
                 "orc_alphaB_HeIII", "orc_ci_HI", "orc_ci_HeI", "orc_ci_HeII",
                 "orc_planck_shape"):
        f = getattr(lib, name)
        f.restype = C.c_double
        f.argtypes = [C.c_double] * (2 if name == "orc_planck_shape" else 1)
    lib.orc_ci_model_set.argtypes = [C.c_int]
    lib.orc_ci_model_set.restype = C.c_int
    lib.orc_cooling_setup.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.orc_thermal_update.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    lib.orc_set_solve_threads.argtypes = [C.c_void_p, C.c_int]
    for name, args, res in (
            ("orc_ups_eval", [C.c_void_p, C.c_int, C.c_double], C.c_double),
            ("orc_nlevel_cooling_at", [C.c_void_p, C.c_double, C.c_double], C.c_double),
            ("orc_nlevel_emissivities", [C.c_void_p, C.c_double, C.c_double, _dp], None),
            ("orc_nlevel_line_ident", [C.c_void_p, C.c_int], C.c_double),
            ("orc_nlcool_build", [C.c_void_p, C.c_void_p], None),
            ("orc_nlcool_coeff", [C.c_void_p, C.c_double, C.c_double], C.c_double),
            ("orc_nlcool_low_density", [C.c_void_p, C.c_double], C.c_double),
            ("orc_nlcool_apply", [C.c_void_p, C.c_double, C.c_double, C.c_double], C.c_double),
            ("orc_sh95_emis", [C.c_void_p, C.c_int, C.c_double, C.c_double], C.c_double),
            ("orc_set_stage_table", [C.c_void_p, C.c_int, C.c_int, C.c_void_p], None),
            ("orc_set_hi_table", [C.c_void_p, C.c_void_p], None),
            ("orc_lines", [C.c_void_p, C.c_void_p, C.c_void_p, _dp, C.c_int], C.c_int)):
        f = getattr(lib, name)
        f.argtypes, f.restype = args, res
    lib.orc_set_solve_threads.restype = None
    lib.orc_net_rate.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp, _dp]
    lib.orc_net_rate.restype = C.c_double
    lib.orc_gaunt_ff.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.c_int]
    lib.orc_gaunt_ff_mean.argtypes = [C.c_double, C.c_double]
    lib.orc_gaunt_ff_mean.restype = C.c_double
    lib.orc_gff_table.argtypes = [_dp, _dp]
    lib.orc_gbar_ff.argtypes = [_dp, _dp, C.c_double, C.c_int]
    lib.orc_gbar_ff.restype = C.c_double
    lib.orc_cooling_total.argtypes = [C.POINTER(Cooling)] + [C.c_double] * 7 + [C.c_int]
    lib.orc_cooling_total.restype = C.c_double
    lib.orc_milne_setup.argtypes = [C.c_void_p, C.c_double]
    lib.orc_milne_energy_per_recomb.argtypes = [C.c_void_p, C.c_int, C.c_double]
    lib.orc_milne_energy_per_recomb.restype = C.c_double
    lib.orc_milne_sample_energy.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    lib.orc_milne_sample_energy.restype = C.c_double
    lib.orc_diffuse_setup.argtypes = [C.c_void_p]
    lib.orc_diffuse_setup.restype = C.c_int
    lib.orc_diffuse_build.argtypes = [C.c_void_p]
    lib.orc_launch_diffuse.argtypes = [C.c_void_p, C.c_int64, _dp]
    lib.orc_diffuse_hei_setup.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    lib.orc_diffuse_hei_setup.restype = C.c_int
    lib.orc_hei_branch_fractions.argtypes = [C.c_double, _dp, _dp, _dp]
    lib.orc_hei_branch_fractions.restype = None
    lib.orc_hei_584_ionizes_H.argtypes = [C.c_double] * 3
    lib.orc_hei_584_ionizes_H.restype = C.c_double
    lib.orc_hei2ph_setup.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double]
    lib.orc_hei2ph_setup.restype = C.c_int
    lib.orc_hei2ph_sample_energy.argtypes = [C.c_void_p, C.c_double]
    lib.orc_hei2ph_sample_energy.restype = C.c_double
    lib.orc_hei_branch_energies.argtypes = [C.c_void_p] + [C.c_double] * 3 + [_dp] * 3
    lib.orc_hei_branch_energies.restype = None
    lib.orc_transport_pass.restype = C.c_int
    vp = C.c_void_p
    lib.orc_species_setup.argtypes = [vp, C.c_int, vp]
    lib.orc_species_setup.restype = C.c_int
    lib.orc_gas_opacity_fill.argtypes = [vp]
    lib.orc_species_ne_prepare.argtypes = [vp, C.c_int, C.c_double, vp]
    lib.orc_species_metal_sums.argtypes = [vp, vp] + [C.c_double] * 4 + [_dp, _dp, _dp]
    lib.orc_solve_ion_cell_leaf.argtypes = ([vp, C.c_int] + [C.c_double] * 5 +
                                            [C.c_int, C.c_double, _dp, _dp, _dp, _dp])
    for name in ("orc_metal_cooling", "orc_metal_cooling_H", "orc_metal_freefree",
                 "orc_metal_heating"):
        getattr(lib, name).argtypes = [vp, C.c_int] + [C.c_double] * 5
        getattr(lib, name).restype = C.c_double
    lib.orc_species_fractions.argtypes = [vp, C.c_int, C.c_int] + [C.c_double] * 4 + [_dp]
    lib.orc_charge_neutrality_ne.argtypes = [vp, C.c_int] + [C.c_double] * 5
    lib.orc_charge_neutrality_ne.restype = C.c_double
    lib.orc_element_size.restype = C.c_int
    lib.orc_sigma_vy95.argtypes = [C.c_double] * 8
    lib.orc_sigma_vy95.restype = C.c_double
    lib.orc_sigma_vfky96.argtypes = [C.c_double] * 9
    lib.orc_sigma_vfky96.restype = C.c_double
    lib.orc_ci_dere_ratio.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.orc_ci_dere_ratio.restype = C.c_double
    lib.orc_voronov_ci.argtypes = [_dp, C.c_double]
    lib.orc_voronov_ci.restype = C.c_double
    lib.orc_alpha_rec.argtypes = [vp, C.c_int, C.c_double]
    lib.orc_alpha_rec.restype = C.c_double
    lib.orc_cx_rate.argtypes = [C.c_int, _dp, C.c_double]
    lib.orc_cx_rate.restype = C.c_double
    lib.orc_species_sigma.argtypes = [vp, C.c_int, C.c_double]
    lib.orc_species_sigma.restype = C.c_double
    lib.orc_build_planck.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, C.c_int]
    lib.orc_build_tabulated.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    lib.orc_sample_energy.argtypes = [C.c_void_p, C.c_double]
    lib.orc_sample_energy.restype = C.c_double
    lib.orc_energy_cdf_value.argtypes = [C.c_void_p, C.c_double]
    lib.orc_energy_cdf_value.restype = C.c_double
    lib.orc_sampler_free.argtypes = [C.c_void_p]
    # dust thermal-infrared band
    for name in ("orc_dust_temp_setup", "orc_dust_emis_transport_setup"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = C.c_int
    lib.orc_dust_femit_of.argtypes = [vp, C.c_double]
    lib.orc_dust_femit_of.restype = C.c_double
    lib.orc_dust_temp_compute.argtypes = [vp, _dp]
    lib.orc_dust_heat_compute.argtypes = [vp, _dp]
    lib.orc_ir_leaf_spectrum.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    lib.orc_ir_leaf_spectrum.restype = C.c_double
    lib.orc_sample_linear_bin.argtypes = [C.c_double] * 5
    lib.orc_sample_linear_bin.restype = C.c_double
    lib.orc_ir_counts.argtypes = [vp, C.c_int, C.c_int64, _dp, _dp]
    lib.orc_ir_counts.restype = C.c_double
    lib.orc_ir_launch_state.argtypes = [vp, C.c_int, _dp, _dp, C.c_uint64, _dp]
    lib.orc_ir_transport.argtypes = [vp, C.c_int, C.c_int64, C.c_int]
    lib.orc_ir_transport.restype = C.c_int
    lib.orc_transport_dust_emission.argtypes = [vp, _dp, C.c_int64, C.c_int, C.c_double,
                                                C.c_int, _dp]
    lib.orc_transport_dust_emission.restype = C.c_int
    lib.orc_dust_ir_arrays.argtypes = [vp] + [_dp] * 8
    lib.orc_dust_ir_arrays.restype = C.c_int
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


# ---- QMC ----------------------------------------------------------------
class Qmc:
    SIZE = 32 * 12 * 4 + 12 * 2 * 4

    def __init__(self, seed):
        self.buf = C.create_string_buffer(self.SIZE)
        lib().orc_qmc_setup(self.buf, seed)

    def uniforms(self, idx, n, stream=1):
        u = np.zeros(n)
        lib().orc_qmc_uniforms(self.buf, idx, n, stream, ptr(u))
        return u

    def raw(self, idx, n):
        u = np.zeros(n)
        lib().orc_qmc_raw(self.buf, idx, n, ptr(u))
        return u

    def code(self, idx, d):
        return lib().orc_sobol_code(self.buf, idx, d)


# ---- samplers -------------------------------------------------------------
class _SamplerStruct(C.Structure):
    _fields_ = [("n", C.c_int), ("energy", _dp), ("density", _dp), ("cdf", _dp),
                ("total_integral", C.c_double)]


class Sampler:
    def __init__(self):
        self.s = _SamplerStruct()

    @classmethod
    def planck(cls, T, emin, emax, rtol=1e-8, thresholds=()):
        o = cls()
        thr = np.ascontiguousarray(thresholds, dtype=np.float64)
        rc = lib().orc_build_planck(C.byref(o.s), T, emin, emax, rtol,
                                    ptr(thr) if thr.size else None, thr.size)
        if rc:
            raise RuntimeError(f"orc_build_planck rc={rc}")
        return o

    @classmethod
    def tabulated(cls, e, f):
        o = cls()
        e = np.ascontiguousarray(e, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        rc = lib().orc_build_tabulated(C.byref(o.s), ptr(e), ptr(f), e.size)
        if rc:
            raise RuntimeError(f"orc_build_tabulated rc={rc}")
        return o

    def arrays(self):
        n = self.s.n
        return (np.ctypeslib.as_array(self.s.energy, (n,)).copy(),
                np.ctypeslib.as_array(self.s.density, (n,)).copy(),
                np.ctypeslib.as_array(self.s.cdf, (n,)).copy(), self.s.total_integral)

    def sample(self, u):
        return lib().orc_sample_energy(C.byref(self.s), u)

    def cdf(self, e):
        return lib().orc_energy_cdf_value(C.byref(self.s), e)

    def __del__(self):
        try:
            lib().orc_sampler_free(C.byref(self.s))
        except Exception:
            pass


# ---- model ----------------------------------------------------------------
class Model:
    """One oracle context: params, band, grid, leaf state, tallies."""

    def __init__(self, params: dict, octree=None, nH=None, dust_table=None, elements=None):
        p = OrcParams()
        params = dict(OrcParams.METAL_DEFAULTS, **params)
        if elements:
            from mochii_amd.atomic import metal_thresholds
            params.setdefault("metal_eth", metal_thresholds(elements))
        if "metal_eth" in params:
            params["nmetal"] = len(params["metal_eth"])
        for k, v in params.items():
            if k in ("tstar", "lum", "sx", "sy", "sz", "metal_eth"):
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(p, k, v)
        self.p = p
        self.params = dict(params)
        L = lib()
        self.c = L.orc_create(C.byref(p))
        rc = L.orc_setup_band(self.c)
        if rc:
            raise RuntimeError(f"orc_setup_band rc={rc}")
        if octree is None:
            L.orc_grid_car(self.c)
        else:
            xl, yl, zl, lev, half = octree
            self._oct = [np.ascontiguousarray(a, dtype=np.float64) for a in (xl, yl, zl)]
            self._lev = np.ascontiguousarray(lev, dtype=np.int32)
            L.orc_grid_octree(self.c, self._lev.size, ptr(self._oct[0]), ptr(self._oct[1]),
                              ptr(self._oct[2]), ptr(self._lev, C.c_int), half)
        self.nleaf = C.cast(self.c, C.POINTER(_CtxHead)).contents.nleaf
        if dust_table is not None:
            lam, alb, g, cext = [np.ascontiguousarray(a, dtype=np.float64) for a in dust_table]
            L.orc_set_dust_table(self.c, lam.size, ptr(lam), ptr(alb), ptr(g), ptr(cext))
        self.elements = list(elements or [])
        if self.elements:
            self.species_setup(self.elements)
        if nH is not None:
            self.state_init(nH)

    # ---- trace metals -------------------------------------------------------
    def species_setup(self, elements):
        from mochii_amd.atomic import to_struct
        self.elements = list(elements)
        self._els = to_struct(self.elements)
        rc = lib().orc_species_setup(self.c, len(self.elements), self._els)
        if rc:
            raise RuntimeError(f"orc_species_setup rc={rc}")

    def opacity_fill(self):
        lib().orc_gas_opacity_fill(self.c)

    def _flat(self):
        h = self._head()
        return h.ntr, h.nst, list(h.tr_off)[:h.nel], list(h.st_off)[:h.nel]

    def species_fractions(self, ie):
        """stage_frac of element ie: (nleaf, nstage)."""
        ntr, nst, toff, soff = self._flat()
        ns = self.elements[ie]["nstage"]
        a = np.ctypeslib.as_array(self._head().mfrac, (self.nleaf * nst,)).reshape(self.nleaf, nst)
        return a[:, soff[ie]:soff[ie] + ns].copy()

    def species_rates(self, ie):
        """(egam, eheat) of element ie: (nleaf, nstage-1) each."""
        ntr, nst, toff, soff = self._flat()
        nt = self.elements[ie]["nstage"] - 1
        h = self._head()
        g = np.ctypeslib.as_array(h.mgam, (self.nleaf * ntr,)).reshape(self.nleaf, ntr)
        q = np.ctypeslib.as_array(h.mheat, (self.nleaf * ntr,)).reshape(self.nleaf, ntr)
        return g[:, toff[ie]:toff[ie] + nt].copy(), q[:, toff[ie]:toff[ie] + nt].copy()

    def species_tallies(self, ie):
        """raw(2, nstage-1) of element ie per leaf: (nleaf, nstage-1, 2)."""
        h = self._head()
        ntr = h.ntr
        raw = np.ctypeslib.as_array(h.mraw, (self.nleaf * ntr * 2,)).reshape(self.nleaf, ntr, 2)
        nt = self.elements[ie]["nstage"] - 1
        out = np.zeros((self.nleaf, nt, 2))
        for k in range(ntr):
            if h.flat_ie[k] == ie:
                out[:, h.flat_it[k], :] = raw[:, k, :]
        return out

    def set_species_rates(self, ie, gam, heat=None):
        ntr, nst, toff, soff = self._flat()
        h = self._head()
        nt = self.elements[ie]["nstage"] - 1
        g = np.ctypeslib.as_array(h.mgam, (self.nleaf * ntr,)).reshape(self.nleaf, ntr)
        g[:, toff[ie]:toff[ie] + nt] = gam
        if heat is not None:
            q = np.ctypeslib.as_array(h.mheat, (self.nleaf * ntr,)).reshape(self.nleaf, ntr)
            q[:, toff[ie]:toff[ie] + nt] = heat

    def metal_sums(self, il, T, ne, xHI):
        """species_ne_prepare + species_metal_sums: (nem, r_cx_in, r_cx_out)."""
        cc = C.create_string_buffer(8 * 5 * 64)
        L = lib()
        L.orc_species_ne_prepare(self.c, il, T, cc)
        nH = self.arr("nH")[il]
        a, b, d = C.c_double(), C.c_double(), C.c_double()
        L.orc_species_metal_sums(self.c, cc, nH, ne, nH * xHI, nH * (1.0 - xHI), C.byref(a),
                                 C.byref(b), C.byref(d))
        return a.value, b.value, d.value

    def solve_cell(self, il, T, x=None, gam=None):
        """solve_ion_cell for leaf il with its metal terms:
        (xHI, xHeI, xHeII, ne)."""
        st = self.state()
        x0 = x if x is not None else (st["xHI"][il], st["xHeI"][il], st["xHeII"][il])
        r = self.rates()
        g = gam if gam is not None else (r["gHI"][il], r["gHeI"][il], r["gHeII"][il])
        a, b, c, ne = C.c_double(x0[0]), C.c_double(x0[1]), C.c_double(x0[2]), C.c_double()
        lib().orc_solve_ion_cell_leaf(self.c, il, g[0], g[1], g[2], st["nH"][il], T,
                                      int(self.p.caseA), self.p.He_abund, C.byref(a),
                                      C.byref(b), C.byref(c), C.byref(ne))
        return a.value, b.value, c.value, ne.value

    def metal_terms(self, il, T, ne, xHI):
        """(metal_cooling, metal_cooling_H, metal_freefree, metal_heating) of
        leaf il at T, n_e, x_HI."""
        L = lib()
        nH = self.arr("nH")[il]
        args = (self.c, il, T, nH, ne, nH * xHI, nH * (1.0 - xHI))
        return (L.orc_metal_cooling(*args), L.orc_metal_cooling_H(*args),
                L.orc_metal_freefree(*args), L.orc_metal_heating(*args))

    def species_fractions_at(self, ie, il, T, ne, xHI):
        out = np.zeros(6)
        nH = self.arr("nH")[il]
        lib().orc_species_fractions(self.c, ie, il, T, ne, nH * xHI, nH * (1.0 - xHI), ptr(out))
        return out[:self.elements[ie]["nstage"]].copy()

    # raw pointer access to the context arrays
    def _head(self):
        return C.cast(self.c, C.POINTER(_CtxHead)).contents

    def arr(self, name, n=None):
        h = self._head()
        n = self.nleaf if n is None else n
        return np.ctypeslib.as_array(getattr(h, name), (n,))

    def state_init(self, nH, rhokap=None):
        nH = np.ascontiguousarray(nH, dtype=np.float64)
        P = self.params
        lib().orc_state_init(self.c, ptr(nH), P.get("xHI_init", 1.0), P.get("xHeI_init", 1.0),
                             P.get("xHeII_init", 0.0))
        if rhokap is not None:
            self.arr("rhokap")[:] = rhokap
        if self.elements:
            self.opacity_fill()          # gas_opacity_setup -> gas_opacity_fill

    def transport_pass(self, first=1, last=None, nthreads=1):
        if last is None:
            last = self.p.nphotons + 1
        lib().orc_transport_pass(self.c, first, last, nthreads)

    def rates_compute(self):
        lib().orc_rates_compute(self.c)

    def equilibrium_update(self):
        a, b = C.c_double(), C.c_double()
        lib().orc_equilibrium_update(self.c, C.byref(a), C.byref(b))
        return a.value, b.value

    def cooling_setup(self, A, Ti):
        A = np.ascontiguousarray(A, dtype=np.float64)
        Ti = np.ascontiguousarray(Ti, dtype=np.float64)
        lib().orc_cooling_setup(self.c, A.size, ptr(A), ptr(Ti))

    def nlevel_setup(self, atoms):
        """cooling_model 'local_ne': the (T, n_e) tables of H I and of every
        registry stage with a cooling fit whose atom is in ``atoms``
        ({'<el>_<stage>': atom dict}); tables are built once per atom."""
        L = lib()
        self._nlc = []

        def table(key):
            if key not in atoms:
                return None
            t = _NLC_CACHE.get((key, id(atoms[key])))
            if t is None:
                t = NlCool.build(NlAtom.make(atoms[key]))
                _NLC_CACHE[(key, id(atoms[key]))] = t
            self._nlc.append(t)
            return t

        t = table("h_1")
        L.orc_set_hi_table(self.c, C.byref(t) if t is not None else None)
        for ie, el in enumerate(self.elements):
            for st in range(1, el["nstage"] + 1):
                t = table(f"{el['name']}_{st}") if st in el.get("cooling", {}) else None
                L.orc_set_stage_table(self.c, ie, st, C.byref(t) if t is not None else None)

    def lines(self, tables, atoms, nthreads=None):
        """lines_write's luminosities [erg/s]: tables = (H I, He II, He I)
        recombination tables (dicts or None), atoms = {'<el>_<stage>': atom};
        the reference's table order."""
        self._lt = [Sh95.make(t) if t is not None else None for t in tables]
        tp = (C.c_void_p * 3)(*[C.addressof(t) if t is not None else None for t in self._lt])
        self._la = []
        ap = (C.c_void_p * (ORC_MAX_EL_PY * ORC_MAX_ST_PY))()
        for ie, el in enumerate(self.elements):
            for st in range(1, el["nstage"] + 1):
                key = f"{el['name']}_{st}"
                if key in atoms:
                    a = NlAtom.make(atoms[key])
                    self._la.append(a)
                    ap[ie * ORC_MAX_ST_PY + st - 1] = C.addressof(a)
        L = np.zeros(4096)
        n = lib().orc_lines(self.c, tp, ap, ptr(L), int(nthreads or min(os.cpu_count() or 1, 16)))
        return L[:n].copy()

    def thermal_update(self, nthreads=None):
        """gas_thermal_update; the per-leaf solves on nthreads threads (default
        all host cores, capped at 16); the result does not depend on it."""
        n = nthreads or min(os.cpu_count() or 1, 16)
        lib().orc_set_solve_threads(self.c, int(n))
        a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        lib().orc_thermal_update(self.c, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def net_rate(self, il, T, x=None):
        """Net heating - cooling of leaf il at T (ionisation solved at T,
        warm start from x = (xHI, xHeI, xHeII) or the leaf's state)."""
        st = self.state()
        x0 = x if x is not None else (st["xHI"][il], st["xHeI"][il], st["xHeII"][il])
        a, b, c, ne = C.c_double(x0[0]), C.c_double(x0[1]), C.c_double(x0[2]), C.c_double()
        net = lib().orc_net_rate(self.c, il, T, C.byref(a), C.byref(b), C.byref(c), C.byref(ne))
        return net, (a.value, b.value, c.value, ne.value)

    def residual_grid(self):
        a, b = C.c_double(), C.c_double()
        lib().orc_residual_grid(self.c, C.byref(a), C.byref(b))
        return a.value, b.value

    def tallies(self):
        h = self._head()
        nl, nnu = self.nleaf, h.nnu_band
        return dict(
            hhe=np.ctypeslib.as_array(h.hhe, (nl * 6,)).reshape(nl, 6).copy(),
            jt=np.ctypeslib.as_array(h.jt, (nl * nnu,)).reshape(nl, nnu).copy(),
            dust_abs=np.ctypeslib.as_array(h.dust_abs, (nl,)).copy(),
            seg_direct=h.seg_direct, seg_scattered=h.seg_scattered,
            ledger=(h.e_emit, h.e_abs, h.e_esc))

    def rates(self):
        return {k: self.arr(k).copy() for k in ("gHI", "gHeI", "gHeII", "hHI", "hHeI", "hHeII")}

    def state(self):
        return {k: self.arr(k).copy() for k in ("nH", "xHI", "xHeI", "xHeII", "ne", "Te")}

    def set_state(self, **kw):
        for k, v in kw.items():
            self.arr(k)[:] = v

    def band(self):
        h = self._head()
        return dict(nnu_band=h.nnu_band,
                    eedge=np.ctypeslib.as_array(h.eedge, (h.nnu_band + 1,)).copy(),
                    Ltot=h.Ltot, qmc_ndim=h.qmc_ndim, ncomp=h.ncomp)

    def bin_of(self, E):
        return lib().orc_bin_of(self.c, E)

    def dust_at(self, E):
        a, b, g = C.c_double(), C.c_double(), C.c_double()
        lib().orc_dust_at(self.c, E, C.byref(a), C.byref(b), C.byref(g))
        return a.value, b.value, g.value

    def trace_ray(self, x, y, z, kx, ky, kz, maxn=100000):
        cells = np.zeros(maxn, dtype=np.int32)
        segs = np.zeros(maxn)
        n = lib().orc_trace_ray(self.c, x, y, z, kx, ky, kz, ptr(cells, C.c_int), ptr(segs), maxn)
        return cells[:n].copy(), segs[:n].copy()

    def launch(self, pid):
        out = np.zeros(9)
        lib().orc_launch(self.c, pid, ptr(out))
        return out

    def diffuse_build(self):
        """Emission table from the current state; returns (L_diffuse, n_packets)."""
        L = lib()
        if not self._head().milne and L.orc_diffuse_setup(self.c):
            raise RuntimeError("orc_diffuse_setup failed")
        L.orc_diffuse_build(self.c)
        h = self._head()
        return h.diffuse_lum, h.diffuse_nphot

    def diffuse_tables(self):
        h = self._head()
        return (np.ctypeslib.as_array(h.dif_ch, (self.nleaf * 4,)).reshape(self.nleaf, 4).copy(),
                np.ctypeslib.as_array(h.dif_cdf, (self.nleaf,)).copy())

    def diffuse_hei_setup(self, y, A):
        """The He I two-photon shape for the diffuse field's excited channel
        (hei_2ph_setup, src/hei_twophoton_mod.f90:98-183)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        A = np.ascontiguousarray(A, dtype=np.float64)
        if lib().orc_diffuse_hei_setup(self.c, y.size, ptr(y), ptr(A)):
            raise RuntimeError("orc_diffuse_hei_setup failed")

    def launch_diffuse(self, pid):
        out = np.zeros(9)
        lib().orc_launch_diffuse(self.c, pid, ptr(out))
        return out

    # ---- dust thermal-infrared band -----------------------------------------
    def dust_temp_setup(self):
        rc = lib().orc_dust_temp_setup(self.c)
        if rc:
            raise RuntimeError(f"orc_dust_temp_setup rc={rc}")

    def femit_of(self, T):
        return lib().orc_dust_femit_of(self.c, T)

    def dust_heat(self):
        """Heat_dust of the last ionizing pass (ion_score_apply_dust_heat)."""
        out = np.zeros(self.nleaf)
        lib().orc_dust_heat_compute(self.c, ptr(out))
        return out

    def dust_temp_compute(self, heat):
        heat = np.ascontiguousarray(heat, dtype=np.float64)
        lib().orc_dust_temp_compute(self.c, ptr(heat))
        return self.ir_arrays()["tdust"]

    def dust_emis_transport_setup(self):
        rc = lib().orc_dust_emis_transport_setup(self.c)
        if rc:
            raise RuntimeError(f"orc_dust_emis_transport_setup rc={rc}")

    def _ir_head(self):
        return C.cast(self._head().ir, C.POINTER(_IrHead)).contents

    def ir_arrays(self):
        n = self.nleaf
        ih = self._ir_head()
        nl = ih.nl
        out = dict(tdust=np.zeros(n), heat_ir=np.zeros(n), irabs=np.zeros(n),
                   lesc=np.zeros(nl), lam=np.zeros(nl), sabs=np.zeros(nl),
                   edge=np.zeros(nl + 1), femit=np.zeros(2 * 61))
        lib().orc_dust_ir_arrays(self.c, *[ptr(out[k]) for k in
                                           ("tdust", "heat_ir", "irabs", "lesc", "lam", "sabs",
                                            "edge", "femit")])
        out["tgrid"], out["femit"] = out["femit"][:61].copy(), out["femit"][61:].copy()
        out.update(nzero=ih.nzero, lam_hardest=ih.lam_hardest,
                   frac_discarded=ih.frac_discarded, emitted=ih.emitted, extinct=ih.extinct,
                   escaped=ih.escaped, Ltot=ih.Ltot, Lpacket=ih.Lpacket, xi=ih.xi,
                   seg_direct=ih.seg_direct, seg_scattered=ih.seg_scattered,
                   khi=np.ctypeslib.as_array(ih.khi, (n,)).copy() if ih.khi else None)
        return out

    def ir_leaf_spectrum(self, il):
        nl = self._ir_head().nl
        pdf, cdf, fd = np.zeros(nl), np.zeros(nl), C.c_double()
        esum = lib().orc_ir_leaf_spectrum(self.c, il, ptr(pdf), ptr(cdf), C.byref(fd))
        return pdf, cdf, esum, fd.value

    def ir_counts(self, npass, N, heat):
        heat = np.ascontiguousarray(heat, dtype=np.float64)
        xi = C.c_double()
        L = lib().orc_ir_counts(self.c, npass, N, ptr(heat), C.byref(xi))
        return L, xi.value, self.ir_arrays()["khi"]

    def ir_launch(self, il, pdf, cdf, key):
        out = np.zeros(7)
        lib().orc_ir_launch_state(self.c, il, ptr(np.ascontiguousarray(pdf)),
                                  ptr(np.ascontiguousarray(cdf)), key, ptr(out))
        return out

    def transport_dust_emission(self, heat, npacket=1000000, niter=2, tol=1e-3, nthreads=4):
        """transport_dust_emission: returns (heat_dust out, report dict)."""
        heat = np.array(heat, dtype=np.float64)
        rep = np.zeros(10)
        rc = lib().orc_transport_dust_emission(self.c, ptr(heat), npacket, niter, tol, nthreads,
                                               ptr(rep))
        if rc:
            raise RuntimeError(f"orc_transport_dust_emission rc={rc}")
        return heat, ir_report(rep)

    def __del__(self):
        try:
            if self.c:
                lib().orc_destroy(self.c)
                self.c = None
        except Exception:
            pass


class _CtxHead(C.Structure):
    """Mirror of orc_ctx (mc_oracle.h)."""
    _fields_ = [
        ("p", OrcParams),
        ("qmc", C.c_uint32 * (32 * 12 + 12 * 2)),
        ("nnu_band", C.c_int), ("eedge", _dp),
        ("ncomp", C.c_int), ("multi_src", C.c_int),
        ("comp_kind", C.c_int * (MAXSRC + 1)),
        ("src_cdf", C.c_double * (MAXSRC + 1)),
        ("samp", _SamplerStruct * (MAXSRC + 1)),
        ("Ltot", C.c_double), ("qmc_ndim", C.c_int),
        ("nleaf", C.c_int), ("ncells", C.c_int), ("car", C.c_int),
        ("xmin", C.c_double), ("xmaxb", C.c_double), ("ymin", C.c_double), ("ymaxb", C.c_double),
        ("zmin", C.c_double), ("zmaxb", C.c_double), ("xrange", C.c_double),
        ("yrange", C.c_double), ("zrange", C.c_double), ("dcell", C.c_double),
        ("parent", _ip), ("children", _ip), ("level", _ip), ("ileaf", _ip),
        ("icell_of_leaf", _ip), ("neighbor", _ip),
        ("cx", _dp), ("cy", _dp), ("cz", _dp), ("ch", _dp),
        ("nH", _dp), ("xHI", _dp), ("xHeI", _dp), ("xHeII", _dp), ("ne", _dp), ("Te", _dp),
        ("rhokap", _dp),
        ("dn", C.c_int), ("dlam", _dp), ("dalb", _dp), ("dg", _dp), ("dcext", _dp),
        ("dcref", C.c_double),
        ("hhe", _dp), ("jt", _dp), ("dust_abs", _dp),
        ("seg_direct", C.c_int64), ("seg_scattered", C.c_int64),
        ("e_emit", C.c_double), ("e_abs", C.c_double), ("e_esc", C.c_double),
        ("gHI", _dp), ("gHeI", _dp), ("gHeII", _dp), ("hHI", _dp), ("hHeI", _dp), ("hHeII", _dp),
        ("cool", Cooling),
        ("milne", C.c_void_p), ("hei", C.c_void_p), ("dif_ch", _dp), ("dif_cdf", _dp),
        ("diffuse_lum", C.c_double), ("diffuse_nphot", C.c_int64),
        ("use_metals", C.c_int), ("nel", C.c_int), ("ntr", C.c_int), ("nst", C.c_int),
        ("el", C.c_byte * (11 * 3648)),
        ("tr_off", C.c_int * 11), ("st_off", C.c_int * 11),
        ("flat_ie", C.c_int * 64), ("flat_it", C.c_int * 64), ("flat_eth", C.c_double * 64),
        ("mraw", _dp), ("mgam", _dp), ("mheat", _dp), ("mfrac", _dp),
        ("ir", C.c_void_p),
    ]


class _IrHead(C.Structure):
    """Mirror of orc_dustir (mc_oracle.h)."""
    _fields_ = [
        ("tgrid", C.c_double * 61), ("femit", C.c_double * 61),
        ("temp_ready", C.c_int), ("setup_ready", C.c_int),
        ("nl", C.c_int), ("nzero", C.c_int),
        ("lam", _dp), ("sabs", _dp), ("edge", _dp),
        ("lam_hardest", C.c_double), ("frac_discarded", C.c_double),
        ("tdust", _dp), ("irabs", _dp), ("heat_ir", _dp), ("lesc", _dp),
        ("khi", C.POINTER(C.c_int64)),
        ("emitted", C.c_double), ("extinct", C.c_double), ("escaped", C.c_double),
        ("Ltot", C.c_double), ("Lpacket", C.c_double), ("xi", C.c_double),
        ("seg_direct", C.c_int64), ("seg_scattered", C.c_int64),
    ]


IR_REPORT_KEYS = ("passes", "converged", "L_emit", "emitted", "extinguished", "escaped",
                  "closure", "L_abs_path", "drel", "frac_discarded")


def ir_report(rep):
    return dict(zip(IR_REPORT_KEYS, [float(v) for v in rep]))


def sample_linear_bin(lam0, lam1, p0, p1, u):
    return lib().orc_sample_linear_bin(lam0, lam1, p0, p1, u)


def solve_ion_cell(gH, gHe1, gHe2, nH, T, caseA, yHe, xHI, xHeI, xHeII):
    a, b, c, ne = C.c_double(xHI), C.c_double(xHeI), C.c_double(xHeII), C.c_double()
    lib().orc_solve_ion_cell(gH, gHe1, gHe2, nH, T, int(caseA), yHe, C.byref(a), C.byref(b),
                             C.byref(c), C.byref(ne))
    return a.value, b.value, c.value, ne.value


def rng_uniform(seed, stream, pid, k):
    return lib().orc_rng_uniform(seed, stream, pid, k)


# ---- the reference's own modules compiled from source (optional) ----------
class Hei2ph(C.Structure):
    """Mirror of orc_hei2ph (the He I two-photon shape and its moments)."""
    _fields_ = [("ready", C.c_int), ("n_pts", C.c_int),
                ("tab_y", C.c_double * 128), ("tab_A", C.c_double * 128),
                ("n_per_decay", C.c_double), ("e_per_decay", C.c_double),
                ("y_lo", C.c_double), ("y_hi", C.c_double),
                ("grid_y", C.c_double * 1024), ("cdf_e", C.c_double * 1024)]


def hei_branch_fractions(T):
    """(f_3S, f_1S, f_1P) of the He I cascade at T (src/hei_cascade_mod.f90:72-81)."""
    f = np.zeros(3)
    lib().orc_hei_branch_fractions(T, ptr(f[0:]), ptr(f[1:]), ptr(f[2:]))
    return tuple(float(v) for v in f)


def hei_584_ionizes_H(T, xHI, xHeI):
    return float(lib().orc_hei_584_ionizes_H(T, xHI, xHeI))


def hei_two_photon(y, A, eion_max=100.0):
    """hei_2ph_setup of the given table: the Hei2ph record."""
    h = Hei2ph()
    y = np.ascontiguousarray(y, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    if lib().orc_hei2ph_setup(C.byref(h), y.size, ptr(y), ptr(A), eion_max):
        raise RuntimeError("orc_hei2ph_setup failed")
    return h


def hei2ph_sample_energy(h, u):
    return float(lib().orc_hei2ph_sample_energy(C.byref(h), u))


def hei_branch_energies(h, T, xHI, xHeI):
    e = np.zeros(3)
    lib().orc_hei_branch_energies(C.byref(h), T, xHI, xHeI, ptr(e[0:]), ptr(e[1:]), ptr(e[2:]))
    return tuple(float(v) for v in e)


def ref_lib():
    """The reference Fortran modules built into oracle/_ref, or None."""
    if not os.path.exists(REF_PATH):
        return None
    L = C.CDLL(REF_PATH)
    L.ref_qmc_setup.argtypes = [C.c_int]
    L.ref_qmc_uniforms.argtypes = [C.c_int64, C.c_int, C.c_int, _dp]
    L.ref_qmc_raw.argtypes = [C.c_int64, C.c_int, _dp]
    L.ref_build_planck.argtypes = [C.c_double] * 4 + [_dp, C.c_int]
    L.ref_build_planck.restype = C.c_int
    L.ref_build_tabulated.argtypes = [_dp, _dp, C.c_int]
    L.ref_build_tabulated.restype = C.c_int
    L.ref_sampler_get.argtypes = [_dp, _dp, _dp, _dp]
    L.ref_sample_energy.argtypes = [C.c_double]
    L.ref_sample_energy.restype = C.c_double
    L.ref_cdf_value.argtypes = [C.c_double]
    L.ref_cdf_value.restype = C.c_double
    L.ref_sigma.argtypes = [C.c_int, C.c_double]
    L.ref_sigma.restype = C.c_double
    L.ref_rate.argtypes = [C.c_int, C.c_double]
    L.ref_rate.restype = C.c_double
    L.ref_gaunt_ff.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float),
                               C.POINTER(C.c_float), C.c_int]
    L.ref_hei_branch_fractions.argtypes = [C.c_double, _dp, _dp, _dp]
    L.ref_hei_branch_fractions.restype = None
    L.ref_hei_584_ionizes_H.argtypes = [C.c_double] * 3
    L.ref_hei_584_ionizes_H.restype = C.c_double
    L.ref_milne_setup.argtypes = [C.c_double] * 3
    L.ref_milne_sample.argtypes = [C.c_int, C.c_double, C.c_double]
    L.ref_milne_sample.restype = C.c_double
    L.ref_milne_eper.argtypes = [C.c_int, C.c_double]
    L.ref_milne_eper.restype = C.c_double
    L.ref_vfky96.argtypes = [C.c_double, _dp]
    L.ref_vfky96.restype = C.c_double
    L.ref_vy95.argtypes = [C.c_double, _dp]
    L.ref_vy95.restype = C.c_double
    L.ref_rate_dere.argtypes = [C.c_int, C.c_double]
    L.ref_rate_dere.restype = C.c_double
    L.ref_ci_dere_ratio.argtypes = [C.c_int, C.c_int]
    L.ref_ci_dere_ratio.restype = C.c_double
    L.ref_nlevel_load.argtypes = [C.c_char_p, C.c_int]
    L.ref_nlevel_load.restype = C.c_int
    L.ref_nlevel_cooling_at.argtypes = [C.c_double, C.c_double]
    L.ref_nlevel_cooling_at.restype = C.c_double
    L.ref_nlevel_emissivities.argtypes = [C.c_double, C.c_double, _dp]
    L.ref_nlevel_line_ident.argtypes = [C.c_int]
    L.ref_nlevel_line_ident.restype = C.c_double
    L.ref_nlevel_cooling_add.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.c_int]
    L.ref_nlevel_cooling_add.restype = C.c_int
    L.ref_nlevel_cooling_apply.argtypes = [C.c_int] + [C.c_double] * 3
    L.ref_nlevel_cooling_apply.restype = C.c_double
    L.ref_nlevel_cooling_lam_se.argtypes = [C.c_int] + [C.c_double] * 2
    L.ref_nlevel_cooling_lam_se.restype = C.c_double
    return L


# ---- n-level atoms, local-n_e cooling tables, recombination lines ---------
NL_MAXLEV, NL_MAXTR, NLC_NT, NLC_NNE = 40, 600, 401, 161
ORC_MAX_EL_PY, ORC_MAX_ST_PY = 11, 6


class NlAtom(C.Structure):
    """Mirror of orc_nlatom."""
    _fields_ = [("nlev", C.c_int), ("g", C.c_double * NL_MAXLEV), ("e_cm1", C.c_double * NL_MAXLEV),
                ("nrad", C.c_int), ("rl", C.c_int * NL_MAXTR), ("ru", C.c_int * NL_MAXTR),
                ("A", C.c_double * NL_MAXTR), ("nups", C.c_int),
                ("ul_l", C.c_int * NL_MAXTR), ("ul_u", C.c_int * NL_MAXTR),
                ("ttype", C.c_int * NL_MAXTR), ("logf", C.c_int * NL_MAXTR),
                ("nc", C.c_int * NL_MAXTR), ("cups", C.c_double * NL_MAXTR),
                ("dery", C.c_double * NL_MAXTR), ("st_lo", C.c_double * NL_MAXTR),
                ("st_hi", C.c_double * NL_MAXTR), ("coef", (C.c_double * 16) * NL_MAXTR)]

    @classmethod
    def make(cls, atom):
        """From an atom dict (mochii_amd.atomic.read_nlevel)."""
        a = cls()
        a.nlev = len(atom["g"])
        for i, (g, e) in enumerate(zip(atom["g"], atom["e_cm1"])):
            a.g[i], a.e_cm1[i] = g, e
        a.nrad = len(atom["rad"])
        for k, (l, u, A) in enumerate(atom["rad"]):
            a.rl[k], a.ru[k], a.A[k] = l, u, A
        a.nups = len(atom["ups"])
        for k, r in enumerate(atom["ups"]):
            a.ul_l[k], a.ul_u[k], a.ttype[k], a.logf[k] = r["l"], r["u"], r["type"], r["logf"]
            a.nc[k], a.cups[k], a.dery[k] = len(r["coef"]), r["C"], r["dE"]
            a.st_lo[k], a.st_hi[k] = r["st_lo"], r["st_hi"]
            for j, v in enumerate(r["coef"]):
                a.coef[k][j] = v
        return a

    def cooling_at(self, T, ne):
        return lib().orc_nlevel_cooling_at(C.byref(self), T, ne)

    def emis(self, T, ne):
        out = np.zeros(NL_MAXTR)
        lib().orc_nlevel_emissivities(C.byref(self), T, ne, ptr(out))
        return out[:self.nrad].copy()

    def line_ident(self):
        return np.array([lib().orc_nlevel_line_ident(C.byref(self), k) for k in range(self.nrad)])

    def ups(self, k, T):
        return lib().orc_ups_eval(C.byref(self), k, T)


class Sh95(C.Structure):
    """Mirror of orc_sh95 (etab compact: [line][T][n_e])."""
    _fields_ = [("nT", C.c_int), ("nD", C.c_int), ("nlin", C.c_int),
                ("tgrid", C.c_double * 32), ("dgrid", C.c_double * 32), ("wls", C.c_double * 12),
                ("etab", C.c_double * (12 * 32 * 32))]

    @classmethod
    def make(cls, tab):
        """From a table dict (mochii_amd.atomic.read_line_table)."""
        t = cls()
        t.nT, t.nD, t.nlin = len(tab["T"]), len(tab["NE"]), len(tab["lines"])
        t.tgrid[:t.nT] = tab["T"]
        t.dgrid[:t.nD] = tab["NE"]
        for k, ln in enumerate(tab["lines"]):
            t.wls[k] = ln["wl"]
            e = np.asarray(ln["e"], dtype=np.float64).ravel()
            t.etab[k * t.nT * t.nD:(k + 1) * t.nT * t.nD] = e
        return t

    def emis(self, k, T, ne):
        return lib().orc_sh95_emis(C.byref(self), k, T, ne)


_NLC_CACHE = {}        # (atom key, atom dict id) -> table: built once per process


class NlCool(C.Structure):
    """Mirror of orc_nlcool: the (T, n_e) table of one ion."""
    _fields_ = [("ready", C.c_int), ("lamref", C.c_double * NLC_NT),
                ("Ssup", C.c_double * (NLC_NNE * NLC_NT))]

    @classmethod
    def build(cls, atom):
        t = cls()
        lib().orc_nlcool_build(C.byref(atom), C.byref(t))
        return t

    def coeff(self, T, ne):
        return lib().orc_nlcool_coeff(C.byref(self), T, ne)

    def apply(self, lam_fit, T, ne):
        return lib().orc_nlcool_apply(C.byref(self), lam_fit, T, ne)

    def low_density(self, T):
        return lib().orc_nlcool_low_density(C.byref(self), T)


class Milne:
    """Milne ground-continuum tables of the oracle (channels 0 H I, 1 He I,
    2 He II)."""
    SIZE = 8 * (192 + 512 + 3 * 192 * 512 + 2 * 3 * 192 + 3 + 1)

    def __init__(self, eion_max=100.0):
        self.buf = C.create_string_buffer(self.SIZE)
        lib().orc_milne_setup(self.buf, eion_max)

    def energy_per_recomb(self, ich, T):
        return lib().orc_milne_energy_per_recomb(self.buf, ich, T)

    def sample(self, ich, T, u):
        return lib().orc_milne_sample_energy(self.buf, ich, T, u)
