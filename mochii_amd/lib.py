"""ctypes binding of the C-ABI declared in include/mochii_gpu.h."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
MC_MAXSRC = 16
MC_FP64, MC_MIXED = 0, 1
MC_GRID_CAR, MC_GRID_AMR = 0, 1
MC_EXT_SPH, MC_EXT_REC = 1, 2
MC_DIFFUSE_THRESHOLD, MC_DIFFUSE_EXPONENTIAL, MC_DIFFUSE_MILNE = 0, 1, 2
MC_WARM_NONE, MC_WARM_NEWTON, MC_WARM_MLP = 0, 1, 2
MC_ORDER_LAUNCH, MC_ORDER_DIRECTION = 0, 1
MC_ROWS_AUTO, MC_ROWS_LANE, MC_ROWS_TOGETHER = 0, 1, 2
MC_MLP_NPARAM = 9180
MC_MAX_METAL_TR = 64
MC_CX_TWO_WAY, MC_CX_METAL_ONLY = 0, 1
MC_CI_VORONOV, MC_CI_DERE_HYBRID = 0, 1
MC_COOL_LOW_DENSITY, MC_COOL_LOCAL_NE = 0, 1
FIELDS = {"nH": 0, "xHI": 1, "xHeI": 2, "xHeII": 3, "ne": 4, "Te": 5, "rhokap": 6}
DUST_FIELDS = {"heat_dust": 0, "t_dust": 1, "heat_ir": 2, "ir_abs_path": 3}
MC_DUSTRT_NREPORT = 10
DUSTRT_KEYS = ("passes", "converged", "L_emit", "emitted", "extinguished", "escaped",
               "closure", "L_abs_path", "drel", "frac_discarded")


class MochiiError(RuntimeError):
    """A C-ABI call returned an error (the reference would abort the job)."""


class Params(C.Structure):
    """Mirror of mc_params (the subset of the reference's par on this path)."""
    _fields_ = [
        ("nphotons", C.c_int64), ("qmc_seed", C.c_int32), ("precision", C.c_int32),
        ("eion_min", C.c_double), ("eion_max", C.c_double), ("source_cdf_tol", C.c_double),
        ("nnu_ion", C.c_int32), ("ion_align_edges", C.c_int32),
        ("nsource", C.c_int32), ("ext_on", C.c_int32),
        ("tstar", C.c_double * MC_MAXSRC), ("src_lum", C.c_double * MC_MAXSRC),
        ("src_x", C.c_double * MC_MAXSRC), ("src_y", C.c_double * MC_MAXSRC),
        ("src_z", C.c_double * MC_MAXSRC),
        ("ext_geometry", C.c_int32), ("car_walk_amr", C.c_int32),
        ("ext_tstar", C.c_double), ("ext_intensity", C.c_double), ("rmax", C.c_double),
        ("grid_type", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
        ("distance2cm", C.c_double),
        ("He_abund", C.c_double),
        ("te_fixed", C.c_double), ("ion_relax", C.c_double), ("te_min", C.c_double),
        ("te_max", C.c_double), ("gas_tol", C.c_double), ("gas_tol_te", C.c_double),
        ("case_A", C.c_int32), ("solve_te", C.c_int32), ("conv_vol", C.c_int32),
        ("gas_niter", C.c_int32),
        ("xHI_init", C.c_double), ("xHeI_init", C.c_double), ("xHeII_init", C.c_double),
        ("ion_add_dust", C.c_int32), ("ion_dust_scatter", C.c_int32),
        ("lambda_ref", C.c_double),
        ("diffuse_field", C.c_int32), ("diffuse_model", C.c_int32),
        ("nmetal", C.c_int32), ("metal_eth", C.c_double * MC_MAX_METAL_TR),
        ("ion_metal_abs", C.c_int32), ("metal_ne", C.c_int32), ("metal_heat", C.c_int32),
        ("metal_freefree", C.c_int32), ("charge_exchange", C.c_int32), ("ci_model", C.c_int32),
        ("cooling_model", C.c_int32),
        ("dust_niter", C.c_int32), ("dust_no_photons", C.c_double), ("dust_tol", C.c_double),
        ("hei_diffuse", C.c_int32),
    ]

    # reference defaults (src/define.f90) for the fields this path reads
    DEFAULTS = dict(nphotons=100000, qmc_seed=12345, precision=MC_FP64, eion_min=13.598,
                    eion_max=100.0, source_cdf_tol=1e-8, nnu_ion=16, ion_align_edges=1,
                    nsource=1, ext_on=0, ext_geometry=MC_EXT_REC, car_walk_amr=0,
                    ext_tstar=-999.0, ext_intensity=-1.0, rmax=-999.0, grid_type=MC_GRID_CAR,
                    nx=1, ny=1, nz=11, xmax=1.0, ymax=1.0, zmax=1.0,
                    distance2cm=3.0856776e21, He_abund=0.1, te_fixed=1e4, ion_relax=1.0,
                    te_min=1e3, te_max=5e4, gas_tol=1e-3, gas_tol_te=1e-3, case_A=0,
                    solve_te=0, conv_vol=0, gas_niter=0, xHI_init=1.0, xHeI_init=1.0,
                    xHeII_init=0.0, ion_add_dust=0, ion_dust_scatter=0, lambda_ref=0.55,
                    diffuse_field=0, diffuse_model=MC_DIFFUSE_MILNE, nmetal=0, ion_metal_abs=1,
                    metal_ne=1, metal_heat=1, metal_freefree=1, charge_exchange=MC_CX_TWO_WAY,
                    ci_model=MC_CI_VORONOV, cooling_model=MC_COOL_LOW_DENSITY,
                    dust_niter=2, dust_no_photons=1.0e6, dust_tol=1.0e-3, hei_diffuse=0)

    @classmethod
    def make(cls, **kw):
        p = cls()
        vals = dict(cls.DEFAULTS)
        vals.update(kw)
        for k, v in vals.items():
            if k == "metal_eth":
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
                p.nmetal = len(v)
            elif k in ("tstar", "src_lum", "src_x", "src_y", "src_z"):
                arr = getattr(p, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(p, k, v)
        return p


def library_path():
    return os.environ.get("MOCHII_GPU_LIB", os.path.join(HERE, "libmochii_gpu.so"))


_LIB = None


def library():
    """Load libmochii_gpu.so (raises MochiiError if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise MochiiError(f"HIP library not built: {path} (run make -C mochii_amd)")
    L = C.CDLL(path)
    vp, dp, ip, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sig = {
        "mc_last_error": ([], C.c_char_p),
        "mc_init": ([C.POINTER(Params), C.POINTER(vp)], C.c_int),
        "mc_finalize": ([vp], C.c_int),
        "mc_grid_octree": ([vp, C.c_int, dp, dp, dp, ip, C.c_double], C.c_int),
        "mc_gas_state_setup": ([vp, dp, dp], C.c_int),
        "mc_state_set": ([vp, C.c_int, dp], C.c_int),
        "mc_state_get": ([vp, C.c_int, dp], C.c_int),
        "mc_ion_dust_table": ([vp, C.c_int, dp, dp, dp, dp], C.c_int),
        "mc_gas_opacity_fill": ([vp], C.c_int),
        "mc_ionizing_field_and_rates": ([vp, C.c_int], C.c_int),
        "mc_gas_equilibrium_update": ([vp, dp, dp], C.c_int),
        "mc_hydrogen_balance_residual_grid": ([vp, dp, dp], C.c_int),
        "mc_diffuse_info": ([vp, dp, i64p], C.c_int),
        "mc_diffuse_hei_setup": ([vp, C.c_int, dp, dp], C.c_int),
        "mc_diffuse_hei_info": ([vp, dp, dp, dp, dp], C.c_int),
        "mc_set_warm_start": ([vp, C.c_int], C.c_int),
        "mc_set_warm_levels": ([vp, C.c_int], C.c_int),
        "mc_nlevel_setup": ([vp, C.c_int, vp], C.c_int),
        "mc_lines_setup": ([vp, vp, vp, vp], C.c_int),
        "mc_lines_compute": ([vp, C.c_int, dp, dp, ip, ip], C.c_int),
        "mc_qmc_uniforms": ([vp, C.c_int, C.c_int64, C.c_int, C.c_int, dp], C.c_int),
        "mc_launch_records": ([vp, C.c_int64, C.c_int, dp, ip], C.c_int),
        "mc_set_packet_order": ([vp, C.c_int], C.c_int),
        "mc_set_tally_moments": ([vp, C.c_int], C.c_int),
        "mc_set_launch_batch": ([vp, C.c_int64], C.c_int),
        "mc_set_row_issue": ([vp, C.c_int], C.c_int),
        "mc_row_issue": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "mc_tally_moment_info": ([vp, C.POINTER(C.c_int32), dp], C.c_int),
        "mc_thermal_stats": ([vp, i64p], C.c_int),
        "mc_mlp_set": ([vp, fp, C.c_int], C.c_int),
        "mc_mlp_get": ([vp, fp], C.c_int),
        "mc_mlp_features": ([vp, fp], C.c_int),
        "mc_mlp_train": ([vp, fp, fp, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint32, fp],
                         C.c_int),
        "mc_mlp_grad": ([vp, fp, fp, C.c_int, fp, fp], C.c_int),
        "mc_mlp_infer": ([vp, fp, C.c_int, fp], C.c_int),
        "mc_cooling_setup": ([vp, C.c_int, dp, dp], C.c_int),
        "mc_gas_thermal_update": ([vp, dp, dp, dp, dp], C.c_int),
        "mc_tallies_get": ([vp, dp, dp, dp, i64p], C.c_int),
        "mc_rates_get": ([vp, dp], C.c_int),
        "mc_energy_ledger": ([vp, dp], C.c_int),
        "mc_info": ([vp, ip, ip, dp, dp], C.c_int),
        "mc_set_precision": ([vp, C.c_int], C.c_int),
        "mc_timing": ([vp, dp], C.c_int),
        "mc_trace_ray": ([vp, C.c_int, dp, dp, ip, C.c_int, ip], C.c_int),
        "mc_trace_ray_noalloc": ([vp, C.c_int, dp, dp, ip, C.c_int, ip], C.c_int),
        "mc_probe_scattered_rmw": ([C.c_int64, C.c_int, C.c_int64, dp, dp], C.c_int),
        "mc_probe_row_atomics": ([C.c_int64, C.c_int, C.c_int64, dp, dp], C.c_int),
        "mc_leaf_list_read": ([C.c_char_p, C.c_int64, i64p, dp, dp, dp, dp, ip, dp, dp, dp, dp,
                               ip], C.c_int),
        "mc_species_setup": ([vp, C.c_int, vp], C.c_int),
        "mc_species_fractions_get": ([vp, C.c_int, dp], C.c_int),
        "mc_species_rates_get": ([vp, C.c_int, dp, dp], C.c_int),
        "mc_species_tallies_get": ([vp, C.c_int, dp], C.c_int),
        "mc_metal_segment_transitions": ([vp, i64p], C.c_int),
        "mc_walk_atomics": ([vp, i64p], C.c_int),
        "mc_dust_temp_setup": ([vp], C.c_int),
        "mc_dust_temp_compute": ([vp], C.c_int),
        "mc_dust_emis_transport_setup": ([vp], C.c_int),
        "mc_transport_dust_emission": ([vp, dp], C.c_int),
        "mc_dust_get": ([vp, C.c_int, dp], C.c_int),
        "mc_dust_set": ([vp, C.c_int, dp], C.c_int),
        "mc_dust_ir_energy_budget": ([vp, dp], C.c_int),
        "mc_dust_ir_escaped_spectrum": ([vp, dp, dp, dp], C.c_int),
        "mc_dust_ir_grid": ([vp, ip, dp, dp, dp, ip, dp, dp], C.c_int),
        "mc_dust_ir_timing": ([vp, dp, i64p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = library().mc_last_error().decode(errors="replace")
        raise MochiiError(f"{what} failed (rc={rc}): {msg}")
