"""Test helpers: build the same model on the oracle and on the HIP library."""
import numpy as np

from oracle import pyoracle as po

_MAP = {  # library Params name -> oracle param name
    "nphotons": "nphotons", "qmc_seed": "qmc_seed", "eion_min": "eion_min",
    "eion_max": "eion_max", "source_cdf_tol": "source_cdf_tol", "nnu_ion": "nnu_ion",
    "ion_align_edges": "align_edges", "nsource": "nsource", "tstar": "tstar",
    "src_lum": "lum", "src_x": "sx", "src_y": "sy", "src_z": "sz", "ext_on": "ext_on",
    "ext_geometry": "ext_geometry", "ext_tstar": "ext_tstar", "ext_intensity": "ext_intensity",
    "rmax": "rmax", "grid_type": "grid_type", "car_walk_amr": "car_walk_amr",
    "nx": "nx", "ny": "ny", "nz": "nz", "xmax": "xmax", "ymax": "ymax", "zmax": "zmax",
    "distance2cm": "distance2cm", "He_abund": "He_abund", "te_fixed": "te_fixed",
    "ion_relax": "ion_relax", "te_min": "te_min", "te_max": "te_max", "case_A": "caseA",
    "ion_add_dust": "add_dust", "ion_dust_scatter": "dust_scatter", "lambda_ref": "lambda_ref",
    "xHI_init": "xHI_init", "xHeI_init": "xHeI_init", "xHeII_init": "xHeII_init",
    "diffuse_field": "diffuse_field", "diffuse_model": "diffuse_model",
    "metal_eth": "metal_eth", "ion_metal_abs": "ion_metal_abs", "metal_ne": "metal_ne",
    "metal_heat": "metal_heat", "metal_freefree": "metal_freefree",
    "hei_diffuse": "hei_diffuse",
}
_ORC_DEFAULTS = dict(source_cdf_tol=1e-8, align_edges=1, lambda_ref=0.55, te_min=1e3,
                     te_max=5e4, ext_geometry=2)


def oracle_params(params):
    out = dict(_ORC_DEFAULTS)
    for k, v in params.items():
        if k in _MAP:
            out[_MAP[k]] = v
    if out.get("diffuse_field"):
        out["caseA"] = 1                   # the diffuse field forces case A
        out.setdefault("diffuse_model", 2)
    if "charge_exchange" in params:
        out["cx_two_way"] = int(params["charge_exchange"] == 0)
    if "ci_model" in params:
        out["ci_dere"] = int(params["ci_model"] == 1)          # MC_CI_DERE_HYBRID
    return out


def oracle_model(params, nH, rhokap=None, dust_table=None, octree=None, elements=None):
    from mochii_amd.workloads import elements_of
    op = oracle_params(params)
    els = elements if elements is not None else elements_of(params)
    m = po.Model({k: v for k, v in op.items()
                  if k not in ("xHI_init", "xHeI_init", "xHeII_init")} | {}, octree=octree,
                 dust_table=dust_table, elements=els)
    m.params.update(op)
    m.state_init(nH, rhokap=rhokap)
    if op.get("diffuse_field") and op.get("hei_diffuse"):
        from mochii_amd.atomic import load_hei_two_photon
        m.diffuse_hei_setup(*load_hei_two_photon())
    if params.get("cooling_model") == 1:          # 'local_ne': the (T, n_e) tables
        from mochii_amd.atomic import load_nlevel
        m.nlevel_setup(load_nlevel())
    return m


def gpu_model(params, nH, rhokap=None, dust_table=None, octree=None, elements=None):
    import mochii_amd as mc
    from mochii_amd.workloads import elements_of
    m = mc.Model(dict(params))
    if octree is not None:
        m.grid_octree(*octree)
    if dust_table is not None:
        m.ion_dust_table(*dust_table)
    els = elements if elements is not None else elements_of(params)
    if els:
        m.species_setup(els)
    m.gas_state_setup(nH)
    if rhokap is not None:
        m.set_state(rhokap=rhokap)
    if params.get("cooling_model") == 1:          # 'local_ne': the n-level atoms
        m.nlevel_setup()
    return m


def rel_close(a, b, rtol, floor_frac=1e-6):
    """Elementwise |a-b| <= rtol*|b| on entries above floor_frac*max|b|;
    returns (ok, worst relative error)."""
    a, b = np.asarray(a), np.asarray(b)
    big = np.abs(b) > floor_frac * np.abs(b).max() if b.size and np.abs(b).max() > 0 else np.zeros_like(b, bool)
    if not big.any():
        return np.allclose(a, b), 0.0
    err = np.abs(a[big] - b[big]) / np.abs(b[big])
    return bool(err.max() <= rtol), float(err.max())


def read_fits_images(path):
    """Every HDU of a FITS file of images: [(header dict, numpy array)],
    header values parsed (T/F, ints, floats, quoted strings), the array in C
    order (NAXISn reversed).  Written here independently of the package."""
    import gzip as _gz
    b = _gz.open(path).read() if str(path).endswith(".gz") else open(path, "rb").read()
    out, p = [], 0
    while p < len(b):
        h, order = {}, []
        while True:
            c = b[p:p + 80].decode("ascii")
            p += 80
            if c.startswith("END"):
                break
            if c[8:10] != "= ":
                continue
            k, v = c[:8].strip(), c[10:]
            if v.strip().startswith("'"):
                val = v.strip()[1:].split("'")[0].rstrip()
            else:
                t = v.split("/")[0].strip()
                if t in ("T", "F"):
                    val = t == "T"
                else:
                    try:
                        val = int(t)
                    except ValueError:
                        val = float(t)
            h[k] = val
            order.append(k)
        p = -(-p // 2880) * 2880
        shape = tuple(h[f"NAXIS{i}"] for i in range(h["NAXIS"], 0, -1))
        n = int(np.prod(shape)) if shape else 0
        assert h["BITPIX"] == -64
        a = np.frombuffer(b, dtype=">f8", count=n, offset=p).reshape(shape).astype(np.float64)
        p += -(-(8 * n) // 2880) * 2880
        h["_order"] = order
        out.append((h, a))
    return out
