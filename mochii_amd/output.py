"""The rates file of a converged model, as gas_rates_write writes it
(src/gas_rates_mod.f90:216-366, with species_write,
src/species_mod.f90:873-915): a gzip-compressed FITS file of BITPIX -64
image HDUs in the reference's order -- Gamma_HI (primary, with the run's
keywords), Gamma_HeI, Gamma_HeII, Heat_HI, Heat_HeI, Heat_HeII, Heat_dust and
T_dust (grains), J_nu(nnu_band, leaf), x_HI, x_HeI, x_HeII, n_e, T_e, E_bin,
dE_bin, LeafXYZ(leaf, 3), LeafSize, then per registry element
x_<el>_stages(stage, leaf), Gamma_<el>_stages and Heat_<el>_stages.  Arrays
keep the Fortran layout: the first index runs fastest (NAXIS1).

rates_fields(model, ...) gathers the arrays from a device model;
write_rates(path, hdus) writes them (no device needed)."""
import gzip
import os

import numpy as np

FOURPI = 12.56637061435917295385057353311801153679
EV2ERG = 1.602176634e-12
H_PLANCK = 6.62607015e-27


def rates_filename(out_file, file_format="fits"):
    """get_base_name(out_file)//'_rates'//io_file_extension(file_format)
    (src/utility.f90:222-250, src/iofile_mod.f90:121-135); HDF5 output is
    not written by this package."""
    if file_format.strip().lower() in ("hdf5", "h5"):
        raise ValueError("the rates file is written as FITS only (file_format 'fits')")
    base = os.path.basename(out_file)
    for ext in (".fits.gz", ".hdf5", ".fits", ".h5"):
        if base.endswith(ext):
            base = base[:-len(ext)]
            break
    return base + "_rates.fits.gz"


def leaf_geometry(model):
    """Leaf centres (nleaf, 3) and full widths of the model's grid: the leaf
    list it was built from, or the namelist car grid in raster order
    (src/grid_mod_amr.f90:95-107)."""
    leaves = getattr(model, "leaves", None)
    if leaves is not None:
        x, y, z, level, half = leaves
        return np.stack([x, y, z], axis=1), 2.0 * half / 2.0 ** np.asarray(level, np.float64)
    p = model.params
    dc = 2.0 * p.xmax / p.nx
    ix = np.arange(p.nx * p.ny * p.nz)
    cx = -p.xmax + (ix % p.nx + 0.5) * dc
    cy = -p.ymax + ((ix // p.nx) % p.ny + 0.5) * (2.0 * p.ymax / p.ny)
    cz = -p.zmax + (ix // (p.nx * p.ny) + 0.5) * (2.0 * p.zmax / p.nz)
    return np.stack([cx, cy, cz], axis=1), np.full(ix.size, dc)


def rates_fields(model, run=None, luminosity=None, no_photons=None, launch_sequence="sobol"):
    """The HDUs and primary keywords of the rates file of a device model
    after its last pass.  run: run_model's result (CONVERGD, NITERDON,
    FINALDX, FINALDTE); luminosity: par%luminosity (default the band total,
    as ion_setup fills an unset one).  The metal stage fractions are those
    of the stage refresh at the written state and the last pass's rates
    (species_write evaluates species_fractions there): the call refreshes
    the stage cache (gas_opacity_fill) first."""
    p = model.params
    rates = model.rates()
    state = model.state("nH", "xHI", "xHeI", "xHeII", "ne", "Te")
    tal = model.tallies()
    edges = model.band_edges()
    e_bin = np.sqrt(edges[:-1] * edges[1:])
    de_bin = edges[1:] - edges[:-1]
    dnu = de_bin * EV2ERG / H_PLANCK
    xyz, size = leaf_geometry(model)
    jnu = tal["jt"] / (FOURPI * size[:, None] ** 3 * dnu[None, :] * p.distance2cm ** 2)
    led = model.energy_ledger()
    hist = (run or {}).get("history") or []
    last = hist[-1] if hist else None
    head = [("EXTNAME", "Gamma_HI", "H I photoionization rate [s^-1]"),
            ("NNU_ION", int(p.nnu_ion), "ionizing energy-bin count"),
            ("NNU_BAND", int(model.nnu_band), "total band bins (ionizing+FUV)"),
            ("EION_MIN", float(p.eion_min), "band lower edge [eV]"),
            ("EION_MAX", float(p.eion_max), "band upper edge [eV]"),
            ("TOT_LUM", float(luminosity if luminosity and luminosity > 0 else model.ion_Ltot),
             "band luminosity [erg/s]"),
            ("DIST_CM", float(p.distance2cm), "distance unit (cm)"),
            ("NPHOTONS", float(no_photons if no_photons is not None else p.nphotons),
             "number of photons"),
            ("CONVERGD", bool((run or {}).get("converged", False)), "gas iteration converged"),
            ("NITERDON", int((run or {}).get("iterations", 0)), "gas iterations run"),
            ("FINALDX", float(last[1]) if last else 0.0, "final max|delta x_HII|"),
            ("FINALDTE", float(last[3]) if last and len(last) > 3 else 0.0,
             "final max|delta Te|/Te"),
            ("FLDCONS", True, "rates rebuilt by transporting the written state"),
            ("IONEMODE", "continuous", "ionizing packet energy sampling"),
            ("CDFRTOL", float(p.source_cdf_tol), "source sampler relative tolerance"),
            ("NUBINUSE", "diagnostic_only", "role of ionizing spectral bins"),
            ("ENRGSAMP", launch_sequence, "energy/source sampling sequence"),
            ("L_EMIT", float(led[0]), "transported ionizing packet luminosity"),
            ("L_ABS", float(led[1]), "absorbed ionizing packet luminosity"),
            ("L_ESC", float(led[2]), "escaped ionizing packet luminosity")]
    if launch_sequence == "sobol":
        head.append(("QMCSEED", int(p.qmc_seed), "Sobol scramble seed"))
    hdus = [(rates["gHI"], head)]

    def add(a, name, comment, extra=()):
        hdus.append((np.asarray(a, np.float64), [("EXTNAME", name, comment), *extra]))

    add(rates["gHeI"], "Gamma_HeI", "He I photoionization rate [s^-1]")
    add(rates["gHeII"], "Gamma_HeII", "He II photoionization rate [s^-1]")
    add(rates["hHI"], "Heat_HI", "H I photoheating [erg/s per HI]")
    add(rates["hHeI"], "Heat_HeI", "He I photoheating [erg/s per HeI]")
    add(rates["hHeII"], "Heat_HeII", "He II photoheating [erg/s per HeII]")
    if p.ion_add_dust:
        add(model.dust("heat_dust"), "Heat_dust", "grain heating [erg/s/cm^3]")
        if getattr(model, "dust_temp_ready", False):
            add(model.dust("t_dust"), "T_dust", "equilibrium mixture dust temperature [K]")
    # (nnu_band, nleaf) in Fortran order = C array (nleaf, nnu_band)
    add(jnu, "J_nu", "J(nu,leaf) [erg/s/cm^2/Hz/sr]")
    add(state["xHI"], "x_HI", "n_HI/n_H per leaf")
    add(state["xHeI"], "x_HeI", "n_HeI/n_He per leaf")
    add(state["xHeII"], "x_HeII", "n_HeII/n_He per leaf")
    add(state["ne"], "n_e", "electron density [cm^-3]")
    add(state["Te"], "T_e", "electron temperature [K]")
    add(e_bin, "E_bin", "bin centers [eV]")
    add(de_bin, "dE_bin", "bin widths [eV]")
    # leafxyz(nleaf, 3) in Fortran order = C array (3, nleaf)
    add(xyz.T, "LeafXYZ", "leaf center x,y,z (code units)")
    add(size, "LeafSize", "leaf full width (code units)")
    els = getattr(model, "elements", None) or []
    if els:
        model.gas_opacity_fill()
        empty = state["nH"] <= 0.0
        for ie, el in enumerate(els):
            fr = model.species_fractions(ie)
            fr[empty] = 0.0                 # an empty leaf is written all neutral
            fr[empty, 0] = 1.0
            add(fr, f"x_{el['name']}_stages", "ion fractions (stage, leaf)",
                [("ABUND", float(el["abund"]), "n(X)/n(H)")])
            g, h = model.species_rates(ie)
            add(g, f"Gamma_{el['name']}_stages",
                "photoionization rates (transition, leaf) [s^-1]")
            add(h, f"Heat_{el['name']}_stages",
                "photoheating rates (transition, leaf) [erg/s per ion]")
    return hdus


# ---- FITS writer -----------------------------------------------------------
def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        v = "%20d" % value
    elif isinstance(value, (float, np.floating)):
        v = "%20s" % repr(float(value)).upper()       # shortest round-trip digits
    else:
        v = "%-20s" % ("'%-8s'" % str(value).replace("'", "''"))
    c = "%-8s= %s" % (key[:8].upper(), v)
    if comment:
        c += " / " + comment
    return c[:80].ljust(80)


def _hdu(a, cards, primary):
    a = np.asarray(a, dtype=np.float64)
    # numpy C order (slow, ..., fast) -> FITS NAXIS1 = fastest
    shape = a.shape[::-1]
    head = ([_card("SIMPLE", True, "conforms to FITS standard")] if primary else
            [_card("XTENSION", "IMAGE", "image extension")])
    head += [_card("BITPIX", -64, "array data type"), _card("NAXIS", len(shape))]
    head += [_card(f"NAXIS{i + 1}", int(n)) for i, n in enumerate(shape)]
    head += ([_card("EXTEND", True)] if primary else
             [_card("PCOUNT", 0), _card("GCOUNT", 1)])
    head += [_card(k, v, c) for k, v, c in cards]
    text = "".join(head) + "END".ljust(80)
    text += " " * (-len(text) % 2880)
    data = np.ascontiguousarray(a).astype(">f8").tobytes()
    data += b"\0" * (-len(data) % 2880)
    return text.encode("ascii") + data


def write_rates(path, hdus):
    """Write [(array, [(key, value, comment), ...]), ...] as one FITS file,
    gzip-compressed when path ends with .gz."""
    blob = b"".join(_hdu(a, c, i == 0) for i, (a, c) in enumerate(hdus))
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(blob)
    return path


def gas_rates_write(model, out_dir=".", out_file="mochii", file_format="fits", **kw):
    """gas_rates_write: the rates file of the model in out_dir; kw as for
    rates_fields.  Returns the path written."""
    path = os.path.join(out_dir, rates_filename(out_file, file_format))
    return write_rates(path, rates_fields(model, **kw))
