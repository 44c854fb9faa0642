"""Trace-metal registry on the host: element data, the C struct, thresholds.

Mirrors species_setup / read_element (src/species_mod.f90:143-228, 358-422)
and cooling_load (src/cooling_mod.f90:98-123).  The element and cooling files
use the reference's own text formats (data/atomic/element_<el>.txt: keyword
rows ELEMENT, TRANSITION, ETH, PHOTO/PHOTO2, CI, RR/RR2, DR/DR2, CXI, CXR;
cooling_<el>_<i>.txt: the term count then A_i T_i rows), so a driver can read
a data directory in that format directly with ``read_registry``.  A snapshot
of the eleven registry elements is kept as data in
``mochii_amd/data/atomic_registry.json`` (written by
``mochii_amd/data/make_atomic_json.py``) and is what ``registry`` loads.
"""
import ctypes as C
import json
import os
import numpy as np

MC_MAX_EL, MC_MAX_ST, MC_MAX_METAL_TR, MC_MAX_COOLFIT = 11, 6, 64, 8
MC_CX_TWO_WAY, MC_CX_METAL_ONLY = 0, 1
MC_CI_VORONOV, MC_CI_DERE_HYBRID = 0, 1
MC_COOL_LOW_DENSITY, MC_COOL_LOCAL_NE = 0, 1

# species_setup's element order and the par%abund_* defaults
# (src/species_mod.f90:153-157, src/define.f90:578-580, 847-857)
ORDER = ("c", "n", "o", "ne", "s", "ar", "mg", "fe", "si", "cl", "ca")
DEFAULT_ABUND = {"c": 2.2e-4, "n": 4.0e-5, "o": 3.3e-4, "ne": 5.0e-5, "s": 9.0e-6,
                 "ar": 0.0, "mg": 0.0, "fe": 0.0, "si": 0.0, "cl": 0.0, "ca": 0.0}

HERE = os.path.dirname(os.path.abspath(__file__))
SNAPSHOT = os.path.join(HERE, "data", "atomic_registry.json")

_ST = MC_MAX_ST


def _d(*dims):
    t = C.c_double
    for n in reversed(dims):
        t = t * n
    return t


class Element(C.Structure):
    """Mirror of mc_element (include/mochii_gpu.h)."""
    _fields_ = [
        ("name", C.c_char * 4), ("Z", C.c_int32), ("nstage", C.c_int32),
        ("abund", C.c_double),
        ("eth", _d(_ST)), ("photo_form", C.c_int32 * _ST), ("photo", _d(_ST, 8)),
        ("ci", _d(_ST, 5)), ("rr", _d(_ST, 6)),
        ("ndr", C.c_int32 * _ST), ("dr_c", _d(_ST, 9)), ("dr_e", _d(_ST, 9)),
        ("has_dr2", C.c_int32 * _ST), ("dr2", _d(_ST, 4)),
        ("has_rr2", C.c_int32 * _ST), ("rr2", _d(_ST, 2)),
        ("cxi_form", C.c_int32 * _ST), ("cxi", _d(_ST, 6)),
        ("cxr_form", C.c_int32 * _ST), ("cxr", _d(_ST, 6)),
        ("ncool", C.c_int32 * _ST),
        ("cool_A", _d(_ST, MC_MAX_COOLFIT)), ("cool_T", _d(_ST, MC_MAX_COOLFIT)),
    ]


def _floats(tok):
    return [float(t.replace("D", "E").replace("d", "e")) for t in tok]


def read_element(path):
    """Parse one element file (read_element, src/species_mod.f90:358-422) into
    a dict: name, Z, nstage, transitions (list of per-transition dicts)."""
    el = {"name": "", "Z": 0, "nstage": 0, "transitions": []}
    cur = None
    toks = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            toks.append(s.split())
    i = 0
    while i < len(toks):
        t = toks[i]
        key = t[0]
        if key == "ELEMENT":
            el["name"], el["Z"], el["nstage"] = t[1].lower(), int(t[2]), int(t[3])
        elif key == "TRANSITION":
            it = int(t[1])
            while len(el["transitions"]) < it:
                el["transitions"].append(dict(eth=0.0, photo_form=1, photo=[0.0] * 8,
                                              ci=[0.0] * 5, rr=[0.0] * 6, ndr=0, dr_c=[],
                                              dr_e=[], dr2=None, rr2=None, cxi_form=0,
                                              cxi=[0.0] * 6, cxr_form=0, cxr=[0.0] * 6))
            cur = el["transitions"][it - 1]
        elif key == "ETH":
            cur["eth"] = _floats(t[1:2])[0]
        elif key == "PHOTO":
            cur["photo"], cur["photo_form"] = _floats(t[1:9]), 1
        elif key == "PHOTO2":
            cur["photo"], cur["photo_form"] = _floats(t[1:8]) + [0.0], 2
        elif key == "CI":
            cur["ci"] = _floats(t[1:6])
        elif key == "RR":
            cur["rr"] = _floats(t[1:7])
        elif key == "DR":
            n = int(t[1])
            vals = _floats(t[2:])
            # list-directed read continues over following lines when needed
            while len(vals) < 2 * n:
                i += 1
                vals += _floats(toks[i])
            cur["ndr"], cur["dr_c"], cur["dr_e"] = n, vals[:n], vals[n:2 * n]
        elif key == "DR2":
            cur["dr2"] = _floats(t[1:5])
        elif key == "RR2":
            cur["rr2"] = _floats(t[1:3])
        elif key in ("CXI", "CXR"):
            form = int(float(t[1]))
            vals = _floats(t[2:8]) if form > 0 else [0.0] * 6
            vals += [0.0] * (6 - len(vals))
            if key == "CXI":
                cur["cxi_form"], cur["cxi"] = form, vals
            else:
                cur["cxr_form"], cur["cxr"] = form, vals
        i += 1
    return el


def read_cooling_fit(path):
    """cooling_load (src/cooling_mod.f90:98-123): [(A_i, T_i), ...]."""
    with open(path) as f:
        rows = [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    n = int(rows[0][0])
    return [tuple(_floats(r[:2])) for r in rows[1:1 + n]]


def read_nlevel(path):
    """nlevel_load (src/nlevel_mod.f90:61-108): one n-level atom file
    (data/atomic/nlevel_<el>_<stage>.txt) -> dict with the level list
    (g, E [cm^-1]), the radiative rows (l, u, A [s^-1]) and the collision
    strength fits (l, u, type, C, dE [Ry], st_lo, st_hi, logf, Chebyshev
    coefficients).  Level indices stay 1-based as in the file."""
    with open(path) as f:
        rows = [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    at = {"g": [], "e_cm1": [], "rad": [], "ups": []}
    i = 0
    while i < len(rows):
        key, n = rows[i][0], int(rows[i][1])
        body = rows[i + 1:i + 1 + n]
        if key == "NLEV":
            at["g"] = [float(r[1]) for r in body]
            at["e_cm1"] = [float(r[2]) for r in body]
        elif key == "NRAD":
            at["rad"] = [[int(r[0]), int(r[1]), float(r[2])] for r in body]
        elif key == "NUPS":
            for r in body:
                nc = int(r[8])
                at["ups"].append({"l": int(r[0]), "u": int(r[1]), "type": int(r[2]),
                                  "C": float(r[3]), "dE": float(r[4]), "st_lo": float(r[5]),
                                  "st_hi": float(r[6]), "logf": int(r[7]),
                                  "coef": [float(v) for v in r[9:9 + nc]]})
        i += 1 + n
    return at


MC_NL_MAXLEV, MC_NL_MAXTR = 40, 600


class NlevelAtom(C.Structure):
    """Mirror of mc_nlevel_atom (include/mochii_gpu.h)."""
    _fields_ = [("name", C.c_char * 4), ("stage", C.c_int32), ("nlev", C.c_int32),
                ("g", C.c_double * MC_NL_MAXLEV), ("e_cm1", C.c_double * MC_NL_MAXLEV),
                ("nrad", C.c_int32), ("rl", C.c_int32 * MC_NL_MAXTR),
                ("ru", C.c_int32 * MC_NL_MAXTR), ("A", C.c_double * MC_NL_MAXTR),
                ("nups", C.c_int32), ("ul_l", C.c_int32 * MC_NL_MAXTR),
                ("ul_u", C.c_int32 * MC_NL_MAXTR), ("ttype", C.c_int32 * MC_NL_MAXTR),
                ("logf", C.c_int32 * MC_NL_MAXTR), ("nc", C.c_int32 * MC_NL_MAXTR),
                ("cups", C.c_double * MC_NL_MAXTR), ("dery", C.c_double * MC_NL_MAXTR),
                ("st_lo", C.c_double * MC_NL_MAXTR), ("st_hi", C.c_double * MC_NL_MAXTR),
                ("coef", (C.c_double * 16) * MC_NL_MAXTR)]


def nlevel_struct(atoms):
    """ctypes array of NlevelAtom for mc_nlevel_setup from {'<el>_<stage>':
    atom dict} (read_nlevel / load_nlevel)."""
    keys = sorted(atoms)
    arr = (NlevelAtom * max(len(keys), 1))()
    for n, key in enumerate(keys):
        name, stage = key.rsplit("_", 1)
        at, a = atoms[key], arr[n]
        a.name, a.stage, a.nlev = name.encode()[:3], int(stage), len(at["g"])
        for i, (g, e) in enumerate(zip(at["g"], at["e_cm1"])):
            a.g[i], a.e_cm1[i] = g, e
        a.nrad = len(at["rad"])
        for k, (l, u, A) in enumerate(at["rad"]):
            a.rl[k], a.ru[k], a.A[k] = l, u, A
        a.nups = len(at["ups"])
        for k, r in enumerate(at["ups"]):
            a.ul_l[k], a.ul_u[k], a.ttype[k], a.logf[k] = r["l"], r["u"], r["type"], r["logf"]
            a.nc[k], a.cups[k], a.dery[k] = len(r["coef"]), r["C"], r["dE"]
            a.st_lo[k], a.st_hi[k] = r["st_lo"], r["st_hi"]
            for j, v in enumerate(r["coef"]):
                a.coef[k][j] = v
    return arr, len(keys)


def nlevel_files(atomic_dir):
    """The n-level atoms of a data directory: {'<el>_<stage>': atom}."""
    out = {}
    for fn in sorted(os.listdir(atomic_dir)):
        if fn.startswith("nlevel_") and fn.endswith(".txt") and not fn.endswith("_full.txt"):
            out[fn[len("nlevel_"):-len(".txt")]] = read_nlevel(os.path.join(atomic_dir, fn))
    return out


def read_line_table(path):
    """sh95_load (src/sh95_mod.f90:93-136): a recombination-line table
    (GRID nT nD, T row, NE row, then LINE label nu nl lambda[A] blocks of
    nT rows of nD emissivities 4 pi j / (n_e n_ion) [erg cm^3 s^-1])."""
    with open(path) as f:
        rows = [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("#")]
    t = {"T": [], "NE": [], "lines": []}
    i = 0
    while i < len(rows):
        r = rows[i]
        if r[0] == "GRID":
            nT, nD = int(r[1]), int(r[2])
            i += 1
        elif r[0] == "T":
            t["T"] = [float(v) for v in r[1:]]
            i += 1
        elif r[0] == "NE":
            t["NE"] = [float(v) for v in r[1:]]
            i += 1
        elif r[0] == "LINE":
            e = [[float(v) for v in rows[i + 1 + a]] for a in range(nT)]
            t["lines"].append({"label": r[1], "nu": int(r[2]), "nl": int(r[3]),
                               "wl": float(r[4]), "e": e})
            i += 1 + nT
        else:
            i += 1
    return t


LINE_TABLES = ("sh95_hi_caseA", "sh95_hi_caseB", "sh95_heii_caseA", "sh95_heii_caseB",
               "hei_porter_caseB")


def line_table_files(atomic_dir):
    return {k: read_line_table(os.path.join(atomic_dir, k + ".txt")) for k in LINE_TABLES
            if os.path.exists(os.path.join(atomic_dir, k + ".txt"))}


_LINES = {}


def load_line_tables(path=SNAPSHOT):
    """The recombination-line tables of the data snapshot (read once)."""
    if path not in _LINES:
        with open(path) as f:
            _LINES[path] = json.load(f).get("line_tables", {})
    return _LINES[path]


class LineTable(C.Structure):
    """Mirror of mc_line_table (include/mochii_gpu.h)."""
    _fields_ = [("nT", C.c_int32), ("nD", C.c_int32), ("nlin", C.c_int32),
                ("tgrid", C.c_double * 32), ("dgrid", C.c_double * 32), ("wl", C.c_double * 12),
                ("etab", ((C.c_double * 32) * 32) * 12)]

    @classmethod
    def make(cls, tab):
        t = cls()
        t.nT, t.nD, t.nlin = len(tab["T"]), len(tab["NE"]), len(tab["lines"])
        t.tgrid[:t.nT] = tab["T"]
        t.dgrid[:t.nD] = tab["NE"]
        for k, ln in enumerate(tab["lines"]):
            t.wl[k] = ln["wl"]
            for a, row in enumerate(ln["e"]):
                t.etab[k][a][:len(row)] = row
        return t


_NLEVEL = {}


def load_nlevel(path=SNAPSHOT):
    """The n-level atoms of the data snapshot: {'<el>_<stage>': atom}, read
    once per path (the same dict on every call; treat it as read-only)."""
    if path not in _NLEVEL:
        with open(path) as f:
            _NLEVEL[path] = json.load(f).get("nlevel", {})
    return _NLEVEL[path]


def load_hei_two_photon(path=SNAPSHOT):
    """The He I 2^1S two-photon shape of the data snapshot (data/HeI2phot.dat,
    Drake, Victor & Dalgarno 1969): (y, A) float64 arrays, y = E / 20.62 eV."""
    with open(path) as f:
        t = json.load(f).get("hei_two_photon")
    if not t:
        raise KeyError(f"{path} holds no He I two-photon table")
    return np.asarray(t["y"], dtype=np.float64), np.asarray(t["A"], dtype=np.float64)


def read_registry(atomic_dir):
    """Every element file of a data directory with its stage cooling fits:
    {name: element dict with 'cooling': {stage (1-based): fit}}."""
    out = {}
    for name in ORDER:
        path = os.path.join(atomic_dir, f"element_{name}.txt")
        if not os.path.exists(path):
            continue
        el = read_element(path)
        el["cooling"] = {}
        for i in range(1, el["nstage"] + 1):
            cp = os.path.join(atomic_dir, f"cooling_{name}_{i}.txt")
            if os.path.exists(cp):
                el["cooling"][i] = read_cooling_fit(cp)
        out[name] = el
    return out


def load_snapshot(path=SNAPSHOT):
    with open(path) as f:
        data = json.load(f)
    els = data["elements"]
    for el in els.values():
        el["cooling"] = {int(k): [tuple(r) for r in v] for k, v in el["cooling"].items()}
    return els


def registry(abund=None, data=None):
    """Active registry in species_setup order: the elements with abundance
    > 0 (par%abund_* defaults overridden by ``abund``), each a dict with its
    'abund' set."""
    ab = dict(DEFAULT_ABUND)
    ab.update(abund or {})
    data = data if data is not None else load_snapshot()
    out = []
    for name in ORDER:
        if ab.get(name, 0.0) <= 0.0:
            continue
        el = dict(data[name])
        el["abund"] = float(ab[name])
        out.append(el)
    if len(out) > MC_MAX_EL:
        raise ValueError("too many elements")
    return out


def metal_thresholds(elements):
    """metal_eth(1:nmetal) as src/main.f90:66-82 gathers them."""
    return [tr["eth"] for el in elements if el["abund"] > 0.0 for tr in el["transitions"]]


def to_struct(elements):
    """ctypes array of Element for mc_species_setup / the oracle."""
    arr = (Element * max(len(elements), 1))()
    for k, el in enumerate(elements):
        e = arr[k]
        e.name = el["name"].encode()[:3]
        e.Z, e.nstage, e.abund = el["Z"], el["nstage"], el.get("abund", 0.0)
        if el["nstage"] > MC_MAX_ST:
            raise ValueError(f"{el['name']}: {el['nstage']} stages > {MC_MAX_ST}")
        for it, tr in enumerate(el["transitions"][:el["nstage"] - 1]):
            e.eth[it] = tr["eth"]
            e.photo_form[it] = tr["photo_form"]
            for j, v in enumerate(tr["photo"][:8]):
                e.photo[it][j] = v
            for j, v in enumerate(tr["ci"]):
                e.ci[it][j] = v
            for j, v in enumerate(tr["rr"]):
                e.rr[it][j] = v
            e.ndr[it] = tr["ndr"]
            for j in range(tr["ndr"]):
                e.dr_c[it][j], e.dr_e[it][j] = tr["dr_c"][j], tr["dr_e"][j]
            if tr["dr2"] is not None:
                e.has_dr2[it] = 1
                for j, v in enumerate(tr["dr2"]):
                    e.dr2[it][j] = v
            if tr["rr2"] is not None:
                e.has_rr2[it] = 1
                for j, v in enumerate(tr["rr2"]):
                    e.rr2[it][j] = v
            e.cxi_form[it], e.cxr_form[it] = tr["cxi_form"], tr["cxr_form"]
            for j in range(6):
                e.cxi[it][j], e.cxr[it][j] = tr["cxi"][j], tr["cxr"][j]
        for st, fit in el.get("cooling", {}).items():
            if len(fit) > MC_MAX_COOLFIT:
                raise ValueError(f"{el['name']} {st}: {len(fit)} cooling terms")
            e.ncool[st - 1] = len(fit)
            for j, (a, t) in enumerate(fit):
                e.cool_A[st - 1][j], e.cool_T[st - 1][j] = a, t
    return arr
