"""Write mochii_amd/data/atomic_registry.json from a MoCHII data/atomic
directory (element_<el>.txt and cooling_<el>_<i>.txt in the reference's text
formats) the n-level atoms nlevel_<el>_<stage>.txt, the recombination-line tables
(Storey & Hummer 1995 H I / He II, Porter et al. 2012 He I) and, from the data
directory above it, the He I two-photon shape HeI2phot.dat.  The JSON is a data snapshot of the published coefficients those
files hold (VFKY96 / VY95 photoionisation, Voronov collisional ionisation,
Badnell / CHIANTI recombination, Huang et al. 2023 and Kingdon & Ferland 1996
charge exchange, CHIANTI line-cooling fits, CHIANTI level energies,
A-values and Burgess-Tully fits of the collision strengths), with each file's provenance
header kept beside it.

    python mochii_amd/data/make_atomic_json.py /path/to/MoCHII/data/atomic
"""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from mochii_amd.atomic import (ORDER, SNAPSHOT, line_table_files, nlevel_files,  # noqa: E402
                               read_registry)


def provenance(path):
    with open(path) as f:
        return [ln.rstrip("\n") for ln in f if ln.startswith("#")]


def hei_two_photon(atomic_dir):
    """data/HeI2phot.dat beside data/atomic: the He I 2^1S two-photon shape
    A(y) of Drake, Victor & Dalgarno (1969) on y = E / 20.62 eV, two columns."""
    path = os.path.join(os.path.dirname(os.path.abspath(atomic_dir.rstrip("/"))), "HeI2phot.dat")
    if not os.path.exists(path):
        return None
    # rows until the first that is not two numbers (the reference reads the
    # file list-directed and stops there: blank lines are skipped, the '!'
    # trailer ends it)
    y, a, note = [], [], []
    with open(path) as f:
        lines = f.readlines()
    k = 0
    for k, ln in enumerate(lines):
        t = ln.split()
        if not t:
            continue
        try:
            y.append(float(t[0]))
            a.append(float(t[1]))
        except (ValueError, IndexError):
            break
    note = [ln.rstrip("\n") for ln in lines[k:] if ln.strip()]
    return {"y": y, "A": a, "provenance": note,
            "source": "data/HeI2phot.dat: He I 2^1S two-photon spectral shape, "
                      "Drake, Victor & Dalgarno (1969), y = E / 20.62 eV"}


def main(atomic_dir, out=SNAPSHOT):
    reg = read_registry(atomic_dir)
    prov = {}
    for name in ORDER:
        if name not in reg:
            continue
        prov[f"element_{name}.txt"] = provenance(os.path.join(atomic_dir, f"element_{name}.txt"))
        for st in reg[name]["cooling"]:
            fn = f"cooling_{name}_{st}.txt"
            prov[fn] = provenance(os.path.join(atomic_dir, fn))[:4]
    nlev = nlevel_files(atomic_dir)
    for key in nlev:
        prov[f"nlevel_{key}.txt"] = provenance(os.path.join(atomic_dir, f"nlevel_{key}.txt"))[:3]
    lt = line_table_files(atomic_dir)
    for key in lt:
        prov[f"{key}.txt"] = provenance(os.path.join(atomic_dir, f"{key}.txt"))[:3]
    doc = {"source": "MoCHII data/atomic (element, cooling-fit, n-level atom and "
                     "recombination-line files)",
           "elements": {k: dict(v, cooling={str(s): f for s, f in v["cooling"].items()})
                        for k, v in reg.items()},
           "nlevel": nlev,
           "line_tables": lt,
           "provenance": prov}
    h2 = hei_two_photon(atomic_dir)
    if h2:
        doc["hei_two_photon"] = h2
    with open(out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {out}: {len(reg)} elements, {len(nlev)} n-level atoms")


if __name__ == "__main__":
    main(sys.argv[1])
