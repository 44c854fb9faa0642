"""lines_write (src/lines_mod.f90:37-352): the line-luminosity table of a
converged model, '<base>_lines.txt', in the reference's text layout:
comment header, then one row per line 'elem stage lambda[A] L[erg/s]
L/L(Hbeta) [label]' (Fortran format a4,i4,f12.2,2es14.5,2x,a)."""
import numpy as np

ELEM = {1: "h", 2: "he", 6: "c", 7: "n", 8: "o", 10: "ne", 12: "mg", 14: "si", 16: "s", 17: "cl",
        18: "ar", 20: "ca", 26: "fe"}


def _es(v, d=5):
    """Fortran ES14.5: one digit before the point, two-digit exponent."""
    return f"{v:14.{d}E}"


def format_row(elem, stage, wl, L, ratio, label=""):
    row = f"{elem:>4s}{stage:4d}{wl:12.2f}{_es(L)}{_es(ratio)}"
    return row + ("  " + label if label else "")


def lines_write(model, path, lines=None):
    """Write the line table of model's current state (model.lines_setup must
    have been called); returns the lines dict of model.lines_compute."""
    lines = lines or model.lines_compute()
    L, wl, ion, lab = lines["L"], lines["wl"], lines["ion"], lines["labels"]
    ihb = [k for k, s in enumerate(lab) if s == "Hbeta"]
    LHb = L[ihb[0]] if ihb else np.nan
    case = getattr(model, "line_case", "B")
    with open(path, "w") as f:
        f.write("# MoCHII line luminosities (converged state)\n")
        f.write(f"# H I recombination lines: Storey & Hummer (1995) case {case}, bilinear "
                f"(log T, log ne) interpolation\n")
        f.write(f"# (SH95 H I/He II lines use case {case}; He I Porter lines are case B)\n")
        f.write("# metals: n-level solve leaf by leaf\n")
        f.write(f"# L(Hbeta) = {_es(LHb, 6)} erg/s\n")
        f.write("# elem stage  lambda[A]      L[erg/s]     L/L(Hbeta)\n")
        for k in range(L.size):
            z, st = divmod(int(ion[k]), 100)
            f.write(format_row(ELEM.get(z, str(z)), st, wl[k], L[k], L[k] / LHb, lab[k]) + "\n")
    return lines
