"""Write tests/golden/rates/cooling_rates.txt: the reference's recombination
cooling coefficients and free-free Gaunt factor as its Python mirror of the
Fortran evaluates them (tools/python/mochii_rates.py, which the reference's
tests/rates/check_python_rates.py holds to the Fortran at 1e-10):
betaA_HII, betaB_HII, beta_HeII and beta_HeIII (case A and B) of
src/cooling_mod.f90:148-190 and gbar_ff(T, Z = 1..4) of
src/cooling_mod.f90:220-240, on 81 temperatures log-spaced over 1e3-1e5 K
plus the 41 nodes of the Gaunt table.

Run in a checkout that holds the reference at /root/reference:
    python tests/golden/make_rates_fixtures.py
"""
import os
import sys

import numpy as np

REF = "/root/reference/tools/python"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rates", "cooling_rates.txt")


def main():
    sys.path.insert(0, REF)
    import mochii_rates as R
    T = np.unique(np.concatenate([np.geomspace(1e3, 1e5, 81), 10 ** (2.0 + 0.1 * np.arange(41))]))
    cols = [T, R.betaA_HII(T), R.betaB_HII(T), R.beta_HeII(T, True), R.beta_HeII(T, False),
            R.beta_HeIII(T, True), R.beta_HeIII(T, False)]
    cols += [R.gbar_ff(T, z) for z in (1, 2, 3, 4)]
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hdr = ("written by tests/golden/make_rates_fixtures.py from the reference's "
           "tools/python/mochii_rates.py\n"
           "T[K] betaA_HII betaB_HII beta_HeII(A) beta_HeII(B) beta_HeIII(A) beta_HeIII(B) "
           "gbar_ff(Z=1) gbar_ff(Z=2) gbar_ff(Z=3) gbar_ff(Z=4)")
    np.savetxt(OUT, np.column_stack(cols), fmt="%.17e", header=hdr)
    print(f"wrote {OUT}: {T.size} rows")


if __name__ == "__main__":
    main()
