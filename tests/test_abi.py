"""The C-ABI library builds/loads and exports every entry point declared in
include/mochii_gpu.h; its struct layout matches the Python mirror.  No
compute calls (no GPU in this test)."""
import ctypes as C
import os
import re

import pytest

import mochii_amd
from mochii_amd.lib import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mochii_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mc_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = mochii_amd.library()          # raises if the library was not built
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_params_layout_matches_header():
    """sizeof(mc_params) as laid out by the C compiler equals the ctypes mirror."""
    import subprocess
    import tempfile
    src = ('#include "mochii_gpu.h"\n#include <stdio.h>\n#include <stddef.h>\n'
           'int main(){printf("%zu %zu %zu %zu\\n", sizeof(mc_params),'
           ' offsetof(mc_params, ext_geometry), offsetof(mc_params, He_abund),'
           ' offsetof(mc_params, lambda_ref)); return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "t.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe,
                               os.path.join(d, "t.c")])
        got = [int(x) for x in subprocess.check_output([exe]).split()]
    assert got == [C.sizeof(Params), Params.ext_geometry.offset, Params.He_abund.offset,
                   Params.lambda_ref.offset]


def test_errors_are_reported_not_aborted():
    lib = mochii_amd.library()
    h = C.c_void_p()
    p = Params.make(nphotons=0)
    rc = lib.mc_init(C.byref(p), C.byref(h))
    assert rc < 0 and lib.mc_last_error()
    assert lib.mc_ionizing_field_and_rates(None, 0) < 0
    assert b"null context" in lib.mc_last_error()


def test_fortran_binding_layout_matches_header():
    """The bind(C) mc_params of the Fortran module has the C struct's size and
    field offsets (the driver prints them; no GPU needed)."""
    import subprocess
    exe = os.path.join(ROOT, "mochii_amd", "fortran", "_build", "drive_stromgren")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (amdflang absent)")
    out = subprocess.check_output([exe, "layout"], text=True).split("\n")
    got = dict(line.split() for line in out if line.strip() and not line.startswith("Fortran"))
    el = {k.split(".", 1)[1]: v for k, v in got.items() if k.startswith("element.")}
    got = {k: v for k, v in got.items() if not k.startswith("element.")}
    assert int(got.pop("sizeof")) == C.sizeof(Params)
    for name, off in got.items():
        assert int(off) == getattr(Params, name).offset, name
    from mochii_amd.atomic import Element
    assert int(el.pop("sizeof")) == C.sizeof(Element)
    for name, off in el.items():
        assert int(off) == getattr(Element, name).offset, name


def test_element_layout_matches_header_and_oracle():
    """sizeof/offsets of mc_element (C compiler) equal the ctypes mirror and
    the oracle's orc_element, so one host record serves both."""
    import subprocess
    import tempfile
    from mochii_amd.atomic import Element
    fields = ["abund", "photo", "ndr", "dr2", "rr2", "cxr", "ncool", "cool_T"]
    src = ('#include "mochii_gpu.h"\n#include <stdio.h>\n#include <stddef.h>\n'
           'int main(){printf("%zu", sizeof(mc_element));' +
           "".join(f'printf(" %zu", offsetof(mc_element, {f}));' for f in fields) +
           'printf("\\n"); return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "t.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe,
                               os.path.join(d, "t.c")])
        got = [int(x) for x in subprocess.check_output([exe]).split()]
    assert got == [C.sizeof(Element)] + [getattr(Element, f).offset for f in fields]
    from oracle import pyoracle as po
    assert po.lib().orc_element_size() == C.sizeof(Element)
