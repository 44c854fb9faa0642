"""The scattered read-modify-write probe (mc_probe_scattered_rmw): every
update lands (the array sums to 2 n_ops width exactly, two runs of n_ops
adds of 1.0), and arguments outside its limits are refused."""
import ctypes as C

import pytest

from mochii_amd.lib import MochiiError, check, library

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,width,n_ops", [(1 << 20, 4, 10_000_000), (4099, 1, 3_000_001),
                                              (1, 16, 1000)])
def test_probe_counts_every_update(rows, width, n_ops):
    L = library()
    ms, s = C.c_double(), C.c_double()
    check(L.mc_probe_scattered_rmw(rows, width, n_ops, C.byref(ms), C.byref(s)),
          "mc_probe_scattered_rmw")
    assert s.value == 2.0 * n_ops * width
    assert ms.value > 0.0


def test_probe_refuses_bad_shapes():
    L = library()
    ms = C.c_double()
    for args in ((0, 4, 10), (10, 0, 10), (10, 17, 10), (10, 4, 0)):
        with pytest.raises(MochiiError, match="width"):
            check(L.mc_probe_scattered_rmw(*args, C.byref(ms), None), "mc_probe_scattered_rmw")
