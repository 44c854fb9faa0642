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


@pytest.mark.parametrize("rows,width,n_rows", [(1 << 20, 4, 10_000_000), (4099, 8, 3_000_001),
                                               (1, 2, 1000), (77, 1, 12345)])
def test_row_probe_counts_every_update(rows, width, n_rows):
    """mc_probe_row_atomics: `width` lanes per row, every add lands."""
    L = library()
    ms, s = C.c_double(), C.c_double()
    check(L.mc_probe_row_atomics(rows, width, n_rows, C.byref(ms), C.byref(s)),
          "mc_probe_row_atomics")
    assert s.value == 2.0 * n_rows * width
    assert ms.value > 0.0


def test_row_probe_costs_one_request_per_row():
    """The rows of four doubles issued by four lanes together complete at
    least twice the doubles per second of one lane adding them in turn (the
    device counts one atomic request per 64-byte line an instruction touches;
    measured about four times)."""
    L = library()
    ms1, ms4 = C.c_double(), C.c_double()
    rows, n = 4_000_000, 200_000_000
    check(L.mc_probe_scattered_rmw(rows, 4, n, C.byref(ms1), None), "mc_probe_scattered_rmw")
    check(L.mc_probe_row_atomics(rows, 4, n, C.byref(ms4), None), "mc_probe_row_atomics")
    print(f"4-double rows: one lane {ms1.value:.2f} ms, four lanes {ms4.value:.2f} ms")
    assert ms4.value * 2.0 < ms1.value


def test_row_probe_refuses_bad_shapes():
    L = library()
    ms = C.c_double()
    for args in ((0, 4, 10), (10, 3, 10), (10, 16, 10), (10, 4, 0)):
        with pytest.raises(MochiiError, match="width"):
            check(L.mc_probe_row_atomics(*args, C.byref(ms), None), "mc_probe_row_atomics")
