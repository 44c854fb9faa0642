"""The oracle's Sobol/Owen launch sequence against the reference's golden
vectors (tests/qmc/q0_*.txt, q2_*.txt of the reference; src/qmc_mod.f90)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

QDIR = os.path.join(os.path.dirname(__file__), "golden", "qmc")


def _codes_from_u(u):
    """Recover the 32-bit integer code from u = (code + 0.5) 2^-32."""
    c = np.rint(u * 2.0**32 - 0.5).astype(np.int64)
    # the file prints 16 significant digits: the recovered code must
    # reproduce the printed value to that precision
    assert np.all(np.abs((c + 0.5) / 2.0**32 - u) <= 1e-15 * np.maximum(u, 1e-300) + 1e-17)
    return c


@pytest.mark.parametrize("fname,ndim", [("q0_raw.txt", 3), ("q2_raw.txt", 12)])
def test_raw_codes_bit_exact(fname, ndim):
    raw = np.loadtxt(os.path.join(QDIR, fname), dtype=np.int64)
    q = po.Qmc(12345)
    for row in raw:
        idx = int(row[0])
        got = [q.code(idx, d) for d in range(ndim)]
        assert got == list(row[1:1 + ndim]), f"index {idx}"


@pytest.mark.parametrize("fname,seed,ndim,stream", [
    ("q0_scr_12345.txt", 12345, 3, 1), ("q0_scr_777.txt", 777, 3, 1),
    ("q0_scr_2024.txt", 2024, 3, 1),
    ("q2_scr_12345.txt", 12345, 12, 1), ("q2_scr_777.txt", 777, 12, 1),
    ("q2_scr_999983.txt", 999983, 12, 1),
    ("q2_dif_12345.txt", 12345, 12, 2), ("q2_dif_777.txt", 777, 12, 2),
    ("q2_dif_999983.txt", 999983, 12, 2)])
def test_scrambled_bit_exact(fname, seed, ndim, stream):
    tab = np.loadtxt(os.path.join(QDIR, fname))
    want = _codes_from_u(tab[:, 1:1 + ndim])
    q = po.Qmc(seed)
    got = np.array([q.uniforms(int(i), ndim, stream) for i in tab[:, 0]])
    assert np.array_equal(_codes_from_u(got), want)
    # the uniform itself is exactly (code + 0.5) 2^-32
    assert np.array_equal(got, (want + 0.5) / 2.0**32)


def test_owen_net_balance():
    """Every 1-D dyadic interval 2^-j holds N/2^j of the first N=2^m scrambled
    points (the net property Owen scrambling preserves)."""
    q = po.Qmc(4242)
    N = 1024
    u = np.array([q.uniforms(i, 12, s) for s in (1, 2) for i in range(N)]).reshape(2, N, 12)
    for s in range(2):
        for d in range(12):
            for j in range(1, 11):
                h = np.bincount((u[s, :, d] * 2**j).astype(int), minlength=2**j)
                assert np.all(h == N // 2**j)


def test_against_reference_module_if_built():
    ref = po.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    import ctypes as C
    for seed in (12345, 7, -3, 2**31 - 1):
        ref.ref_qmc_setup(seed)
        q = po.Qmc(seed)
        for idx in (0, 1, 5, 1023, 2**20 + 17, 2**32 + 5, 987654321987):
            for stream in (1, 2):
                u = np.zeros(12)
                ref.ref_qmc_uniforms(idx, 12, stream, u.ctypes.data_as(C.POINTER(C.c_double)))
                assert np.array_equal(u, q.uniforms(idx, 12, stream))
