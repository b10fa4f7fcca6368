"""The Schur inverse by block elimination without pivoting (sehf_spd.hip),
through the C ABI's timing entry point, which inverts symmetric positive
definite matrices of the Schur matrices' conditioning and returns the sampled
residual max|A X - I|: single-leaf orders (n <= 64), orders whose recursion
splits unevenly, and the l_cont = 20 and 30 Schur orders.  The scan's use of
it is covered end to end by the parity tests (the default Schur inverse) and,
at full size, by tests/test_gpu_large.py, which also takes the pivoted
Gauss-Jordan inverse as the second rounding order."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,batch", [(1, 2), (17, 3), (64, 2), (65, 2), (129, 3), (441, 4), (961, 5)])
def test_spd_inverse_residual(n, batch):
    from sehf._lib import check, load
    L = load()
    ms, res = ctypes.c_double(), ctypes.c_double()
    check(L.sehf_time_spd_inverse(n, batch, 1, ctypes.byref(ms), ctypes.byref(res)), "sehf_time_spd_inverse")
    print(f"n={n} batch={batch}: {ms.value:.3f} ms, |AX - I| = {res.value:.1e}")
    assert res.value < 1e-13
    assert ms.value > 0.0


def test_spd_inverse_rejects_bad_sizes():
    from sehf._lib import load
    L = load()
    ms, res = ctypes.c_double(), ctypes.c_double()
    assert L.sehf_time_spd_inverse(0, 1, 1, ctypes.byref(ms), ctypes.byref(res)) != 0
    assert b"must be > 0" in L.sehf_last_error()
