"""Kernel-level checks: GEMM shapes/transposes and Gauss-Jordan orders.

n <= 16 runs only the register-resident pivot panel; 17..32 adds one MFMA
block update; larger orders nest the recursion.  GEMM runs through the
dipole-engine entry point (one plain GEMM).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


@pytest.mark.parametrize("ns,Np", [(1, 1), (4, 16), (16, 16), (7, 81), (64, 64), (70, 130), (11, 961)])
def test_gemm_shapes(ctx, ns, Np):
    from sehf.steppers import GpuDipoleEngine
    rng = np.random.default_rng(ns * 1000 + Np)
    V = rng.standard_normal((ns, Np))
    P = rng.standard_normal((Np, Np))
    eng = GpuDipoleEngine(ctx, Np, ns)
    got = eng.step(V, P)
    print(ns, Np, rel_err(got, V @ P))
    assert rel_err(got, V @ P) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 17, 24, 32, 33, 48, 64, 100])
def test_gj_orders(ctx, n):
    from sehf.steppers import GpuSinvStepper
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * 0.0
    st = GpuSinvStepper(ctx, n, 0)
    got, _ = st.step_inverse_only(A)
    ref = np.linalg.inv(A)
    ref = 0.5 * (ref + ref.T)
    print(n, rel_err(got, ref))
    assert rel_err(got, ref) < 1e-10


@pytest.mark.parametrize("n", [1, 5, 16, 17, 33, 100, 1024, 1025, 1566])
def test_gj_orders_partial_pivoting(ctx, n):
    """The one-workgroup leaf with partial pivoting (sehf_set_gj_pivoting(1)),
    one and two rows per thread, against numpy's inverse; the tournament
    leaf's result on the same matrix agrees to rounding."""
    from sehf._lib import check, load
    from sehf.steppers import GpuSinvStepper
    L = load()
    rng = np.random.default_rng(7 * n + 1)
    A = rng.standard_normal((n, n))
    st = GpuSinvStepper(ctx, n, 0)
    check(L.sehf_set_gj_pivoting(1), "sehf_set_gj_pivoting")
    try:
        got, _ = st.step_inverse_only(A)
    finally:
        check(L.sehf_set_gj_pivoting(0), "sehf_set_gj_pivoting")
    tour, _ = st.step_inverse_only(A)
    ref = np.linalg.inv(A)
    ref = 0.5 * (ref + ref.T)
    print(n, rel_err(got, ref), rel_err(got, tour))
    assert rel_err(got, ref) < 1e-10
    assert rel_err(got, tour) < 1e-10


def test_gj_pivoting_mode_rejects_bad_values(ctx):
    from sehf._lib import load
    L = load()
    assert L.sehf_set_gj_pivoting(2) != 0
    assert L.sehf_set_gj_pivoting(-1) != 0
    assert L.sehf_set_gj_pivoting(0) == 0
