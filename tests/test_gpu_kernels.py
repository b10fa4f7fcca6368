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
