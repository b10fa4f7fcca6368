"""Per-step compatibility ABI vs direct references.

Mirrors src/tests/test_gpu_steppers.cpp: random SPD matrices (N=128, N_psi=40,
N_f=88, seeds 42/43/...), one forward step at 1e-11, five stacked steps at
1e-10, back step psi/f at 1e-12, compute_f=false.  The forward reference is the
oracle's restatement of forward_reference (test_gpu_steppers.cpp:85-95).
Larger orders exercise every level of the recursive Gauss-Jordan inverse.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def make_random_spd(N, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N, N))
    S = A @ A.T / N
    lam, V = np.linalg.eigh(S)
    lam = np.maximum(lam, lam.max() / cond)
    return (V * lam) @ V.T


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


def test_context_info(ctx):
    info = ctx.info()
    print("device:", info)
    assert info["is_gpu"] and info["global_mem_bytes"] > 0


@pytest.mark.parametrize("N", [128, 162, 441, 1046, 1566])
def test_forward_single_step(ctx, N):
    import oracle_py as O
    from sehf.steppers import GpuForwardStepper
    W = make_random_spd(N, 42)
    R = make_random_spd(N, 43)
    ref = O.frp_step(W, R)
    st = GpuForwardStepper(ctx, N)
    st.init_R_prev_inv(R)
    got, sym = st.step(W)
    print(f"N={N} rel={rel_err(got, ref):.3e} sym={sym:.1e}")
    assert rel_err(got, ref) < 1e-11
    assert sym == 0.0


def test_forward_five_steps(ctx):
    import oracle_py as O
    from sehf.steppers import GpuForwardStepper
    N = 128
    R = make_random_spd(N, 99)
    st = GpuForwardStepper(ctx, N)
    st.init_R_prev_inv(R)
    Rh = R.copy()
    for k in range(5):
        W = make_random_spd(N, 100 + k)
        Rh = O.frp_step(W, Rh)
        Rg, _ = st.step(W)
    assert rel_err(Rg, Rh) < 1e-10


@pytest.mark.parametrize("compute_f", [True, False])
def test_back_step(ctx, compute_f):
    from sehf.steppers import GpuBackStepper
    N, Np = 128, 40
    Nf = N - Np
    Rinv = make_random_spd(N, 200)
    W = make_random_spd(N, 201)
    Z = np.random.default_rng(300).standard_normal((N, Np))
    st = GpuBackStepper(ctx, N, Np, Nf)
    st.init_Z(Z)
    psi, f = st.step(Rinv, W, compute_f=compute_f)
    Zn = Rinv @ Z
    Y = W @ Zn
    assert rel_err(psi, Y[:Np]) < 1e-12
    assert rel_err(st.get_Z(), Zn) < 1e-12
    if compute_f:
        assert rel_err(f, Y[Np:]) < 1e-12
    else:
        assert f is None


@pytest.mark.parametrize("Np,Nf", [(40, 88), (441, 605), (961, 605)])
def test_sinv_step(ctx, Np, Nf):
    from sehf.steppers import GpuSinvStepper
    rng = np.random.default_rng(7)
    A = make_random_spd(Np, 11) + 2.0 * np.eye(Np)
    B = 0.05 * rng.standard_normal((Np, Nf))
    D = rng.uniform(0.5, 2.0, Nf)
    S = A - (B * D) @ B.T
    ref = np.linalg.inv(S)
    ref = 0.5 * (ref + ref.T)
    st = GpuSinvStepper(ctx, Np, Nf)
    got, sym = st.step(A, B, D)
    assert rel_err(got, ref) < 1e-11
    got2, _ = st.step_inverse_only(A)
    ref2 = np.linalg.inv(A)
    assert rel_err(got2, 0.5 * (ref2 + ref2.T)) < 1e-11


def test_dipole_engine(ctx):
    from sehf.steppers import GpuDipoleEngine
    rng = np.random.default_rng(5)
    Np, ns = 81, 7
    V = rng.standard_normal((ns, Np))
    P = rng.standard_normal((Np, Np))
    eng = GpuDipoleEngine(ctx, Np, ns)
    assert rel_err(eng.step(V, P), V @ P) < 1e-13


def test_nonsymmetric_inverse_via_sinv_inverse_only(ctx):
    """Pivoting on a matrix whose leading entry is tiny (forces row interchanges)."""
    from sehf.steppers import GpuSinvStepper
    rng = np.random.default_rng(9)
    N = 300
    A = rng.standard_normal((N, N))
    A = A + A.T
    A[0, 0] = 1e-12
    st = GpuSinvStepper(ctx, N, 0)
    got, _ = st.step_inverse_only(A)
    ref = np.linalg.inv(A)
    assert rel_err(got, 0.5 * (ref + ref.T)) < 1e-10


def test_stepper_stats(ctx):
    """stats() / reset_stats() (GpuPropagate.hpp Stats, GpuDipoleEngine.hpp
    Stats): per-phase nanoseconds accumulate over steps, n_steps counts the
    steps, and reset clears every field."""
    from sehf.steppers import GpuBackStepper, GpuDipoleEngine, GpuForwardStepper, GpuSinvStepper
    rng = np.random.default_rng(21)
    N, Np, Nf = 96, 40, 56
    fr = GpuForwardStepper(ctx, N)
    fr.init_R_prev_inv(make_random_spd(N, 1))
    for k in range(3):
        fr.step(make_random_spd(N, 2 + k))
    si = GpuSinvStepper(ctx, Np, Nf)
    si.step(make_random_spd(Np, 5) + 5 * np.eye(Np), 0.1 * rng.standard_normal((Np, Nf)), np.full(Nf, 0.5))
    si.step_inverse_only(make_random_spd(Np, 6))
    bp = GpuBackStepper(ctx, N, Np, Nf)
    bp.init_Z(rng.standard_normal((N, Np)))
    bp.step(make_random_spd(N, 7), make_random_spd(N, 8), compute_f=True)
    de = GpuDipoleEngine(ctx, Np, 7)
    de.step(rng.standard_normal((7, Np)), rng.standard_normal((Np, Np)))
    for obj, n, phases in ((fr, 3, ("t_upload_ns", "t_u_combine_ns", "t_inverse_ns", "t_download_ns")),
                           (si, 2, ("t_upload_ns", "t_schur_ns", "t_inverse_ns", "t_download_ns")),
                           (bp, 1, ("t_upload_ns", "t_gemm_z_ns", "t_gemm_y_ns", "t_extract_ns")),
                           (de, 1, ("t_upload_ns", "t_gemm_ns", "t_download_ns"))):
        s = obj.stats()
        print(type(obj).__name__, s)
        assert s["n_steps"] == n and all(s[p] > 0 for p in phases), s
        obj.reset_stats()
        assert all(v == 0 for v in obj.stats().values())
