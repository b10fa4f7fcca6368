"""The allocation-free forward step (sehf_frp_step_noalloc) against sehf_frp_step.

Both run the same kernels, so the results must be bitwise equal: two calls in
a row on the shared static workspace, the second of a smaller order, with
inputs and outputs in caller-owned device buffers.  Orders beyond the
workspace's maximum are rejected before any device work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(N, seed, cond=50.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N, N))
    S = A @ A.T / N
    lam, V = np.linalg.eigh(S)
    return (V * np.maximum(lam, lam.max() / cond)) @ V.T


def _noalloc(L, N, W, R):
    import torch
    # column-major N x N in caller-owned device buffers
    Wd = torch.from_numpy(np.asfortranarray(W).ravel(order="F").copy()).cuda()
    Rd = torch.from_numpy(np.asfortranarray(R).ravel(order="F").copy()).cuda()
    Od = torch.full((N * N,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sym = ctypes.c_double(-1.0)
    rc = L.sehf_frp_step_noalloc(N, ctypes.c_void_p(Wd.data_ptr()), ctypes.c_void_p(Rd.data_ptr()),
                                 ctypes.c_void_p(Od.data_ptr()), ctypes.byref(sym))
    assert rc == 0, L.sehf_last_error().decode()
    return Od.cpu().numpy().reshape(N, N, order="F"), sym.value


def test_frp_step_noalloc_matches_frp_step():
    import torch
    from sehf._lib import load
    from sehf.steppers import GpuContext, GpuForwardStepper
    L = load()
    ctx = GpuContext(True)
    torch.cuda.set_device(0)
    for N, seeds in ((1046, (42, 43)), (441, (44, 45))):   # second call: smaller order
        W, R = _spd(N, seeds[0]), _spd(N, seeds[1])
        st = GpuForwardStepper(ctx, N)
        st.init_R_prev_inv(R)
        ref, sym_ref = st.step(W)
        got, sym = _noalloc(L, N, W, R)
        print(f"N={N}: max|diff|={np.max(np.abs(got - ref)):.1e} sym={sym:.1e} (stepper {sym_ref:.1e})")
        assert np.array_equal(got, ref)
        assert sym == sym_ref
    # beyond the static workspace: a clear error, nothing launched
    rc = L.sehf_frp_step_noalloc(1601, ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1), None)
    assert rc != 0 and b"outside [1, 1600]" in L.sehf_last_error()
