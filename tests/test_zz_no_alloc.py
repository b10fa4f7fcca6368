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


def _hip():
    """The HIP runtime libsehf is linked against (already loaded with it, so
    the same runtime instance): caller-owned device buffers come from it
    directly rather than from a second runtime in the process."""
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    for f in (H.hipMalloc, H.hipMemcpy, H.hipFree, H.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return H


H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _noalloc(L, N, W, R):
    H = _hip()
    nb = N * N * 8
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b in bufs:
        assert H.hipMalloc(ctypes.byref(b), nb) == 0
    try:
        Wd, Rd, Od = bufs
        # column-major N x N in caller-owned device buffers
        w = np.asfortranarray(W).ravel(order="F").copy()
        r = np.asfortranarray(R).ravel(order="F").copy()
        o = np.full(N * N, np.nan)
        assert H.hipMemcpy(Wd, w.ctypes.data, nb, H2D) == 0
        assert H.hipMemcpy(Rd, r.ctypes.data, nb, H2D) == 0
        assert H.hipMemcpy(Od, o.ctypes.data, nb, H2D) == 0
        sym = ctypes.c_double(-1.0)
        rc = L.sehf_frp_step_noalloc(N, Wd, Rd, Od, ctypes.byref(sym))
        assert rc == 0, L.sehf_last_error().decode()
        assert H.hipDeviceSynchronize() == 0
        assert H.hipMemcpy(o.ctypes.data, Od, nb, D2H) == 0
        return o.reshape(N, N, order="F"), sym.value
    finally:
        for b in bufs:
            H.hipFree(b)


def test_frp_step_noalloc_matches_frp_step():
    from sehf._lib import load
    from sehf.steppers import GpuContext, GpuForwardStepper
    L = load()
    ctx = GpuContext(True)
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
