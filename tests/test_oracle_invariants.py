"""CPU pins of the oracle's stage chain (Sinv -> FRP -> K -> BP -> fit -> dipole)
by what the mathematics fixes, independently of the GPU path.

* Numerov three-point identity on the back-propagated psi (exchange off):
  W_{n+1} psi_{n+1} - (12 - 10 W_n) psi_n + W_{n-1} psi_{n-1} = 0 with
  W_n = I + h^2/6 (E - V_n), the Numerov form of -psi''/2 + V psi = E psi
  (ForwardRPropagator.cpp:339-358, BackPropagator.cpp:533-703).  A wrong
  index, sign or scaling in the ratio recursion or the back-propagation
  breaks it at O(1).
* Sinv equals the psi-psi block of the inverse of the full Numerov matrix
  [[A, B], [B^T, D]] (the Schur-complement identity behind
  SchurInverter.cpp:247-415; src/tests/test_schur_inverter.cpp:144-261).
* Free s wave: with V = 0 the reference's radial convention (solver point n
  at r = (n+1) h, wavefunction vanishing one step inside, WavefunctionSetup
  .cpp:327) makes the exact Numerov solution sin(k(r - h)), so the
  two-point K matrix and the asymptotic fit must both give K = tan(k h).
* Gauge identity for an exact eigenstate: for H = -nabla^2/2 - 1/r and the
  hydrogen 1s orbital, <f|d/dz|i> = -(E_f - E_i) <f|z|i>.  The reference's
  dipole radial grid (r = (ir+1) dr against chi at ir dr,
  DipoleMatrixElement.cpp:204-233) makes the ratio converge at first order
  in dr; the Richardson extrapolation of two grids must give -1.  The
  velocity-gauge coefficients c(l_mu, l_nu) and the five-point derivative
  enter here, and a sign or coefficient mistake moves the ratio by O(1).
* Isotropy: for a spherical potential and an s initial state the x, y and z
  slots carry the same norm (the m bookkeeping of the angular couplings).
* Two independent K extractions (the two-point match at N-2 and the
  least-squares fit over the asymptotic window) agree on a molecule with
  exchange.
"""
import dataclasses
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


def _no_exchange(p):
    return dataclasses.replace(p, psi_lm=np.zeros_like(p.psi_lm))


def test_numerov_three_point_identity(O):
    from sehf import synth
    p = _no_exchange(synth.h2_like(l_cont=2, dr=0.02, r_max=12.0, seed=0))
    k = 0.8
    E, h, N = 0.5 * k * k, p.dr, p.n_psi
    ns = [100, 101, 102, 200, 201, 202]          # n >= 60 (no stability shift), < n_transition = 251
    o = O.solve(p, k, dump_psi=ns)
    psi = {n: o["psi"][i] for i, n in enumerate(ns)}

    def W(n):
        return np.eye(N) + h * h / 6.0 * (E * np.eye(N) - O.potential(p, n))

    for n in (101, 201):
        res = W(n + 1) @ psi[n + 1] - (12.0 * np.eye(N) - 10.0 * W(n)) @ psi[n] + W(n - 1) @ psi[n - 1]
        assert np.abs(res).max() < 1e-12 * np.abs(psi[n]).max(), (n, np.abs(res).max())


def test_sinv_is_schur_block_of_full_inverse(O):
    from sehf import synth
    p = synth.h2_like(l_cont=2, dr=0.02, r_max=12.0, seed=0)
    k = 0.8
    E, h, N = 0.5 * k * k, p.dr, p.n_psi
    ns = [80, 150]
    o = O.solve(p, k, dump_sinv=ns)
    l_sig = np.array([int(math.isqrt(s)) for s in range(p.n_sigma)] * p.n_occ, dtype=float)
    for i, n in enumerate(ns):
        A = np.eye(N) + h * h / 6.0 * (E * np.eye(N) - O.potential(p, n))
        B = h * h / 12.0 * O.exchange_Q(p, n)
        r = (n + 1) * h
        D = 1.0 - h * h / 12.0 * l_sig * (l_sig + 1.0) / (r * r)
        assert D.min() > 0.001                    # the D^{-1} clamp is inactive here
        Wf = np.block([[A, B], [B.T, np.diag(D)]])
        top = np.linalg.inv(Wf)[:N, :N]
        assert np.abs(B).max() > 1e-6             # exchange present
        assert np.abs(top - o["sinv"][i]).max() < 1e-12 * np.abs(top).max()


@pytest.mark.parametrize("dr,k", [(0.02, 0.5), (0.01, 1.0), (0.005, 1.3)])
def test_free_s_wave_phase(O, dr, k):
    from sehf import synth
    p = synth.h2_like(l_cont=0, dr=dr, r_max=20.0, seed=0)
    pf = dataclasses.replace(_no_exchange(p), Z=np.zeros(0, dtype=np.int32), xyz=np.zeros((0, 3)),
                             V_H=np.zeros_like(p.V_H))
    o = O.solve(pf, k)
    ref = math.tan(k * dr)
    assert abs(o["K_kme"][0, 0] - ref) < 1e-6 * ref, (o["K_kme"][0, 0], ref)
    assert abs(o["K_fit"][0, 0] - ref) < 1e-6 * ref, (o["K_fit"][0, 0], ref)


def _hydrogen_ratio(O, dr, k):
    from sehf import synth
    p = synth.h2_like(l_cont=2, dr=dr, r_max=20.0, seed=0)
    r = np.arange(p.Nr) * p.dr
    F = np.zeros_like(p.psi_lm)
    F[0, 0] = 1e-6 * 2.0 * np.exp(-r)   # hydrogen 1s as the HOMO; the 1e-6 scale makes exchange ~1e-12
    ph = dataclasses.replace(p, Z=np.array([1], dtype=np.int32), xyz=np.zeros((1, 3)),
                             V_H=np.zeros_like(p.V_H), psi_lm=F, init_psi_lm=None, allow_closed=True)
    o = O.solve(ph, k)
    omega = 0.5 * k * k + 0.5
    iz = 2                              # channel (l=1, m=0)
    norms = [float(np.linalg.norm(o["d_raw"][s])) for s in range(6)]
    return o["d_raw"][5, iz] / (omega * o["d_raw"][2, iz]), norms


@pytest.mark.parametrize("k", [0.5, 1.0])
def test_gauge_identity_hydrogen(O, k):
    r1, n1 = _hydrogen_ratio(O, 0.01, k)
    r2, n2 = _hydrogen_ratio(O, 0.005, k)
    e1, e2 = r1 + 1.0, r2 + 1.0
    assert 1.7 < e1 / e2 < 2.3, (r1, r2)          # first order in dr
    assert abs(2.0 * r2 - r1 + 1.0) < 5e-4, (r1, r2)
    for norms in (n1, n2):                          # isotropy, length and velocity
        assert abs(norms[0] - norms[2]) < 1e-12 * norms[2] and abs(norms[1] - norms[2]) < 1e-12 * norms[2]
        assert abs(norms[3] - norms[5]) < 1e-12 * norms[5] and abs(norms[4] - norms[5]) < 1e-12 * norms[5]


def test_fit_agrees_with_two_point_K(O):
    from sehf import synth
    p = synth.h2_like(l_cont=2, dr=0.01, r_max=20.0, seed=0)
    for k in (0.6, 1.1):
        o = O.solve(p, k)
        K, Kf = o["K_kme"], o["K_fit"]
        assert np.abs(Kf - K).max() < 2e-3 * np.abs(K).max(), (k, np.abs(Kf - K).max() / np.abs(K).max())
        assert o["diag"][11] < 1e-12                # S = (I + iK)(I - iK)^-1 unitary
