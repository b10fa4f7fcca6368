"""CPU pins of the oracle functions that the end-to-end invariants leave
inactive: each case makes the function act, and checks it against what the
mathematics fixes or against an independent evaluation, chosen so that a
plausible mistake (a dropped factor, a sign, an index or radius shifted by
one point, a transposed operand) fails.

* Stability shift below n = 60 (SchurInverter.cpp:30-59, :247-415): cases
  where A (and, with a strong exchange, S) is shifted.  The shift counts match
  a numpy eigen-decomposition of the unshifted matrices, every shifted matrix
  has lambda_min = W_min + 0.01 exactly (the shift is W_min - lambda_min +
  0.01), and every Sinv_n below n = 60 inverts the expected (shifted or not)
  Schur complement.
* D^{-1} clamp (SchurInverter.cpp:332-339): at n = 1 with l_sigma = 7, 8 the
  f-block diagonal 1 - h^2/12 l(l+1)/r^2 is negative; Sinv must equal the
  psi-psi block of the inverse of the full Numerov matrix with D clamped to
  W_min, and differ measurably from the unclamped one.
* Analytic start (ForwardRPropagator.cpp:78-134) with n_start = 5: the
  diagonal ratio M_n j(k r_n) / (M_{n+1} j(k r_{n+1})) against scipy's
  spherical Bessel functions, and the f block's r^{l+1} ratio.
* /initial_state convention (main.cpp:781-802): the initial state is read raw,
  with no r factor, while the bystander orbitals are r F.  Storing r F of the
  HOMO as the initial state gives bit for bit the HOMO run; storing F does not.
* Lagrange orthogonalisation of the dipole (DipoleMatrixElement.cpp:584-965,
  N_alpha = 2): the reduced dipole against a brute-force evaluation that forms
  psi'_beta = psi_beta - 2 sum_alpha b_beta,alpha phi_alpha explicitly from
  the dumped psi_n at every radial point and integrates <psi'|r_q|Phi_i> with
  its own Gaunt table and Simpson sum, then applies M = ((A - iB)^{-1})^dagger.
"""
import dataclasses
import math

import numpy as np
import pytest
from scipy.special import spherical_jn

from conftest import rel_err

W_MIN, MARGIN, STAB_N = 0.001, 0.01, 60


@pytest.fixture(scope="module")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="module")
def H():
    from sehf import _lib
    return _lib.load()


def _l_sig(p):
    return np.array([int(math.isqrt(s)) for s in range(p.n_sigma)] * p.n_occ, dtype=float)


def _A(O, p, n, E):
    N, h = p.n_psi, p.dr
    A = np.eye(N) + h * h / 6.0 * (E * np.eye(N) - O.potential(p, n))
    return 0.5 * (A + A.T) if n < STAB_N else A


def _D(p, n, clamp=True):
    h = p.dr
    r = (n + 1) * h
    ls = _l_sig(p)
    D = 1.0 - h * h / 12.0 * ls * (ls + 1.0) / (r * r)
    return np.maximum(D, W_MIN) if clamp else D


def _shifted(M):
    lam = np.linalg.eigvalsh(0.5 * (M + M.T)).min()
    if lam < W_MIN:
        return M + (W_MIN - lam + MARGIN) * np.eye(M.shape[0]), True
    return M, False


def _expected_schur(O, p, n, E, n_tr):
    """(S, A shifted, S shifted) at solver point n, from numpy eigenvalues."""
    A, sa = _A(O, p, n, E), False
    if n < STAB_N:
        A, sa = _shifted(A)
    if n >= n_tr:
        return A, A, sa, False
    B = p.dr ** 2 / 12.0 * O.exchange_Q(p, n)
    S = A - (B / _D(p, n)) @ B.T
    ss = False
    if n < STAB_N:
        S, ss = _shifted(S)
    return S, A, sa, ss


@pytest.mark.parametrize("case", ["A_shift", "A_and_S_shift"])
def test_stability_shift_below_60(O, case):
    from sehf import synth
    if case == "A_shift":
        p = synth.h2o_like(l_cont=4, dr=0.02, r_max=14.0, n_occ=3, seed=1)
    else:   # coarse grid and a 100x exchange: the Schur complement needs its own shift
        p0 = synth.h2o_like(l_cont=4, dr=0.1, r_max=14.0, n_occ=3, seed=1)
        p = dataclasses.replace(p0, psi_lm=100.0 * p0.psi_lm)
    k = 0.3
    E = 0.5 * k * k
    ns = list(range(min(STAB_N, p.Nr)))
    o = O.solve(p, k, dump_sinv=ns)
    n_tr = int(o["diag"][4])
    cnt_a = cnt_s = 0
    for i, n in enumerate(ns):
        S, A, sa, ss = _expected_schur(O, p, n, E, n_tr)
        cnt_a += sa
        cnt_s += ss
        S_got = np.linalg.inv(o["sinv"][i])
        assert rel_err(S_got, S) < 1e-11, (case, n, rel_err(S_got, S))
        lam = np.linalg.eigvalsh(0.5 * (S_got + S_got.T)).min()
        assert lam >= W_MIN - 1e-12, (case, n, lam)
        if ss or (sa and n >= n_tr):
            assert abs(lam - (W_MIN + MARGIN)) < 1e-11, (case, n, lam)
        if sa:
            assert abs(np.linalg.eigvalsh(A).min() - (W_MIN + MARGIN)) < 1e-12
    assert cnt_a == o["diag"][8] and cnt_s == o["diag"][9], (case, cnt_a, cnt_s, o["diag"][8:10])
    assert cnt_a > 0
    if case == "A_and_S_shift":
        assert cnt_s > 0


def test_dinv_clamp_active(O):
    from sehf import synth
    p = synth.h2o_like(l_cont=8, dr=0.02, r_max=4.0, n_occ=3, seed=1, l_orb_sce=6, allow_closed=True)
    k = 0.6
    E = 0.5 * k * k
    n = 1
    assert (_D(p, n, clamp=False) < W_MIN).sum() > 0          # the clamp acts at this point
    o = O.solve(p, k, dump_sinv=[n])
    n_tr = int(o["diag"][4])
    assert n < n_tr
    S, A, sa, ss = _expected_schur(O, p, n, E, n_tr)
    assert not ss
    B = p.dr ** 2 / 12.0 * O.exchange_Q(p, n)
    N = p.n_psi
    full = np.block([[A, B], [B.T, np.diag(_D(p, n))]])
    top = np.linalg.inv(full)[:N, :N]
    got = o["sinv"][0]
    assert rel_err(got, top) < 1e-11, rel_err(got, top)
    # without the clamp the block differs by far more than the rounding
    full_u = np.block([[A, B], [B.T, np.diag(_D(p, n, clamp=False))]])
    assert rel_err(np.linalg.inv(full_u)[:N, :N], top) > 1e-9


def _riccati_j(l, x):
    return x * spherical_jn(l, x)


def test_analytic_start_closed_form(O):
    from sehf import synth
    from sehf.dipole_io import idx_to_lm
    p = synth.h2o_like(l_cont=8, dr=0.02, r_max=4.0, n_occ=3, seed=1, l_orb_sce=6, allow_closed=True)
    k = 0.9
    E, h = 0.5 * k * k, p.dr
    o0 = O.solve(p, k)
    ns = int(o0["diag"][5])
    assert ns == 5
    idx = list(range(1, ns))
    o = O.solve(p, k, dump_rinv=idx)
    N, Nt = p.n_psi, p.n_tot
    ls = _l_sig(p)
    for i, n in enumerate(idx):
        R = o["rinv"][i]
        assert np.count_nonzero(R - np.diag(np.diag(R))) == 0
        rn, rn1 = (n + 1) * h, (n + 2) * h
        Vn, Vn1 = np.diag(O.potential(p, n)), np.diag(O.potential(p, n + 1))
        want = np.zeros(Nt)
        for mu in range(N):
            l = idx_to_lm(mu)[0]
            Mn = 1.0 + h * h / 6.0 * (E - Vn[mu])
            Mn1 = 1.0 + h * h / 6.0 * (E - Vn1[mu])
            want[mu] = Mn * _riccati_j(l, k * rn) / (Mn1 * _riccati_j(l, k * rn1))
        for f in range(p.n_f):
            l = ls[f]
            Mn = max(1.0 - h * h / 12.0 * l * (l + 1) / rn ** 2, W_MIN)
            Mn1 = max(1.0 - h * h / 12.0 * l * (l + 1) / rn1 ** 2, W_MIN)
            want[N + f] = Mn * rn ** (l + 1) / (Mn1 * rn1 ** (l + 1))
        d = np.diag(R)
        assert np.max(np.abs(d - want) / np.abs(want)) < 1e-12, (n, np.max(np.abs(d - want) / np.abs(want)))


def test_initial_state_read_raw(O):
    from sehf import synth
    p = synth.h2_like(l_cont=2, dr=0.02, r_max=12.0, seed=0)
    r = np.arange(p.Nr) * p.dr
    homo = p.psi_lm[p.n_occ - 1]
    k = 0.7
    base = O.solve(p, k)
    as_rf = O.solve(dataclasses.replace(p, init_psi_lm=r[None, :] * homo), k)
    as_f = O.solve(dataclasses.replace(p, init_psi_lm=homo.copy()), k)
    for key in ("d_raw", "D_raw", "D", "d_corr"):
        assert np.array_equal(base[key], as_rf[key]), key
    assert rel_err(as_f["d_raw"], base["d_raw"]) > 1e-2


def test_lagrange_term_brute_force(O, H):
    """Length-gauge slots (x, y, z) of the reduced dipole from psi' formed
    explicitly; the velocity slots share the Lagrange code and are pinned by
    the gauge identity of the hydrogen case."""
    from sehf import synth
    from sehf.dipole_io import idx_to_lm
    p = synth.h2o_like(l_cont=2, dr=0.02, r_max=10.0, n_occ=2, seed=3)
    k = 0.8
    h, N, no = p.dr, p.n_psi, p.n_occ
    o0 = O.solve(p, k)
    n_tr = int(o0["diag"][4])
    n_pts = n_tr                       # psi kept on [0, n_tr - 1] for the dipole (main.cpp:1267-1270)
    o = O.solve(p, k, dump_psi=list(range(n_pts)))
    psi = o["psi"]                     # [n, mu, beta]
    r_chi = np.arange(n_pts) * h       # the orbitals' grid
    r_sol = (np.arange(n_pts) + 1) * h  # the solver's grid
    nl = p.n_lambda
    phi = np.stack([r_chi[:, None] * p.psi_lm[a, :nl, :n_pts].T for a in range(no)])   # [a, ir, lam]
    chi = phi[no - 1]
    lm_mu = [idx_to_lm(m) for m in range(N)]
    lm_nu = [idx_to_lm(m) for m in range(nl)]

    def simpson(f):
        return O.simpson(np.ascontiguousarray(f), h)

    # b_{beta, alpha} = int phi_alpha . psi_beta
    b = np.zeros((N, no))
    for a in range(no):
        integrand = np.einsum("im,imb->ib", phi[a][:, :N], psi)
        b[:, a] = [simpson(integrand[:, be]) for be in range(N)]
    assert rel_err(b, o["b_overlap"]) < 1e-12
    # psi' = psi - 2 sum_alpha b phi_alpha, on every radial point
    psi_p = psi - 2.0 * np.einsum("ba,aim->imb", b, phi[:, :, :N])
    Minv = o["A"] - 1j * o["B"]
    M = np.linalg.inv(Minv).conj().T
    worst, lag = 0.0, 0.0
    for slot, q in enumerate((+1, -1, 0)):
        G = np.zeros((N, nl))
        for mu, (l1, m1) in enumerate(lm_mu):
            for nu, (l2, m2) in enumerate(lm_nu):
                if abs(l1 - l2) == 1:
                    G[mu, nu] = math.sqrt(4.0 * math.pi / 3.0) * H.sehf_host_real_gaunt(l1, m1, 1, q, l2, m2)
        Xi = chi @ G.T                                    # [ir, mu]
        integrand = r_sol[:, None] * np.einsum("im,imb->ib", Xi, psi_p)
        d_p = np.array([simpson(integrand[:, be]) for be in range(N)])
        D_brute = M @ d_p
        worst = max(worst, rel_err(o["D"][slot], D_brute))
        lag = max(lag, rel_err(o["D"][slot], o["D_raw"][slot]))
        assert rel_err(o["D_raw"][slot], M @ o["d_raw"][slot]) < 1e-12
    assert lag > 1e-3, lag                 # the Lagrange term is not negligible here
    assert worst < 1e-11, worst
