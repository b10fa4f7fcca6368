"""CPU pins of the oracle and of the product's host numerics (no GPU).

Each check is chosen so that a plausible mistake (sign, index, dropped term,
transposed operand) fails it:
* Wigner 3j vs sympy's exact rationals (tests/golden/golden.npz);
* real Gaunt vs the 20 hand-curated values of the reference's
  src/tests/test_real_gaunt_dipole.cpp:44-68, vs G(00;00;00)=1/sqrt(4pi),
  and vs exact Gauss-Legendre x trapezoid quadrature of three real Y;
* Riccati-Bessel/-Neumann vs scipy and the closed forms j0=sin x,
  j1=sin x/x - cos x, y0=-cos x (src/tests/test_freeparticle_dipole.cpp:55-67);
* Simpson weights exact on cubics, 5-point derivative exact on quartics;
* dense inverse / smallest eigenvalue vs numpy;
* the potential matrix vs an independent quadrature assembly, the exchange
  coupling vs the brute-force triple loop (ExchangeCoupling.cpp:153-185) and
  the mu=0 identity Q[0,(i,s)] = sqrt2 chi_s / r (test_exchange_coupling.cpp:142).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from conftest import ROOT, rel_err

GOLD = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))

# (l_mu, m_mu, q, l_nu, m_nu, G^R)  -- src/tests/test_real_gaunt_dipole.cpp:44-68
REF_GAUNT = [
    (0, 0, -1, 1, -1, +0.282094791774), (1, -1, -1, 0, 0, +0.282094791774),
    (1, -1, -1, 2, 0, -0.126156626101), (1, -1, -1, 2, 2, -0.218509686118),
    (1, 0, -1, 2, -1, +0.218509686118), (1, 1, -1, 2, -2, +0.218509686118),
    (2, -2, -1, 1, 1, +0.218509686118), (2, -1, -1, 1, 0, +0.218509686118),
    (2, 0, -1, 1, -1, -0.126156626101), (2, 2, -1, 1, -1, -0.218509686118),
    (0, 0, 0, 1, 0, +0.282094791774), (1, -1, 0, 2, -1, +0.218509686118),
    (1, 0, 0, 2, 0, +0.252313252202), (1, 1, 0, 2, 1, +0.218509686118),
    (0, 0, +1, 1, 1, +0.282094791774), (1, -1, +1, 2, -2, +0.218509686118),
    (1, 0, +1, 2, 1, +0.218509686118), (1, 1, +1, 0, 0, +0.282094791774),
    (1, 1, +1, 2, 0, -0.126156626101), (1, 1, +1, 2, 2, +0.218509686118),
]


@pytest.fixture(scope="module")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="module")
def H():
    from sehf import _lib
    return _lib.load()


def test_w3j_vs_sympy(O):
    w = GOLD["w3j"]
    L = O.lib()
    got = np.array([L.sehf_oracle_w3j(*[int(v) for v in row[:6]]) for row in w])
    assert np.max(np.abs(got - w[:, 6])) < 1e-15


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_gaunt_reference_table(O, H, which):
    f = O.lib().sehf_oracle_gaunt_real if which == "oracle" else H.sehf_host_real_gaunt
    for (lm, mm, q, ln, mn, g) in REF_GAUNT:
        assert abs(f(lm, mm, 1, q, ln, mn) - g) < 1e-9
    assert abs(f(0, 0, 0, 0, 0, 0) - 1.0 / math.sqrt(4 * math.pi)) < 1e-15


def test_gaunt_vs_quadrature(O, H):
    from sehf.synth import angular_grid, real_ylm_table
    lmax = 8
    th, ph, w = angular_grid(3 * lmax)
    Y = real_ylm_table(lmax, th, ph)
    rng = np.random.default_rng(3)
    worst_o = worst_p = 0.0
    for _ in range(400):
        i1, i2, i3 = rng.integers(0, (lmax + 1) ** 2, size=3)
        l1, l2, l3 = (int(math.isqrt(i)) for i in (i1, i2, i3))
        m1, m2, m3 = i1 - l1 * l1 - l1, i2 - l2 * l2 - l2, i3 - l3 * l3 - l3
        quad = float(np.sum(w * Y[i1] * Y[i2] * Y[i3]))
        worst_o = max(worst_o, abs(O.lib().sehf_oracle_gaunt_real(l1, int(m1), l2, int(m2), l3, int(m3)) - quad))
        worst_p = max(worst_p, abs(H.sehf_host_real_gaunt(l1, int(m1), l2, int(m2), l3, int(m3)) - quad))
    assert worst_o < 1e-13 and worst_p < 1e-13


def test_gaunt_high_l_product_vs_oracle(O, H):
    """The product's Gauss-Legendre Gaunt table against the oracle's Racah
    sums (pinned above against sympy) at the l of the l_cont = 30 scan:
    l_mu, l_mu' <= 30, l_sigma <= 60, m triplets that pass the real-harmonic
    selection rule |m_s| in {|m1|+|m2|, ||m1|-|m2||}."""
    rng = np.random.default_rng(11)
    worst = 0.0
    n_nz = 0
    for _ in range(3000):
        l1, l2 = (int(x) for x in rng.integers(0, 31, size=2))
        m1 = int(rng.integers(-l1, l1 + 1))
        m2 = int(rng.integers(-l2, l2 + 1))
        l3 = int(rng.integers(abs(l1 - l2), l1 + l2 + 1))
        am = abs(m1) + abs(m2) if rng.random() < 0.5 else abs(abs(m1) - abs(m2))
        if am > l3:
            continue
        m3 = am if rng.random() < 0.5 else -am
        o = O.lib().sehf_oracle_gaunt_real(l1, m1, l3, m3, l2, m2)
        p = H.sehf_host_real_gaunt(l1, m1, l3, m3, l2, m2)
        worst = max(worst, abs(o - p))
        n_nz += abs(o) > 1e-15
    assert n_nz > 500
    assert worst < 2e-14, worst


@pytest.mark.parametrize("which", ["oracle", "product"])
def test_riccati_bessel(O, H, which):
    L = O.lib()
    fj = L.sehf_oracle_riccati_j if which == "oracle" else H.sehf_host_ric_j
    fy = L.sehf_oracle_riccati_y if which == "oracle" else H.sehf_host_ric_y
    xs, ls = GOLD["bessel_x"], GOLD["bessel_l"]
    J, Y = GOLD["ric_j"], GOLD["ric_y"]
    for a, l in enumerate(ls):
        for b, x in enumerate(xs):
            gj, gy = fj(int(l), float(x)), fy(int(l), float(x))
            if J[a, b] != 0.0:
                assert abs(gj - J[a, b]) <= 2e-13 * abs(J[a, b]) + 1e-300, (l, x, gj, J[a, b])
            if np.isfinite(Y[a, b]) and abs(Y[a, b]) < 1e300:
                assert abs(gy - Y[a, b]) <= 2e-13 * abs(Y[a, b]), (l, x, gy, Y[a, b])
    for x in (0.3, 1.7, 12.5, 77.0):   # closed forms
        assert abs(fj(0, x) - math.sin(x)) < 1e-15
        assert abs(fj(1, x) - (math.sin(x) / x - math.cos(x))) < 1e-14
        assert abs(fy(0, x) + math.cos(x)) < 1e-15


def test_simpson_and_derivative(O):
    for n in (5, 6, 7, 8, 101, 102):
        h = 0.013
        r = np.arange(n) * h
        for poly in (r ** 3 - 2 * r ** 2 + r, 4 * r ** 3):
            exact = np.polyval(np.polyint(np.polyfit(r, poly, 3)), r[-1])
            assert abs(O.simpson(poly, h) - exact) < 1e-13
        f = r ** 4 - 3 * r ** 2
        assert np.max(np.abs(O.five_point_derivative(f, h) - (4 * r ** 3 - 6 * r))) < 1e-10


def test_dense_primitives(O, H):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((70, 70))
    assert rel_err(O.inverse_general(A), np.linalg.inv(A)) < 1e-12
    S = A + A.T
    ex = np.linalg.eigvalsh(S)[0]
    assert abs(O.min_eig_sym(S) - ex) < 1e-12
    Sf = np.asfortranarray(S)
    assert abs(H.sehf_host_min_eig(70, Sf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) - ex) < 1e-12


def _exact_inverse(A):
    """Inverse of the FP64 matrix A in exact rational arithmetic."""
    from fractions import Fraction
    n = A.shape[0]
    M = [[Fraction(float(A[i, j])) for j in range(n)] + [Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for k in range(n):
        p = next(i for i in range(k, n) if M[i][k] != 0)
        M[k], M[p] = M[p], M[k]
        piv = M[k][k]
        M[k] = [v / piv for v in M[k]]
        for i in range(n):
            if i != k and M[i][k] != 0:
                f = M[i][k]
                M[i] = [a - f * b for a, b in zip(M[i], M[k])]
    return np.array([[float(M[i][n + j]) for j in range(n)] for i in range(n)])


def test_rounded_inverse_mode(O):
    """Rounding-probe bit 5 (the referee of the device's K-family outputs):
    the 80-bit LU inverse rounded once to FP64 is within 1e-13 of the exact
    inverse of an ill-conditioned FP64 matrix (Hilbert, order 7, condition
    ~5e8; measured 7e-14), where the FP64 LU is off by 6e-10."""
    n = 7
    Hm = np.array([[1.0 / (i + j + 1) for j in range(n)] for i in range(n)])
    X = _exact_inverse(Hm)
    lu = O.inverse_general(Hm)
    L = O.lib()
    L.sehf_oracle_set_inverse_mode(32)
    try:
        ext = O.inverse_general(Hm)
    finally:
        L.sehf_oracle_set_inverse_mode(0)
    e_lu, e_ext = rel_err(lu, X), rel_err(ext, X)
    print(f"FP64 LU {e_lu:.1e}, rounded {e_ext:.1e}")
    assert e_ext < 1e-13 and e_lu > 100 * e_ext


@pytest.fixture(scope="module")
def small():
    from sehf import synth
    return synth.h2o_like(l_cont=3, dr=0.02, r_max=8.0, n_occ=2, seed=1)


def _host_mat(H, p, fn, ir, shape):
    from sehf.scan import _c_problem
    keep = []
    cp = _c_problem(p, keep)
    M = np.zeros(shape, order="F")
    assert getattr(H, fn)(ctypes.byref(cp), ir, M.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return M


def test_potential_vs_quadrature(O, H, small):
    """V_{mu nu}(r) = int Y_mu (sum_s V_s(r) Y_s) Y_nu dOmega + diagonal terms."""
    from sehf.synth import angular_grid, real_ylm_table
    p = small
    lc = p.l_cont
    th, ph, w = angular_grid(4 * lc)
    Ys = real_ylm_table(2 * lc, th, ph)
    Yc = Ys[: (lc + 1) ** 2]
    for ir in (3, 40, 250):
        r = ir * p.dr
        Vo = O.potential(p, ir)
        Vh = _host_mat(H, p, "sehf_host_potential", ir, (p.n_psi, p.n_psi))
        # V_sigma(r): V_H + nuclear multipoles (Potentials.cpp:40-88)
        vs = p.V_H[: (2 * lc + 1) ** 2, ir].copy()
        for Z, pos in zip(p.Z, p.xyz):
            R = float(np.linalg.norm(pos))
            thR = math.acos(pos[2] / R) if R > 1e-14 else 0.0
            phR = math.atan2(pos[1], pos[0]) if R > 1e-14 else 0.0
            YR = real_ylm_table(2 * lc, np.array([thR]), np.array([phR]))[:, 0]
            for l in range(2 * lc + 1):
                rad = 4 * math.pi / (2 * l + 1) * min(r, R) ** l / max(r, R) ** (l + 1)
                for m in range(-l, l + 1):
                    yv = YR[l * l + l + m] if R > 1e-14 else (1 / (2 * math.sqrt(math.pi)) if l == 0 else 0.0)
                    if abs(yv) > 1e-15:
                        vs[l * l + l + m] += -Z * rad * yv
        vq = vs @ Ys
        Vq = (Yc * (w * vq)) @ Yc.T
        ls = np.array([int(math.isqrt(i)) for i in range(p.n_psi)])
        Vq += np.diag(ls * (ls + 1) * 0.5 / r ** 2)
        assert rel_err(Vo, Vq) < 1e-12 and rel_err(Vh, Vq) < 1e-12
        assert np.array_equal(Vo, Vo.T) and np.array_equal(Vh, Vh.T)


def test_exchange_vs_bruteforce(O, H, small):
    p = small
    nl = (p.l_cont + p.l_exch + 1) ** 2
    L = O.lib()
    trip = []
    for mu in range(p.n_psi):
        lm = int(math.isqrt(mu)); mm = mu - lm * lm - lm
        for lam in range(nl):
            ll = int(math.isqrt(lam)); ml = lam - ll * ll - ll
            for s in range(p.n_sigma):
                ls = int(math.isqrt(s)); ms = s - ls * ls - ls
                g = L.sehf_oracle_gaunt_real(lm, mm, ll, ml, ls, ms)
                if abs(g) > 1e-14:
                    trip.append((mu, lam, s, g))
    for ir in (1, 30, 200):
        r = ir * p.dr
        chi = r * p.psi_lm[: p.n_occ, :nl, ir]
        Qb = np.zeros((p.n_psi, p.n_f))
        for (mu, lam, s, g) in trip:
            for i in range(p.n_occ):
                Qb[mu, i * p.n_sigma + s] += 2 * math.sqrt(2 * math.pi) / r * g * chi[i, lam]
        Qo = O.exchange_Q(p, ir)
        Bh = _host_mat(H, p, "sehf_host_exchange_B", ir, (p.n_psi, p.n_f)) / (p.dr ** 2 / 12)
        assert rel_err(Qo, Qb) < 1e-13 and rel_err(Bh, Qb) < 1e-13
        for i in range(p.n_occ):   # mu = 0 identity
            np.testing.assert_allclose(Qo[0, i * p.n_sigma:(i + 1) * p.n_sigma],
                                       math.sqrt(2) * chi[i, :p.n_sigma] / r, rtol=1e-13, atol=1e-300)


def test_abi_exports_every_declared_symbol(H):
    import re
    hdr = open(os.path.join(ROOT, "include", "sehf.h")).read()
    names = set(re.findall(r"\b(sehf_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(H, n)]
    assert not missing, missing
