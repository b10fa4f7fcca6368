"""CPU pins of the oracle's tangent (forward-mode derivative along E and the
polarisation strength alpha) by what the mathematics fixes: the derivative of
every output must equal the limit of central finite differences of the
oracle's own forward outputs.

* Richardson-extrapolated central differences (steps h and h/2) of K, A, B,
  d_raw, b_overlap and the reduced dipole D agree with the tangent to
  max(3e-7, 3 x the tangent's spread between two rounding orders) relative;
  at k = 0.7 that is below 1e-6 (the extrapolated differences sit at ~1e-7,
  and the mutations below move the outputs by 2e-6 or more), at k = 1.1 the
  case's conditioning puts dD's rounding floor near 1e-6.  The smallest terms
  of the tangent are O(h^2/6) ~ 7e-5 relative: dropping the boundary source
  h^2/6 w_N psi_N moves A, B, d_raw and b_overlap by ~5e-6 in the E case
  (D, normalised by the fit, absorbs it), and a weight taken one radial
  index off in the -h^2/6 w Sinv psi term moves D by 2e-6 in the alpha case;
  both fail here.
* alpha enters with a radius-dependent weight (w_n = -dV_pol/dalpha at
  potential index n), so an index slip between the Schur, ratio and
  back-propagation steps shows in the alpha case even though it cancels for
  E, whose weight is constant.
* The tangent is as stable as the forward outputs: a second rounding order
  (Gauss-Jordan in place of LU) moves dD by far less than 1e-6 relative on a
  case whose ratio matrices pass close to a pole (|Rinv| ~ 7e2 at n = 663).
* The Wigner delay from the analytic dD/dE agrees with the finite-difference
  delay in k (the reference's post-processing definition) in the limit of a
  refined k grid.
"""
import dataclasses

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


def _problem(alpha=9.8):
    from sehf import synth
    return synth.h2o_like(l_cont=3, dr=0.02, r_max=14.0, n_occ=3, seed=1, l_orb_sce=3, pol=(alpha, 1.5),
                          allow_closed=True)


KEYS = (("K_kme", "dK"), ("A", "dA"), ("B", "dB"), ("d_raw", "dd_raw"), ("b_overlap", "db_overlap"), ("D", "dD"))


def _richardson(f, x0, h):
    """(4 D(h/2) - D(h)) / 3 of the central difference D."""
    fp, fm, fp2, fm2 = f(x0 + h), f(x0 - h), f(x0 + h / 2), f(x0 - h / 2)
    out = {}
    for key, _ in KEYS:
        d1 = (fp[key] - fm[key]) / (2 * h)
        d2 = (fp2[key] - fm2[key]) / h
        out[key] = (4 * d2 - d1) / 3
    return out


def _tolerances(O, p, k, tangent):
    """max(3e-7, 3 x the spread of the tangent between two rounding orders):
    the tangent's own rounding floor depends on the case's conditioning."""
    L = O.lib()
    L.sehf_oracle_set_inverse_mode(1)
    try:
        b = O.solve(p, k, tangent=tangent)
    finally:
        L.sehf_oracle_set_inverse_mode(0)
    a = O.solve(p, k, tangent=tangent)
    return a, {dk: max(3e-7, 3.0 * np.abs(a[dk] - b[dk]).max() / np.abs(a[dk]).max()) for _, dk in KEYS}


@pytest.mark.parametrize("k", [0.7, 1.1])
def test_energy_tangent_matches_finite_differences(O, k):
    p = _problem()
    E = 0.5 * k * k
    o, tol = _tolerances(O, p, k, (1.0, 0.0))
    if k == 0.7:
        assert max(tol.values()) < 1e-6           # the tight case keeps the mutation sensitivity
    fd = _richardson(lambda e: O.solve(p, np.sqrt(2 * e)), E, 2e-4)
    for key, dkey in KEYS:
        an = o[dkey]
        err = np.abs(fd[key] - an).max() / np.abs(an).max()
        assert err < tol[dkey], (key, err, tol[dkey])


def test_alpha_tangent_matches_finite_differences(O):
    p = _problem()
    k = 0.7
    o, tol = _tolerances(O, p, k, (0.0, 1.0))
    assert max(tol.values()) < 1e-6
    fd = _richardson(lambda a: O.solve(dataclasses.replace(p, pol_alpha=a), k), 9.8, 0.04)
    for key, dkey in KEYS:
        an = o[dkey]
        err = np.abs(fd[key] - an).max() / np.abs(an).max()
        assert err < tol[dkey], (key, err, tol[dkey])


def test_tangent_linear_in_direction(O):
    p = _problem()
    k = 0.9
    a = O.solve(p, k, tangent=(1.0, 0.0))["dD"]
    b = O.solve(p, k, tangent=(0.0, 1.0))["dD"]
    c = O.solve(p, k, tangent=(0.5, -2.0))["dD"]
    # exact in exact arithmetic; the bound is the tangent's rounding floor on
    # this case (the LU / Gauss-Jordan spread of dD, ~5e-8)
    assert np.abs(c - (0.5 * a - 2.0 * b)).max() < 1e-7 * np.abs(c).max()


def test_tangent_rounding_stable_through_near_pole(O):
    p = _problem()
    k = 0.7
    a = O.solve(p, k, tangent=(1.0, 0.0), dump_rinv=[663])
    assert np.abs(a["rinv"][0]).max() > 100.0          # the case keeps its near-pole step
    L = O.lib()
    L.sehf_oracle_set_inverse_mode(1)
    try:
        b = O.solve(p, k, tangent=(1.0, 0.0))
    finally:
        L.sehf_oracle_set_inverse_mode(0)
    spread = np.abs(a["dD"] - b["dD"]).max() / np.abs(a["dD"]).max()
    assert spread < 1e-6, spread


def test_analytic_delay_matches_refined_fd_in_k(O):
    """tau = Im sum d* dd/dE / sum |d|^2 (cross_section_delay.py:227-263):
    the analytic dd/dE against central differences in k at two spacings,
    extrapolated (the FD-in-k delay converges at second order)."""
    p = _problem()
    k = 0.8
    o = O.solve(p, k, tangent=(1.0, 0.0))
    D, dDE = o["D"], o["dD"]

    def tau(D, dD):
        return (np.conj(D) * dD).sum(axis=1).imag / (np.abs(D) ** 2).sum(axis=1)

    t_an = tau(D, dDE)

    def fd(dk):
        Dp, Dm = O.solve(p, k + dk)["D"], O.solve(p, k - dk)["D"]
        return tau(D, (Dp - Dm) / (2 * dk) / k)           # dd/dE = (dd/dk) / k

    t1, t2 = fd(2e-3), fd(1e-3)
    t_fd = (4 * t2 - t1) / 3
    assert np.abs(t_fd - t_an).max() < 1e-6 * max(1.0, np.abs(t_an).max()), (t_an, t_fd)
