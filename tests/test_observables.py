"""Observables (cross_section_delay.py counterpart) in torch FP64 against the
numpy/scipy restatement, plus what the formulas themselves fix:

* atomic s -> p amplitudes d^q on the l = 1 channel along q give beta = 2
  (cross_section_delay.py:383-384), and beta stays in [-1, 2] for random
  amplitudes;
* beta is invariant under a rotation of the molecular frame about z (the
  real harmonics and the x/y polarisation slots rotate together);
* gauge equivalence d^V = omega d^L gives sigma_L = sigma_V, Q = 1, dphi = 0;
* numpy.gradient's rule is exact for quadratics in the interior and for
  lines at the two ends, so tau of a quadratic D(k) is known in closed form.
"""
import math

import numpy as np
import pytest
import torch


@pytest.fixture(scope="module")
def R():
    import observables_ref
    return observables_ref


@pytest.fixture(scope="module")
def P():
    from sehf import observables
    return observables


def _rand_D(rng, nE, nch):
    return rng.normal(size=(nE, nch)) + 1j * rng.normal(size=(nE, nch))


def test_product_matches_restatement(R, P):
    rng = np.random.default_rng(5)
    lmax = 6
    nch = (lmax + 1) ** 2
    k = np.sort(rng.uniform(0.2, 2.0, size=9))
    om = 0.5 * k * k + 0.7
    DL, DV = _rand_D(rng, 9, nch), _rand_D(rng, 9, nch)
    for g, D in (("length", DL), ("velocity", DV)):
        a = P.sigma(k, D, om, g).numpy()
        b = R.sigma_from_dipole(k, D, om, g)
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    assert np.abs(P.time_delay(k, DL).numpy() - R.time_delay(k, DL)).max() < 1e-12
    assert np.abs(P.q_sigma(DL, DV, om).numpy() - R.q_sigma(DL, DV, om)).max() < 1e-12
    assert np.abs(P.dphi_rms(DL, DV).numpy() - R.dphi_rms(DL, DV)).max() < 1e-12
    Dx, Dy, Dz = _rand_D(rng, 4, nch), _rand_D(rng, 4, nch), _rand_D(rng, 4, nch)
    a = P.beta(Dx, Dy, Dz).numpy()
    b = R.compute_beta(Dx, Dy, Dz, list(range(nch)))
    assert np.abs(a - b).max() < 1e-12, (a, b)
    assert np.all(a >= -1.0 - 1e-12) and np.all(a <= 2.0 + 1e-12)


def test_beta_atomic_s_to_p(R, P):
    nch = 9
    Dx, Dy, Dz = (np.zeros((1, nch), complex) for _ in range(3))
    # Y_{1,+1} ~ x, Y_{1,-1} ~ y, Y_{1,0} ~ z (mu = 3, 1, 2)
    Dx[0, 3] = Dy[0, 1] = Dz[0, 2] = 0.7 - 0.2j
    assert abs(P.beta(Dx, Dy, Dz).item() - 2.0) < 1e-12
    assert abs(R.compute_beta(Dx, Dy, Dz, list(range(nch)))[0] - 2.0) < 1e-12


def _rotz(d, alpha, lmax):
    """Rotate channel amplitudes: Y_{l,m} (cos m phi), Y_{l,-m} (sin m phi)."""
    out = d.copy()
    for l in range(lmax + 1):
        for m in range(1, l + 1):
            c, s = math.cos(m * alpha), math.sin(m * alpha)
            ip, im = l * l + l + m, l * l + l - m
            a, b = d[..., ip].copy(), d[..., im].copy()
            out[..., ip] = c * a - s * b
            out[..., im] = s * a + c * b
    return out


def test_beta_rotation_invariance(P):
    rng = np.random.default_rng(9)
    lmax = 5
    nch = (lmax + 1) ** 2
    Dx, Dy, Dz = _rand_D(rng, 3, nch), _rand_D(rng, 3, nch), _rand_D(rng, 3, nch)
    b0 = P.beta(Dx, Dy, Dz).numpy()
    al = 0.61
    Rx, Ry, Rz = (_rotz(D, al, lmax) for D in (Dx, Dy, Dz))
    # the polarisation vector rotates with the frame
    Ex = math.cos(al) * Rx - math.sin(al) * Ry
    Ey = math.sin(al) * Rx + math.cos(al) * Ry
    b1 = P.beta(Ex, Ey, Rz).numpy()
    assert np.abs(b0 - b1).max() < 1e-11, (b0, b1)
    # and a wrong pairing (frame rotated, polarisation not) changes it
    assert np.abs(b0 - P.beta(Rx, Ry, Rz).numpy()).max() > 1e-3


def test_gauge_equivalence(P):
    rng = np.random.default_rng(2)
    k = np.linspace(0.3, 1.5, 7)
    om = 0.5 * k * k + 0.5
    DL = _rand_D(rng, 7, 16)
    DV = om[:, None] * DL
    sL = P.sigma(k, DL, om, "length").numpy()
    sV = P.sigma(k, DV, om, "velocity").numpy()
    assert np.abs(sL - sV).max() < 1e-13 * sL.max()
    assert np.abs(P.q_sigma(DL, DV, om).numpy() - 1.0).max() < 1e-13
    assert np.abs(P.dphi_rms(DL, DV).numpy()).max() < 1e-7


def test_time_delay_quadratic(P):
    k = np.array([0.4, 0.45, 0.53, 0.6, 0.72, 0.8])
    rng = np.random.default_rng(3)
    a, b, c = _rand_D(rng, 3, 4)
    D = a[None] + b[None] * k[:, None] + c[None] * k[:, None] ** 2
    dDk = b[None] + 2 * c[None] * k[:, None]
    dDk[0] = (D[1] - D[0]) / (k[1] - k[0])       # first-order ends
    dDk[-1] = (D[-1] - D[-2]) / (k[-1] - k[-2])
    ref = np.sum(np.conj(D) * dDk / k[:, None], axis=1).imag / np.sum(np.abs(D) ** 2, axis=1)
    assert np.abs(P.time_delay(k, D).numpy() - ref).max() < 1e-12
    assert np.abs(P.time_delay(k, D, with_minus=True).numpy() + ref).max() < 1e-12
