"""GPU checks of the differentiable dipoles (sehf.autograd.DeviceDipoles):
gradients of the cross section with respect to the photoelectron energy and
the polarisation strength against central finite differences of the device
scan itself (h_E = 1e-5 Ha, h_alpha = 1e-4), and the Wigner delay from the
analytic dd/dE against the reference's finite-difference-in-k delay
(postprocessing/cross_section_delay.py:227-263, numpy.gradient over k) on a
refined k grid, extrapolated to zero spacing.

Case: the small H2O-like molecule with the soft polarisation potential
-alpha / (2 (r^2 + r_c^2)^2), alpha = 9.8, r_c = 1.5 (the adjoint
configuration's model potential), l_cont = 3, 701 radial points.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    from sehf import synth
    from sehf.autograd import DeviceDipoles
    from sehf.steppers import GpuContext
    p = synth.h2o_like(l_cont=3, dr=0.02, r_max=14.0, n_occ=3, seed=1, l_orb_sce=3, pol=(9.8, 1.5),
                       allow_closed=True)
    return p, DeviceDipoles(p, GpuContext(True))


def _sigma_L(E, D, E_init):
    from sehf import observables as ob
    k = torch.sqrt(2.0 * E)
    return ob.sigma(k, D[:, 0:3, :].reshape(len(E), -1), E - E_init, "length")


def test_sigma_gradients_vs_central_differences(setup):
    """Autograd of sigma_L through the device tangent against central
    differences of the device scan.  Two step sets:
    * the adjoint configuration's steps h_E = 1e-5, h_alpha = 1e-4: there the
      difference quotient carries the forward outputs' rounding noise
      (~1e-12 relative, uncorrelated between the two alpha set-ups) divided
      by 2h, which is ~1e-7 relative for E and ~1e-5 for the small
      d sigma / d alpha; the bounds are 2e-6 and 1e-4;
    * Richardson-extrapolated steps (h, h/2) = (1e-3, 5e-4) in E and
      (0.1, 0.05) in alpha, whose truncation (O(h^4) after extrapolation) and
      noise are both below 1e-7: bound 1e-6."""
    p, dev = setup
    E_init = float(p.orb_energies[p.n_occ - 1])
    E0 = torch.tensor([0.245, 0.405, 0.605], dtype=torch.float64)
    E = E0.clone().requires_grad_(True)
    alpha = torch.tensor(p.pol_alpha, dtype=torch.float64, requires_grad=True)
    sig = _sigma_L(E, dev(E, alpha), E_init)
    sig.sum().backward()

    def s_of(Ev, av):
        with torch.no_grad():
            return _sigma_L(Ev, dev(Ev, torch.tensor(av, dtype=torch.float64)), E_init)

    def cdE(h):
        return (s_of(E0 + h, p.pol_alpha) - s_of(E0 - h, p.pol_alpha)) / (2 * h)

    def cda(h):
        return ((s_of(E0, p.pol_alpha + h) - s_of(E0, p.pol_alpha - h)) / (2 * h)).sum()

    def rel(a, b):
        return float((a - b).abs().max() / b.abs().max())

    fdE, fda = cdE(1e-5), cda(1e-4)
    rE = (4 * cdE(5e-4) - cdE(1e-3)) / 3
    ra = (4 * cda(0.05) - cda(0.1)) / 3
    print(f"dsigma/dE {E.grad.tolist()} fd(1e-5) rel {rel(E.grad, fdE):.1e} richardson rel {rel(E.grad, rE):.1e}; "
          f"dsigma/dalpha {float(alpha.grad):.10e} fd(1e-4) {float(fda):.10e} rel {rel(alpha.grad, fda):.1e} "
          f"richardson {float(ra):.10e} rel {rel(alpha.grad, ra):.1e}")
    assert rel(E.grad, fdE) < 2e-6
    assert rel(alpha.grad, fda) < 1e-4
    assert rel(E.grad, rE) < 1e-6
    assert rel(alpha.grad, ra) < 1e-6


def test_analytic_delay_vs_refined_fd_in_k(setup):
    from sehf import observables as ob
    p, dev = setup
    k0 = np.array([0.7, 0.9, 1.1])
    D, dD = dev.with_energy_derivative(0.5 * k0 ** 2)
    tau = ob.time_delay_analytic(D[:, 0:3].reshape(3, -1), dD[:, 0:3].reshape(3, -1))   # length gauge

    def tau_fd(dk):
        # numpy.gradient over the k grid (k - dk, k, k + dk) at its midpoint
        taus = []
        for k in k0:
            ks = np.array([k - dk, k, k + dk])
            Dk = dev(torch.as_tensor(0.5 * ks ** 2))[:, 0:3].reshape(3, -1)
            taus.append(float(ob.time_delay(torch.as_tensor(ks), Dk)[1]))
        return np.array(taus)

    # dk = 0.005 and the grid refined 4x; the extrapolation removes the
    # O(dk^2) term, leaving O(dk^4) ~ 1e-7 here (at dk = 0.01 the dk^4 term
    # is still ~2e-6 at k = 0.9)
    dk = 0.005
    t1, t2 = tau_fd(dk), tau_fd(dk / 4)
    t_conv = (16 * t2 - t1) / 15
    print("tau analytic", tau.tolist(), "fd(dk/4)", t2.tolist(), "extrapolated", t_conv.tolist(),
          "|analytic - extrapolated|", np.abs(tau.numpy() - t_conv).tolist())
    assert np.abs(tau.numpy() - t_conv).max() < 1e-6           # atomic units
    # the refined grid alone is already within its O(dk^2) truncation
    assert np.abs(tau.numpy() - t2).max() < 1e-3 * max(1.0, float(np.abs(t_conv).max()))
