"""CPU checks of the torch.autograd plumbing of sehf.autograd.DeviceDipoles.

The device scan is replaced by a stand-in with the same run() contract whose
dipoles are a closed-form function d(E, alpha) with known derivatives; the
gradients that flow through DeviceDipoles must then equal torch's own
autograd of the same closed form (complex outputs, real inputs: torch's
convention grad_x = Re sum conj(grad_d) dd/dx is what is being pinned), for
losses that mix |d|^2 (cross sections), Im d* dd (phases) and plain linear
forms.
"""
import dataclasses

import numpy as np
import pytest
import torch


@dataclasses.dataclass
class _P:
    n_psi: int = 4
    n_occ: int = 1
    pol_kind: int = 2
    pol_alpha: float = 1.3


_RNG = np.random.default_rng(7)
_C0 = _RNG.normal(size=(6, 4)) + 1j * _RNG.normal(size=(6, 4))
_C1 = _RNG.normal(size=(6, 4)) + 1j * _RNG.normal(size=(6, 4))


def _closed_form(E, a, xp):
    """d = C0 exp(i a E) + C1 E^2 a (broadcast over energies)."""
    if xp is torch:
        C0 = torch.as_tensor(_C0)
        C1 = torch.as_tensor(_C1)
        ph = torch.exp(1j * (a * E)).reshape(-1, 1, 1)
        return C0 * ph + C1 * (E * E * a).reshape(-1, 1, 1)
    ph = np.exp(1j * a * E)[:, None, None]
    return _C0 * ph + _C1 * (E * E * a)[:, None, None]


class _FakeScan:
    calls = []

    def __init__(self, p, ctx=None):
        self.p = p

    def run(self, ks, batch=0, tangent=None):
        ks = np.asarray(ks, dtype=np.float64)
        E = 0.5 * ks * ks
        a = self.p.pol_alpha
        D = _closed_form(E, a, np)
        _FakeScan.calls.append((len(ks), tangent is not None))
        out = []
        for e in range(len(ks)):
            r = {"k": ks[e], "D": D[e]}
            if tangent is not None:
                tE, tA = np.asarray(tangent)[e]
                dDE = _C0 * 1j * a * np.exp(1j * a * E[e]) + _C1 * 2 * E[e] * a
                dDa = _C0 * 1j * E[e] * np.exp(1j * a * E[e]) + _C1 * E[e] ** 2
                r["dD"] = tE * dDE + tA * dDa
            out.append(r)
        return out


def _losses(D):
    s = (D.abs() ** 2).sum()
    ph = (D[:, :, 1:].conj() * D[:, :, :-1]).imag.sum()
    lin = (D.real * 0.3 - D.imag * 1.7).sum()
    return s + 0.5 * ph + lin


def test_gradients_match_torch_autograd_of_the_closed_form():
    from sehf.autograd import DeviceDipoles
    E0 = torch.tensor([0.2, 0.55, 1.1], dtype=torch.float64)
    a0 = 1.3
    dev = DeviceDipoles(_P(pol_alpha=a0), scan_cls=_FakeScan)
    E = E0.clone().requires_grad_(True)
    alpha = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
    _losses(dev(E, alpha)).backward()
    E2 = E0.clone().requires_grad_(True)
    a2 = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
    _losses(_closed_form(E2, a2, torch)).backward()
    assert torch.allclose(E.grad, E2.grad, rtol=1e-12, atol=1e-12), (E.grad, E2.grad)
    assert torch.allclose(alpha.grad, a2.grad, rtol=1e-12, atol=1e-12), (alpha.grad, a2.grad)


def test_only_requested_directions_run():
    from sehf.autograd import DeviceDipoles
    _FakeScan.calls.clear()
    dev = DeviceDipoles(_P(), scan_cls=_FakeScan)
    E = torch.tensor([0.3, 0.4], dtype=torch.float64)
    dev(E)                                        # no gradient: plain scan
    dev(E.clone().requires_grad_(True))           # E only: one direction per energy
    a = torch.tensor(1.3, dtype=torch.float64, requires_grad=True)
    dev(E.clone().requires_grad_(True), a)        # both: every energy twice
    assert _FakeScan.calls == [(2, False), (2, True), (4, True)]


def test_alpha_other_than_the_problem_builds_a_new_scan():
    from sehf.autograd import DeviceDipoles
    dev = DeviceDipoles(_P(pol_alpha=1.3), scan_cls=_FakeScan)
    E = torch.tensor([0.3], dtype=torch.float64)
    D = dev(E, torch.tensor(2.0, dtype=torch.float64))
    ref = _closed_form(E, torch.tensor(2.0, dtype=torch.float64), torch)
    assert torch.allclose(D, ref, rtol=1e-14, atol=0)
    assert sorted(dev._scans) == [2.0]
    with pytest.raises(ValueError):
        DeviceDipoles(_P(pol_kind=0, pol_alpha=0.0), scan_cls=_FakeScan)(E, torch.tensor(1.0, dtype=torch.float64))


def test_energy_derivative_helper():
    from sehf.autograd import DeviceDipoles
    dev = DeviceDipoles(_P(), scan_cls=_FakeScan)
    E = np.array([0.25, 0.75])
    D, dD = dev.with_energy_derivative(E)
    a = 1.3
    ref = _C0[None] * 1j * a * np.exp(1j * a * E)[:, None, None] + _C1[None] * 2 * E[:, None, None] * a
    assert np.allclose(dD.numpy(), ref, rtol=1e-14, atol=0)
    assert np.allclose(D.numpy(), _closed_form(E, a, np), rtol=1e-14, atol=0)
