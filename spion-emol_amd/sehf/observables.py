"""Photoionization observables from the scan's reduced dipoles, in torch FP64.

The counterpart of the reference's post-processing formulas
(postprocessing/cross_section_delay.py):

* sigma  -- sigma_from_dipole  (cross_section_delay.py:195-220)
            sigma_L = 8 pi omega / (c k) sum |d^L|^2,
            sigma_V = 8 pi / (c k omega) sum |d^V|^2, c = 137.036 (:120)
* tau    -- time_delay          (:227-263)
            tau = + Im[sum d* dd/dE] / sum |d|^2, dd/dE = (dd/dk) / k with
            dd/dk from numpy.gradient's rule on the k grid (second order
            in the interior, first order at the two ends)
* q_sigma, dphi_rms             (:266-278)
* beta   -- compute_beta        (:348-429)
            beta = N1 / sigma_tot - 1 with sigma_tot = (1/3) int sum_q |D^q|^2
            and N1 = int |sum_q D^q(n) n_q|^2 over a Gauss-Legendre (cos
            theta) x uniform (phi) grid, D^q(n) = sum_lm d^q_lm Y^R_lm(n).

Everything runs on the tensors' device; the dipoles of a GPU scan stay on
the GPU.  Channel order mu = l^2 + l + m; slot order of the scan's D is
L x, L y, L z, V x, V y, V z.
"""
from __future__ import annotations

import math

import numpy as np
import torch

C_AU = 137.036   # cross_section_delay.py:120


def _t(x, device=None, dtype=torch.float64):
    if torch.is_tensor(x):
        return x.to(device=device or x.device, dtype=dtype if not x.is_complex() else torch.complex128)
    a = np.asarray(x)
    if np.iscomplexobj(a):
        return torch.as_tensor(a, dtype=torch.complex128, device=device)
    return torch.as_tensor(a, dtype=dtype, device=device)


def sigma(k, D, omega, gauge: str):
    """Total cross section [bohr^2], D: [n_E, n_ch] complex."""
    D = _t(D)
    k = _t(k, D.device)
    omega = _t(omega, D.device)
    s2 = (D.abs() ** 2).sum(dim=1)
    if gauge == "length":
        return 8.0 * math.pi * omega / (C_AU * k) * s2
    if gauge == "velocity":
        return 8.0 * math.pi / (C_AU * k * omega) * s2
    raise ValueError(f"unknown gauge {gauge!r}")


def gradient(f, x):
    """d f / d x along dim 0 with numpy.gradient's rule for a non-uniform
    grid (edge_order=1): interior
      (h1^2 f[i+1] + (h0^2 - h1^2) f[i] - h0^2 f[i-1]) / (h0 h1 (h0 + h1)),
    h1 = x[i] - x[i-1], h0 = x[i+1] - x[i]; one-sided first differences at
    the ends."""
    n = f.shape[0]
    if n < 2:
        raise ValueError("gradient needs at least two points")
    out = torch.empty_like(f)
    dx = (x[1:] - x[:-1])
    shape = (-1,) + (1,) * (f.dim() - 1)
    out[0] = (f[1] - f[0]) / dx[0]
    out[-1] = (f[-1] - f[-2]) / dx[-1]
    if n > 2:
        h1 = dx[:-1].reshape(shape)
        h0 = dx[1:].reshape(shape)
        a = -h0 / (h1 * (h0 + h1))
        b = (h0 - h1) / (h0 * h1)
        c = h1 / (h0 * (h0 + h1))
        out[1:-1] = a * f[:-2] + b * f[1:-1] + c * f[2:]
    return out


def time_delay(k, D, with_minus: bool = False):
    """Wigner delay [a.u.] per energy from finite differences in k."""
    D = _t(D)
    k = _t(k, D.device)
    dD_dE = gradient(D, k) / k.reshape(-1, 1)
    num = (D.conj() * dD_dE).sum(dim=1).imag
    den = (D.abs() ** 2).sum(dim=1)
    tau = torch.where(den > 1e-30, num / torch.where(den > 1e-30, den, torch.ones_like(den)),
                      torch.zeros_like(num))
    return -tau if with_minus else tau


def time_delay_analytic(D, dD_dE, with_minus: bool = False):
    """The same delay with an analytic energy derivative of D."""
    D = _t(D)
    dD_dE = _t(dD_dE, D.device)
    num = (D.conj() * dD_dE).sum(dim=1).imag
    den = (D.abs() ** 2).sum(dim=1)
    tau = torch.where(den > 1e-30, num / torch.where(den > 1e-30, den, torch.ones_like(den)),
                      torch.zeros_like(num))
    return -tau if with_minus else tau


def q_sigma(D_L, D_V, omega):
    D_L, D_V = _t(D_L), _t(D_V)
    omega = _t(omega, D_L.device)
    den = omega ** 2 * (D_L.abs() ** 2).sum(dim=1)
    num = (D_V.abs() ** 2).sum(dim=1)
    return torch.where(den > 1e-30, num / torch.where(den > 1e-30, den, torch.ones_like(den)),
                       torch.full_like(num, float("nan")))


def dphi_rms(D_L, D_V):
    D_L, D_V = _t(D_L), _t(D_V)
    dphi = torch.angle(D_V * D_L.conj())
    w = D_L.abs() * D_V.abs()
    sw = w.sum(dim=1)
    swp = (w * dphi ** 2).sum(dim=1)
    return torch.sqrt(torch.where(sw > 1e-30, swp / torch.where(sw > 1e-30, sw, torch.ones_like(sw)),
                                  torch.zeros_like(sw)))


def real_ylm_table(lmax: int, x, phi):
    """Y^R_lm on the product grid (x = cos theta) x phi, rows mu = l^2+l+m:
    normalised associated Legendre functions without the Condon-Shortley
    phase by the standard three-term recursion, times sqrt2 cos(m phi)
    (m > 0) or sqrt2 sin(|m| phi) (m < 0)."""
    x = _t(x)
    phi = _t(phi, x.device)
    nt, nphi = x.shape[0], phi.shape[0]
    st = torch.sqrt(torch.clamp(1.0 - x * x, min=0.0))
    P = {}
    pmm = torch.full_like(x, 1.0 / math.sqrt(4.0 * math.pi))
    for m in range(lmax + 1):
        if m > 0:
            pmm = pmm * math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * st
        P[(m, m)] = pmm
        if m + 1 <= lmax:
            P[(m + 1, m)] = math.sqrt(2.0 * m + 3.0) * x * pmm
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            P[(l, m)] = a * (x * P[(l - 1, m)] - b * P[(l - 2, m)])
    Y = torch.empty(((lmax + 1) ** 2, nt, nphi), dtype=torch.float64, device=x.device)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            am = abs(m)
            if m == 0:
                ang = torch.ones_like(phi)
            elif m > 0:
                ang = math.sqrt(2.0) * torch.cos(m * phi)
            else:
                ang = math.sqrt(2.0) * torch.sin(am * phi)
            Y[l * l + l + m] = P[(l, am)].reshape(-1, 1) * ang.reshape(1, -1)
    return Y


def beta(D_x, D_y, D_z, mu_list=None, n_theta: int = 40, n_phi: int = 80):
    """Orientation-averaged asymmetry parameter per energy (bounded in [-1, 2])."""
    D_x, D_y, D_z = _t(D_x), _t(D_y), _t(D_z)
    dev = D_x.device
    n_ch = D_x.shape[1]
    mu = list(range(n_ch)) if mu_list is None else list(mu_list)
    lmax = int(math.isqrt(max(mu)))
    xg, wg = np.polynomial.legendre.leggauss(n_theta)
    x = torch.as_tensor(xg, dtype=torch.float64, device=dev)
    w = torch.as_tensor(wg, dtype=torch.float64, device=dev)
    phi = torch.arange(n_phi, dtype=torch.float64, device=dev) * (2.0 * math.pi / n_phi)
    Yall = real_ylm_table(lmax, x, phi)
    Y = Yall[torch.as_tensor(mu, device=dev)].reshape(len(mu), -1).to(torch.complex128)
    st = torch.sqrt(torch.clamp(1.0 - x * x, min=0.0)).reshape(-1, 1)
    nx = (st * torch.cos(phi).reshape(1, -1)).reshape(-1)
    ny = (st * torch.sin(phi).reshape(1, -1)).reshape(-1)
    nz = x.reshape(-1, 1).expand(n_theta, n_phi).reshape(-1)
    wt = (w.reshape(-1, 1) * (2.0 * math.pi / n_phi)).expand(n_theta, n_phi).reshape(-1)
    Mx, My, Mz = D_x @ Y, D_y @ Y, D_z @ Y                     # [n_E, n_grid]
    sig = (1.0 / 3.0) * (((Mx.abs() ** 2 + My.abs() ** 2 + Mz.abs() ** 2) * wt).sum(dim=1))
    dot = Mx * nx + My * ny + Mz * nz
    N1 = (dot.abs() ** 2 * wt).sum(dim=1)
    return torch.where(sig > 1e-30, N1 / torch.where(sig > 1e-30, sig, torch.ones_like(sig)) - 1.0,
                       torch.zeros_like(sig))


def from_scan(results, E_init: float, device=None):
    """Observables of a list of scan results (one dict per energy, as
    returned by Scan.run) for the initial-state energy E_init (Hartree).
    Results of a tangent run along E (Scan.run(..., tangent=(1, 0))) also
    give the delay from the analytic energy derivative (tau_*_analytic)."""
    dev = device or ("cuda" if torch.cuda.is_available() else "cpu")
    k = torch.tensor([r["k"] for r in results], dtype=torch.float64, device=dev)
    D = torch.as_tensor(np.stack([r["D"] for r in results]), dtype=torch.complex128, device=dev)  # [nE, 6, N]
    omega = 0.5 * k * k - E_init
    DL = D[:, 0:3, :].reshape(len(results), -1)
    DV = D[:, 3:6, :].reshape(len(results), -1)
    out = {"k": k, "omega": omega,
           "sigma_L": sigma(k, DL, omega, "length"), "sigma_V": sigma(k, DV, omega, "velocity"),
           "q_sigma": q_sigma(DL, DV, omega), "dphi_rms": dphi_rms(DL, DV),
           "beta_L": beta(D[:, 0], D[:, 1], D[:, 2]), "beta_V": beta(D[:, 3], D[:, 4], D[:, 5])}
    if len(results) >= 2:
        out["tau_L"] = time_delay(k, DL)
        out["tau_V"] = time_delay(k, DV)
    if all("dD" in r for r in results):
        # results of a tangent run along E (t_E = 1): the analytic delay
        dD = torch.as_tensor(np.stack([r["dD"] for r in results]), dtype=torch.complex128, device=dev)
        out["tau_L_analytic"] = time_delay_analytic(DL, dD[:, 0:3, :].reshape(len(results), -1))
        out["tau_V_analytic"] = time_delay_analytic(DV, dD[:, 3:6, :].reshape(len(results), -1))
    return out
