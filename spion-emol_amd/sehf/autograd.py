"""Differentiable reduced dipoles d_lm(E, alpha) for torch post-processing.

`DeviceDipoles(problem)(E, alpha)` returns the reduced dipoles of every energy
E (Hartree) as a complex128 tensor [n_E, 6, N] (slots L x, y, z, V x, y, z),
computed by the device scan (sehf_scan_run / sehf_scan_run_tangent through
the C ABI).  The tensor is differentiable with respect to E and to the
polarisation strength alpha of the problem's model potential: the Jacobian
columns dd/dE and dd/dalpha come from the scan's forward-mode derivative of
the whole stage chain, so the Wigner delay tau = Im sum d* dd/dE / sum |d|^2
and d sigma / d alpha need no finite differences over the k grid (the
reference takes dd/dE from numpy.gradient over k,
postprocessing/cross_section_delay.py:227-263).

Each energy depends on its own E only, so one tangent direction per energy
(t_E = 1) gives the whole E-Jacobian; alpha is shared by all energies and
takes a second direction (t_alpha = 1).  The scan computes both in one call
when both are needed, with every energy listed once per direction.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import scan as _scan


def _run(sc, E: np.ndarray, want_E: bool, want_a: bool, batch: int):
    """D [nE, 6, N] and the requested Jacobian columns as numpy arrays."""
    ks = np.sqrt(2.0 * E)
    nE = len(ks)
    dirs = []
    if want_E:
        dirs.append(np.tile([1.0, 0.0], (nE, 1)))
    if want_a:
        dirs.append(np.tile([0.0, 1.0], (nE, 1)))
    if not dirs:
        res = sc.run(ks, batch=batch)
        return np.stack([r["D"] for r in res]), None, None
    kk = np.concatenate([ks] * len(dirs))
    res = sc.run(kk, batch=batch, tangent=np.concatenate(dirs))
    D = np.stack([r["D"] for r in res[:nE]])
    dE = np.stack([r["dD"] for r in res[:nE]]) if want_E else None
    da = np.stack([r["dD"] for r in res[nE * (len(dirs) - 1):]]) if want_a else None
    return D, dE, da


class _DipoleFn(torch.autograd.Function):
    """d(E, alpha) with the Jacobian from the device tangent.

    For a real input x and complex output d, the vector-Jacobian product of
    torch's convention is  grad_x = Re sum conj(grad_d) * (dd/dx)."""

    @staticmethod
    def forward(ctx, E, alpha, runner, batch):
        want_E, want_a = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        sc = runner.scan_for(float(alpha.detach().cpu()))
        Dn, dE, da = _run(sc, E.detach().cpu().numpy().astype(np.float64), want_E, want_a, batch)
        dev = E.device
        ctx.dE = torch.as_tensor(dE, dtype=torch.complex128, device=dev) if dE is not None else None
        ctx.da = torch.as_tensor(da, dtype=torch.complex128, device=dev) if da is not None else None
        return torch.as_tensor(Dn, dtype=torch.complex128, device=dev)

    @staticmethod
    def backward(ctx, gD):
        gE = ga = None
        if ctx.needs_input_grad[0]:
            gE = (gD.conj() * ctx.dE).real.sum(dim=(1, 2))
        if ctx.needs_input_grad[1]:
            ga = (gD.conj() * ctx.da).real.sum().reshape(())
        return gE, ga, None, None


class DeviceDipoles:
    """Reduced dipoles of one molecule as a differentiable function of the
    photoelectron energy E and the polarisation strength alpha.

    The energy-independent device tables are built once per alpha value
    (sehf_scan_create) and cached."""

    def __init__(self, problem, ctx=None, batch: int = 0, scan_cls=None):
        self.problem = problem
        self.ctx = ctx
        self.batch = batch
        self._scan_cls = scan_cls or _scan.Scan
        self._scans = {}

    def scan_for(self, alpha: float):
        key = float(alpha)
        if key not in self._scans:
            p = self.problem
            if key != float(p.pol_alpha):
                if p.pol_kind == 0:
                    raise ValueError("the problem has no polarisation potential (pol_kind = 0); alpha must equal 0")
                p = dataclasses.replace(p, pol_alpha=key)
            self._scans[key] = self._scan_cls(p, self.ctx)
        return self._scans[key]

    def __call__(self, E, alpha=None):
        """E: float64 tensor [n_E] (Hartree); alpha: 0-d float64 tensor
        (default: the problem's pol_alpha, not differentiable)."""
        if not torch.is_tensor(E):
            E = torch.as_tensor(E, dtype=torch.float64)
        if alpha is None:
            alpha = torch.tensor(float(self.problem.pol_alpha), dtype=torch.float64)
        return _DipoleFn.apply(E, alpha, self, self.batch)

    def with_energy_derivative(self, E):
        """(D, dD/dE) as complex128 tensors without building a graph."""
        E = np.asarray(E.detach().cpu() if torch.is_tensor(E) else E, dtype=np.float64)
        Dn, dE, _ = _run(self.scan_for(float(self.problem.pol_alpha)), E, True, False, self.batch)
        return torch.as_tensor(Dn, dtype=torch.complex128), torch.as_tensor(dE, dtype=torch.complex128)
