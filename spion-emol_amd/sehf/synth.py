"""Synthetic preprocessing inputs for the solver (the preproc HDF5 contents).

The solver consumes what the reference's `preprocess_molden` writes
(`io::PreprocData`, src/io/HDF5Reader.hpp:42-86): a radial grid, nuclei,
the Hartree potential's single-centre expansion `/potential/V_H`, the
occupied orbitals' SCE radial coefficients `/orbitals/psi_lm`, orbital
energies and optionally an `/initial_state`.  Real molden inputs are out of
scope; this module builds seeded synthetic molecules of the benchmark shapes:

* ``h2_like``   -- two +1 nuclei at z=+-0.70, one doubly occupied orbital with
  F_l(r) = N r e^{-1.2 r} (1 + c_l r) on (l,m) = (0,0), (2,0).
* ``h2o_like``  -- O(+8) at the origin, H(+1) at (+-1.43, 0, 1.11); n_occ
  orbitals whose SCE radial functions (l <= 10) are sums of three normalised
  Slater terms, exponents U(0.8, 8), coefficients N(0, 1), Gram-Schmidt
  orthonormalised on the grid.
* ``ring_like`` -- planar 8 x (+6) / 8 x (+9) ring (large-molecule shape).

V_H is the Hartree potential of the closed-shell density (2 electrons per
orbital), obtained from the radial Poisson Green's function
  V_lm(r) = 4 pi/(2l+1) [ r^-(l+1) int_0^r rho_lm s^(l+2) ds
                          + r^l int_r^inf rho_lm s^(1-l) ds ]
(the formula of preprocessing/src/potential/Hartree.hpp:69-244), with the
density's SCE rho_lm taken by exact angular quadrature.

Real spherical harmonics follow the reference convention
(src/angular/Gaunt.hpp:143-181): no Condon-Shortley phase,
Y_{l,m>0} = sqrt2 S_l^m cos(m phi), Y_{l,m<0} = sqrt2 S_l^|m| sin(|m| phi),
channel index l*l + l + m.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Problem:
    """One molecule on one radial grid, plus the solver's angular cutoff."""
    name: str
    l_cont: int
    n_occ: int
    dr: float
    Nr: int
    Z: np.ndarray             # [n_atoms] int32
    xyz: np.ndarray           # [n_atoms, 3] bohr
    V_H: np.ndarray           # [nlm_vh, Nr] float64 (row-major)
    psi_lm: np.ndarray        # [n_orb, nlm_orb, Nr] float64, F without the r factor
    orb_energies: np.ndarray  # [n_orb] Hartree
    init_psi_lm: np.ndarray | None = None   # [nlm_orb, Nr]
    init_energy: float = 0.0
    pol_kind: int = 0         # 0 none, 1 damped (reference form), 2 soft
    pol_alpha: float = 0.0
    pol_rc: float = 1.5
    allow_closed: bool = False

    # ---- derived sizes (WavefunctionSetup.cpp:236-246, main.cpp:491-506) ----
    @property
    def l_exch(self) -> int:
        return min(self.l_cont, 10)

    @property
    def n_psi(self) -> int:
        return (self.l_cont + 1) ** 2

    @property
    def n_sigma(self) -> int:
        return (self.l_exch + 1) ** 2

    @property
    def n_f(self) -> int:
        return self.n_occ * self.n_sigma

    @property
    def n_tot(self) -> int:
        return self.n_psi + self.n_f

    @property
    def n_lambda(self) -> int:
        return (self.l_cont + self.l_exch + 1) ** 2

    @property
    def e_homo(self) -> float:
        return float(self.init_energy if self.init_psi_lm is not None else self.orb_energies[self.n_occ - 1])

    def with_lcont(self, l_cont: int) -> "Problem":
        """Same molecule, different continuum cutoff (sizes must still fit)."""
        return dataclasses.replace(self, l_cont=l_cont)


# ---------------------------------------------------------------------------
# Real spherical harmonics on an angular grid (vectorised).
# ---------------------------------------------------------------------------
def real_ylm_table(lmax: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Y^R_{lm}(theta, phi) for all (l, m), l <= lmax; shape [(lmax+1)^2, npts]."""
    ct, st = np.cos(theta), np.sin(theta)
    out = np.zeros(((lmax + 1) ** 2, theta.size))
    # normalised associated Legendre S_l^m without Condon-Shortley phase
    s_mm = np.full(theta.shape, 1.0 / math.sqrt(4.0 * math.pi))
    for m in range(lmax + 1):
        if m > 0:
            s_mm = s_mm * math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * st
        s_prev2 = None
        s_prev = s_mm
        for l in range(m, lmax + 1):
            if l == m:
                s = s_mm
            elif l == m + 1:
                s = math.sqrt(2.0 * m + 3.0) * ct * s_mm
            else:
                a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                s = a * (ct * s_prev - b * s_prev2)
            if l > m:
                s_prev2, s_prev = s_prev, s
            if m == 0:
                out[l * l + l] = s
            else:
                out[l * l + l + m] = math.sqrt(2.0) * s * np.cos(m * phi)
                out[l * l + l - m] = math.sqrt(2.0) * s * np.sin(m * phi)
    return out


def angular_grid(lmax_integrand: int):
    """Gauss-Legendre x trapezoid grid exact for polynomials of degree <= lmax_integrand."""
    n_t = lmax_integrand // 2 + 2
    n_p = lmax_integrand + 2
    x, w = np.polynomial.legendre.leggauss(n_t)
    theta = np.arccos(x)
    phi = 2.0 * np.pi * np.arange(n_p) / n_p
    T, P = np.meshgrid(theta, phi, indexing="ij")
    W = np.outer(w, np.full(n_p, 2.0 * np.pi / n_p))
    return T.ravel(), P.ravel(), W.ravel()


# ---------------------------------------------------------------------------
# Density SCE and Hartree potential.
# ---------------------------------------------------------------------------
def _cumtrapz(f: np.ndarray, dr: float) -> np.ndarray:
    out = np.zeros_like(f)
    out[..., 1:] = np.cumsum(0.5 * dr * (f[..., 1:] + f[..., :-1]), axis=-1)
    return out


def hartree_from_orbitals(psi_lm: np.ndarray, n_occ: int, lmax_orb: int, lmax_vh: int,
                          dr: float, occ: float = 2.0) -> np.ndarray:
    """V_H[lm, ir] for lm up to (lmax_vh+1)^2 from the closed-shell density."""
    Nr = psi_lm.shape[-1]
    l_rho = min(2 * lmax_orb, lmax_vh)
    theta, phi, w = angular_grid(2 * lmax_orb + l_rho)
    n_lo = (lmax_orb + 1) ** 2
    Y_orb = real_ylm_table(lmax_orb, theta, phi)            # [n_lo, nq]
    Y_rho = real_ylm_table(l_rho, theta, phi) * w          # [n_lr, nq] with weights
    rho_lm = np.zeros(((l_rho + 1) ** 2, Nr))
    chunk = 512
    for r0 in range(0, Nr, chunk):
        sl = slice(r0, min(Nr, r0 + chunk))
        dens = np.zeros((sl.stop - sl.start, theta.size))
        for i in range(n_occ):
            phi_i = psi_lm[i, :n_lo, sl].T @ Y_orb              # [nr, nq]
            dens += occ * phi_i * phi_i
        rho_lm[:, sl] = (dens @ Y_rho.T).T
    r = np.arange(Nr) * dr
    V = np.zeros(((lmax_vh + 1) ** 2, Nr))
    rp = r[1:]
    for l in range(l_rho + 1):
        pref = 4.0 * math.pi / (2.0 * l + 1.0)
        for m in range(-l, l + 1):
            ch = l * l + l + m
            rho = rho_lm[ch]
            if np.max(np.abs(rho)) < 1e-14:
                continue
            f_in = rho * r ** (l + 2)
            f_out = np.zeros(Nr)
            if l == 0:
                f_out = rho * r
            else:
                f_out[1:] = rho[1:] * rp ** (1 - l)
            I_in = _cumtrapz(f_in, dr)
            I_out = _cumtrapz(f_out, dr)
            v = np.zeros(Nr)
            v[1:] = pref * (I_in[1:] / rp ** (l + 1) + rp ** l * (I_out[-1] - I_out[1:]))
            if l == 0:
                v[0] = pref * I_out[-1]
            V[ch] = v
    return V


def _simpson_weights(Nr: int, dr: float) -> np.ndarray:
    w = np.ones(Nr)
    w[1:-1:2] = 4.0
    w[2:-1:2] = 2.0
    if Nr % 2 == 0:   # even point count: trapezoid on the last interval
        w[-1] = 1.0
    return w * dr / 3.0


def _gram_schmidt(F: np.ndarray, dr: float) -> np.ndarray:
    """Orthonormalise orbitals F[i, lam, r] under sum_lam int F_i F_j r^2 dr."""
    Nr = F.shape[-1]
    r2w = (np.arange(Nr) * dr) ** 2 * _simpson_weights(Nr, dr)
    out = F.copy()
    for i in range(out.shape[0]):
        for j in range(i):
            ov = np.sum(out[i] * out[j] * r2w)
            out[i] -= ov * out[j]
        nrm = math.sqrt(np.sum(out[i] * out[i] * r2w))
        out[i] /= nrm
    return out


def _slater(l: int, zeta: float, r: np.ndarray) -> np.ndarray:
    """Normalised Slater radial function r^l e^{-zeta r} (int R^2 r^2 dr = 1)."""
    norm = (2.0 * zeta) ** (l + 1.5) / math.sqrt(math.factorial(2 * l + 2))
    return norm * r ** l * np.exp(-zeta * r)


# ---------------------------------------------------------------------------
# Molecules.
# ---------------------------------------------------------------------------
def h2_like(l_cont: int = 8, dr: float = 0.01, r_max: float = 30.0, seed: int = 0,
            lmax_vh: int | None = None) -> Problem:
    Nr = int(round(r_max / dr)) + 1
    r = np.arange(Nr) * dr
    l_exch = min(l_cont, 10)
    l_orb = l_cont + l_exch
    nlm_orb = (l_orb + 1) ** 2
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.2, 0.2, size=2)
    F = np.zeros((1, nlm_orb, Nr))
    for (l, cl) in ((0, c[0]), (2, c[1])):
        if (l + 1) ** 2 <= nlm_orb:
            F[0, l * l + l, :] = r * np.exp(-1.2 * r) * (1.0 + cl * r)
    F = _gram_schmidt(F, dr)
    lvh = 2 * l_cont if lmax_vh is None else lmax_vh
    V_H = hartree_from_orbitals(F, 1, min(2, l_orb), lvh, dr)
    return Problem(name=f"h2_like_l{l_cont}", l_cont=l_cont, n_occ=1, dr=dr, Nr=Nr,
                   Z=np.array([1, 1], dtype=np.int32),
                   xyz=np.array([[0.0, 0.0, 0.70], [0.0, 0.0, -0.70]]),
                   V_H=V_H, psi_lm=F, orb_energies=np.array([-0.594]))


def h2o_like(l_cont: int = 20, dr: float = 0.01, r_max: float = 50.0, n_occ: int = 5,
             seed: int = 1, l_orb_sce: int = 10, pol: tuple | None = None,
             allow_closed: bool = False) -> Problem:
    Nr = int(round(r_max / dr)) + 1
    r = np.arange(Nr) * dr
    l_exch = min(l_cont, 10)
    l_orb = l_cont + l_exch
    nlm_orb = (l_orb + 1) ** 2
    lsce = min(l_orb_sce, l_orb)
    rng = np.random.default_rng(seed)
    F = np.zeros((n_occ, nlm_orb, Nr))
    for i in range(n_occ):
        for lam in range((lsce + 1) ** 2):
            l = int(math.isqrt(lam))
            zetas = rng.uniform(0.8, 8.0, size=3)
            coefs = rng.normal(0.0, 1.0, size=3)
            F[i, lam] = sum(cf * _slater(l, z, r) for cf, z in zip(coefs, zetas))
    F = _gram_schmidt(F, dr)
    V_H = hartree_from_orbitals(F, n_occ, lsce, 2 * l_cont, dr)
    energies = np.array([-20.55, -1.34, -0.70, -0.57, -0.50][:n_occ]) if n_occ <= 5 else \
        np.linspace(-20.0, -0.5, n_occ)
    p = Problem(name=f"h2o_like_l{l_cont}", l_cont=l_cont, n_occ=n_occ, dr=dr, Nr=Nr,
                Z=np.array([8, 1, 1], dtype=np.int32),
                xyz=np.array([[0.0, 0.0, 0.0], [1.43, 0.0, 1.11], [-1.43, 0.0, 1.11]]),
                V_H=V_H, psi_lm=F, orb_energies=energies, allow_closed=allow_closed)
    if pol is not None:
        p.pol_kind, p.pol_alpha, p.pol_rc = 2, float(pol[0]), float(pol[1])
    return p


def k_grid(k_min: float, k_max: float, n: int) -> np.ndarray:
    return np.linspace(k_min, k_max, n)


# The benchmark configurations (grid, cutoffs and scans).
def config(name: str) -> tuple[Problem, np.ndarray]:
    if name == "h2_l8":          # CPU-runnable plain case
        return h2_like(8, 0.01, 30.0, seed=0), k_grid(0.1, 1.0, 10)
    if name == "h2o_l20":        # plain per-ik on one GPU
        return h2o_like(20, 0.01, 50.0, 5, seed=1, allow_closed=True), k_grid(0.1, 2.0, 100)
    if name == "h2o_l30":        # batched scan, closed channels allowed
        return h2o_like(30, 0.01, 60.0, 5, seed=1, allow_closed=True), k_grid(0.05, 2.0, 256)
    if name == "h2o_l20_pol":    # adjoint / polarisation case
        return h2o_like(20, 0.01, 50.0, 5, seed=1, pol=(9.8, 1.5)), k_grid(0.2, 1.5, 64)
    raise KeyError(name)
