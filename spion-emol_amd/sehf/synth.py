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
from decimal import Decimal, localcontext

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
# Host-independent arithmetic.  The generated inputs are bit for bit the same
# on every host: full-size oracle fixtures computed on one machine are
# compared with device runs whose inputs are generated on another, and the
# energy-amplified outputs move by 1e-8 under a last-bit change of the
# inputs.  So nothing below goes through libm transcendentals, BLAS or SIMD
# reductions, whose last bits depend on the CPU model and its vector width:
# only elementwise IEEE + - * / and sqrt (correctly rounded), sequential
# accumulation in a fixed order, math.fsum (correctly rounded sums) and
# decimal arithmetic (software, correctly rounded conversions).
# ---------------------------------------------------------------------------
_PI_DEC = Decimal("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899863")
_PREC = 40


def _dec_cos_sin(x: Decimal) -> tuple:
    """cos x, sin x by their Taylor series at 40 significant digits."""
    with localcontext() as c:
        c.prec = _PREC + 10
        two_pi = 2 * _PI_DEC
        x = x % two_pi
        if x > _PI_DEC:
            x -= two_pi
        x2, term, cs, k = x * x, Decimal(1), Decimal(1), 0
        while True:
            k += 2
            term = -term * x2 / (k * (k - 1))
            if term == 0 or abs(term) < Decimal(10) ** -(_PREC + 5):
                break
            cs += term
        term, sn, k = x, x, 1
        while True:
            k += 2
            term = -term * x2 / (k * (k - 1))
            if term == 0 or abs(term) < Decimal(10) ** -(_PREC + 5):
                break
            sn += term
        return +cs, +sn


def _gauss_legendre(n: int) -> tuple:
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1]: Newton on the
    three-term recurrence of P_n in decimal arithmetic, then rounded once."""
    xs, ws = [], []
    with localcontext() as c:
        c.prec = _PREC + 10
        tol = Decimal(10) ** -(_PREC + 2)
        for i in range(1, n + 1):
            x, _ = _dec_cos_sin(_PI_DEC * (Decimal(i) - Decimal("0.25")) / (Decimal(n) + Decimal("0.5")))
            for _ in range(200):
                p0, p1 = Decimal(1), x
                for k in range(2, n + 1):
                    p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
                dp = n * (x * p1 - p0) / (x * x - 1)
                dx = p1 / dp
                x -= dx
                if abs(dx) < tol:
                    break
            p0, p1 = Decimal(1), x
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = n * (x * p1 - p0) / (x * x - 1)
            xs.append(float(x))
            ws.append(float(2 / ((1 - x * x) * dp * dp)))
    order = np.argsort(xs)
    return np.array(xs)[order], np.array(ws)[order]


def _exp_neg(zeta: float, dr: float, Nr: int) -> np.ndarray:
    """e^{-zeta n dr} for n in [0, Nr): e^{-64 a zeta dr} e^{-b zeta dr} with
    n = 64 a + b, both factors from decimal exp (one rounding each), so every
    value is within an ulp or two of the exact one and the same on every host."""
    with localcontext() as c:
        c.prec = _PREC
        zd = Decimal(zeta) * Decimal(dr)
        lo = np.array([float((-(zd * b)).exp()) for b in range(64)])
        hi = np.array([float((-(zd * 64 * a)).exp()) for a in range((Nr + 63) // 64)])
    n = np.arange(Nr)
    return hi[n // 64] * lo[n % 64]


def _ipow(r: np.ndarray, k: int) -> np.ndarray:
    """r^k for an integer k >= 0 by repeated multiplication."""
    out = np.ones_like(r)
    for _ in range(k):
        out = out * r
    return out


# ---------------------------------------------------------------------------
# Real spherical harmonics on an angular grid (vectorised).
# ---------------------------------------------------------------------------
def _ylm_core(lmax: int, ct: np.ndarray, st: np.ndarray, cos_m, sin_m) -> np.ndarray:
    """Y^R_{lm} from cos(theta), sin(theta) and cos(m phi), sin(m phi) tables;
    shape [(lmax+1)^2, npts]."""
    out = np.zeros(((lmax + 1) ** 2, ct.size))
    # normalised associated Legendre S_l^m without Condon-Shortley phase
    s_mm = np.full(ct.shape, 1.0 / math.sqrt(4.0 * math.pi))
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
                out[l * l + l + m] = math.sqrt(2.0) * s * cos_m(m)
                out[l * l + l - m] = math.sqrt(2.0) * s * sin_m(m)
    return out


def real_ylm_table(lmax: int, theta: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Y^R_{lm}(theta, phi) for all (l, m), l <= lmax; shape [(lmax+1)^2, npts]."""
    return _ylm_core(lmax, np.cos(theta), np.sin(theta), lambda m: np.cos(m * phi), lambda m: np.sin(m * phi))


def angular_grid(lmax_integrand: int):
    """Gauss-Legendre x trapezoid grid exact for polynomials of degree <= lmax_integrand."""
    n_t = lmax_integrand // 2 + 2
    n_p = lmax_integrand + 2
    x, w = _gauss_legendre(n_t)
    theta = np.arccos(x)
    phi = 2.0 * np.pi * np.arange(n_p) / n_p
    T, P = np.meshgrid(theta, phi, indexing="ij")
    W = np.outer(w, np.full(n_p, 2.0 * np.pi / n_p))
    return T.ravel(), P.ravel(), W.ravel()


def _ylm_on_grid(lmax: int, lmax_integrand: int):
    """Y^R_{lm} on angular_grid(lmax_integrand)'s points and its weights, from
    host-independent cos/sin values (decimal), in the grid's point order."""
    n_t = lmax_integrand // 2 + 2
    n_p = lmax_integrand + 2
    x, w = _gauss_legendre(n_t)
    cs = [_dec_cos_sin(2 * _PI_DEC * j / n_p) for j in range(n_p)]
    cos_t = np.array([float(c) for c, _ in cs])
    sin_t = np.array([float(s_) for _, s_ in cs])
    ct = np.repeat(x, n_p)
    st = np.sqrt((1.0 - x) * (1.0 + x))
    st = np.repeat(st, n_p)
    jj = np.tile(np.arange(n_p), n_t)
    W = np.repeat(w, n_p) * (2.0 * math.pi / n_p)
    Y = _ylm_core(lmax, ct, st, lambda m: cos_t[(m * jj) % n_p], lambda m: sin_t[(m * jj) % n_p])
    return Y, W


# ---------------------------------------------------------------------------
# Density SCE and Hartree potential.
# ---------------------------------------------------------------------------
def _cumtrapz(f: np.ndarray, dr: float) -> np.ndarray:
    out = np.zeros_like(f)
    out[..., 1:] = np.cumsum(0.5 * dr * (f[..., 1:] + f[..., :-1]), axis=-1)
    return out


def hartree_from_orbitals(psi_lm: np.ndarray, n_occ: int, lmax_orb: int, lmax_vh: int,
                          dr: float, occ: float = 2.0) -> np.ndarray:
    """V_H[lm, ir] for lm up to (lmax_vh+1)^2 from the closed-shell density.
    The angular contractions accumulate term by term in a fixed order (no
    BLAS), so the result is the same on every host."""
    Nr = psi_lm.shape[-1]
    l_rho = min(2 * lmax_orb, lmax_vh)
    deg = 2 * lmax_orb + l_rho
    n_lo = (lmax_orb + 1) ** 2
    Y_orb, _ = _ylm_on_grid(lmax_orb, deg)                  # [n_lo, nq]
    Y_rho, wq = _ylm_on_grid(l_rho, deg)
    Y_rho = Y_rho * wq                                      # [n_lr, nq] with weights
    nq = wq.size
    rho_lm = np.zeros(((l_rho + 1) ** 2, Nr))
    chunk = 512
    for r0 in range(0, Nr, chunk):
        sl = slice(r0, min(Nr, r0 + chunk))
        dens = np.zeros((sl.stop - sl.start, nq))
        for i in range(n_occ):
            phi_i = np.zeros_like(dens)
            for lam in range(n_lo):
                phi_i += psi_lm[i, lam, sl][:, None] * Y_orb[lam][None, :]
            dens += occ * phi_i * phi_i
        acc = np.zeros((Y_rho.shape[0], sl.stop - sl.start))
        for q in range(nq):
            acc += Y_rho[:, q][:, None] * dens[:, q][None, :]
        rho_lm[:, sl] = acc
    return _hartree_from_rho(rho_lm, l_rho, lmax_vh, dr)


def _hartree_from_rho(rho_lm: np.ndarray, l_rho: int, lmax_vh: int, dr: float) -> np.ndarray:
    """V_H[lm, ir] from the density's multipoles rho_lm[lm, ir] (l <= l_rho) by
    the radial Poisson formula, trapezoid cumulative integrals."""
    Nr = rho_lm.shape[-1]
    r = np.arange(Nr) * dr
    V = np.zeros(((lmax_vh + 1) ** 2, Nr))
    rp = r[1:]
    for l in range(l_rho + 1):
        pref = 4.0 * math.pi / (2.0 * l + 1.0)
        rl, rl1 = _ipow(rp, l), _ipow(rp, l + 1)
        for m in range(-l, l + 1):
            ch = l * l + l + m
            rho = rho_lm[ch]
            if np.max(np.abs(rho)) < 1e-14:
                continue
            f_in = rho * _ipow(r, l + 2)
            f_out = np.zeros(Nr)
            if l == 0:
                f_out = rho * r
            else:
                f_out[1:] = rho[1:] / _ipow(rp, l - 1)
            I_in = _cumtrapz(f_in, dr)
            I_out = _cumtrapz(f_out, dr)
            v = np.zeros(Nr)
            v[1:] = pref * (I_in[1:] / rl1 + rl * (I_out[-1] - I_out[1:]))
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
    """Orthonormalise orbitals F[i, lam, r] under sum_lam int F_i F_j r^2 dr
    (overlaps by math.fsum, correctly rounded)."""
    Nr = F.shape[-1]
    r2w = (np.arange(Nr) * dr) ** 2 * _simpson_weights(Nr, dr)
    out = F.copy()
    for i in range(out.shape[0]):
        for j in range(i):
            ov = math.fsum((out[i] * out[j] * r2w).ravel())
            out[i] -= ov * out[j]
        nrm = math.sqrt(math.fsum((out[i] * out[i] * r2w).ravel()))
        out[i] /= nrm
    return out


def _slater(l: int, zeta: float, dr: float, Nr: int) -> np.ndarray:
    """Normalised Slater radial function r^l e^{-zeta r} (int R^2 r^2 dr = 1)
    on r = n dr."""
    r = np.arange(Nr) * dr
    tz = 2.0 * zeta
    norm = _ipow(np.array([tz]), l + 1)[0] * math.sqrt(tz) / math.sqrt(float(math.factorial(2 * l + 2)))
    return norm * _ipow(r, l) * _exp_neg(zeta, dr, Nr)


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
            F[0, l * l + l, :] = r * _exp_neg(1.2, dr, Nr) * (1.0 + cl * r)
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
            F[i, lam] = sum(cf * _slater(l, z, dr, Nr) for cf, z in zip(coefs, zetas))
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


def ring_like(l_cont: int = 60, dr: float = 0.02, r_max: float = 100.0, n_occ: int = 60, seed: int = 4,
              comps: int = 3, l_sce: int = 60) -> Problem:
    """Planar 16-centre ring: 8 charges +6 on a regular octagon of radius
    2.7 bohr and 8 charges +9 on radius 5.2 at the same angles (the origin is
    the centre of charge), n_occ doubly occupied orbitals and one extra-electron
    initial-state orbital.

    Each orbital's single-centre expansion has `comps` (l, m) components with
    l <= l_sce (the first orbital's first component at l = l_sce), each radial
    function a sum of three Slater terms r^l e^{-zeta r} with zeta ~ U(0.8, 8)
    and N(0, 1) coefficients (seeded).  No two orbitals share a component, so
    they are orthogonal by their angular parts and only normalised.  V_H comes
    from the closed-shell density, whose multipoles are Gaunt contractions of
    the few component pairs per orbital (real Gaunt coefficients by the
    angular quadrature of degree 2 l_sce + 2 l_cont); the Poisson step is the
    other configurations'.

    This molecule is too large for the CPU oracle, so nothing compares its
    inputs across hosts: the density contraction uses BLAS products."""
    Nr = int(round(r_max / dr)) + 1
    l_exch = min(l_cont, 10)
    l_orb = l_cont + l_exch
    nlm_orb = (l_orb + 1) ** 2
    n_ch = (l_sce + 1) ** 2
    rng = np.random.default_rng(seed)
    n_all = n_occ + 1
    top = l_sce * l_sce + rng.integers(0, 2 * l_sce + 1)          # an l = l_sce channel
    rest = rng.permutation(np.setdiff1d(np.arange(n_ch), [top]))[: n_all * comps - 1]
    chans = np.concatenate([[top], rest]).reshape(n_all, comps)
    r = np.arange(Nr) * dr
    r2w = r * r * _simpson_weights(Nr, dr)
    radial = np.zeros((n_all, comps, Nr))
    for i in range(n_all):
        for c in range(comps):
            l = int(math.isqrt(int(chans[i, c])))
            zetas = rng.uniform(0.8, 8.0, size=3)
            coefs = rng.normal(0.0, 1.0, size=3)
            radial[i, c] = sum(cf * _slater(l, z, dr, Nr) for cf, z in zip(coefs, zetas))
        nrm = math.sqrt(math.fsum((radial[i] * radial[i] * r2w).ravel()))
        radial[i] /= nrm
    F = np.zeros((n_occ, nlm_orb, Nr))
    for i in range(n_occ):
        for c in range(comps):
            F[i, chans[i, c]] = radial[i, c]
    init = np.zeros((nlm_orb, Nr))
    for c in range(comps):
        init[chans[n_occ, c]] = radial[n_occ, c]
    # density multipoles rho_LM = 2 sum_i sum_{a,b} F_ia F_ib G(a, b, LM)
    lmax_vh = 2 * l_cont
    l_rho = min(2 * l_sce, lmax_vh)
    deg = 2 * l_sce + l_rho
    Y_a, wq = _ylm_on_grid(l_sce, deg)
    pairs, cols = [], []
    for i in range(n_occ):
        for x in range(comps):
            for y in range(x, comps):
                a, b = int(chans[i, x]), int(chans[i, y])
                pairs.append((i, x, y))
                cols.append((2.0 if x != y else 1.0) * wq * Y_a[a] * Y_a[b])
    del Y_a
    Pm = np.stack(cols, axis=1)                                    # [nq, n_pairs]
    Y_r, _ = _ylm_on_grid(l_rho, deg)
    G = Y_r @ Pm                                                   # [n_LM, n_pairs]
    del Y_r, Pm
    prod = np.stack([radial[i, x] * radial[i, y] for i, x, y in pairs])   # [n_pairs, Nr]
    rho_lm = 2.0 * (G @ prod)
    V_H = _hartree_from_rho(rho_lm, l_rho, lmax_vh, dr)
    ang = 2.0 * math.pi * np.arange(8) / 8.0
    xyz = np.concatenate([np.stack([2.7 * np.cos(ang), 2.7 * np.sin(ang), np.zeros(8)], axis=1),
                          np.stack([5.2 * np.cos(ang), 5.2 * np.sin(ang), np.zeros(8)], axis=1)])
    p = Problem(name=f"ring_like_l{l_cont}", l_cont=l_cont, n_occ=n_occ, dr=dr, Nr=Nr,
                Z=np.array([6] * 8 + [9] * 8, dtype=np.int32), xyz=xyz, V_H=V_H, psi_lm=F,
                orb_energies=np.linspace(-20.0, -0.5, n_occ), init_psi_lm=init, init_energy=-0.25)
    return p


def k_grid(k_min: float, k_max: float, n: int) -> np.ndarray:
    return np.linspace(k_min, k_max, n)


_CONFIG_CACHE: dict = {}


# The benchmark configurations (grid, cutoffs and scans).  Generating one
# takes tens of seconds (host-independent arithmetic), so each is built once
# per process; callers get their own copy of the arrays.
def config(name: str) -> tuple[Problem, np.ndarray]:
    if name not in _CONFIG_CACHE:
        _CONFIG_CACHE[name] = _build_config(name)
    p, ks = _CONFIG_CACHE[name]
    return dataclasses.replace(p, V_H=p.V_H.copy(), psi_lm=p.psi_lm.copy(), xyz=p.xyz.copy(), Z=p.Z.copy(),
                               orb_energies=p.orb_energies.copy()), ks.copy()


def _build_config(name: str) -> tuple[Problem, np.ndarray]:
    if name == "h2_l8":          # CPU-runnable plain case
        return h2_like(8, 0.01, 30.0, seed=0), k_grid(0.1, 1.0, 10)
    if name == "h2o_l20":        # plain per-ik on one GPU
        return h2o_like(20, 0.01, 50.0, 5, seed=1, allow_closed=True), k_grid(0.1, 2.0, 100)
    if name == "h2o_l30":        # batched scan, closed channels allowed
        return h2o_like(30, 0.01, 60.0, 5, seed=1, allow_closed=True), k_grid(0.05, 2.0, 256)
    if name == "h2o_l20_pol":    # adjoint / polarisation case
        return h2o_like(20, 0.01, 50.0, 5, seed=1, pol=(9.8, 1.5)), k_grid(0.2, 1.5, 64)
    if name == "ring_l60":       # large-molecule scale: 16-centre ring, l_cont = 60
        return ring_like(60, 0.02, 100.0, 60, seed=4), k_grid(0.6, 1.6, 32)
    raise KeyError(name)
