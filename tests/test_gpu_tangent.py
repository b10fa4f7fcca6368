"""GPU parity of the scan's forward-mode derivative (sehf_scan_run_tangent)
against the oracle's tangent on the same inputs.

The oracle's tangent is itself pinned on the CPU by Richardson-extrapolated
finite differences and by the analytic-vs-FD Wigner delay
(tests/test_oracle_tangent.py).  Here the device path must reproduce it:

* dK, dA, dB, db_overlap, dd_raw and the reduced-dipole derivative dD along
  E, along the polarisation strength alpha, and along a mixed direction, on
  the small H2O-like case with the soft polarisation potential, closed
  channels allowed;
* the bound is max(1e-8, 4 x the tangent's rounding noise), the noise being
  the larger distance of two other valid evaluations of the oracle's tangent
  from the reference's (Gauss-Jordan inverses with the fit sums from the
  outer end of the window; correctly rounded inverses), relative (max-abs
  error over max-abs value).  The tangent inherits the forward outputs'
  rounding floor amplified by the conditioning of the ratio recursion: the
  noise is 1e-9 .. 2e-6 on these cases, and it changes from host to host
  (the last bit of a few libm results differs between CPU models, which
  moves dK by up to 5e-8 here), so a single realisation is no bound; a
  dropped term (all of which are at least O(h^2/6) ~ 7e-5 relative) cannot
  hide under it;
* dA and dB are measured as one pair (the derivative of the fitted amplitude
  psi ~ A j + B y), normalised by the pair's max-abs: the fit keeps A within
  ~1e-4 of the identity, so max|dA| is ~1e-3 of max|dB| and a per-matrix
  measure of dA would weigh the rounding of dB-sized terms against a
  near-zero scale;
* the forward outputs of a tangent run are bitwise those of a plain run with
  back-propagated dipoles (tangent runs always back-propagate);
* a case with the exchange cut-off inside the grid (n_transition < Nr), where
  W^{-1} switches to block-diagonal form, and the plain CPU-runnable
  configuration (l_cont = 8, 3001 points) along E.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

KEYS = ("dK", "dA", "dB", "db_overlap", "dd_raw", "dD")


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


def _pol_problem():
    from sehf import synth
    return synth.h2o_like(l_cont=3, dr=0.02, r_max=14.0, n_occ=3, seed=1, l_orb_sce=3, pol=(9.8, 1.5),
                          allow_closed=True)


def _errs(got, ref):
    e = {key: rel_err(got[key], ref[key]) for key in KEYS}
    scale = max(np.abs(ref["dA"]).max(), np.abs(ref["dB"]).max(), 1e-300)
    for key in ("dA", "dB"):
        e[key] = float(np.abs(np.asarray(got[key]) - ref[key]).max() / scale)
    return e


def _oracle_with_floor(O, p, k, direction):
    """The oracle's tangent and the bound max(1e-8, 4 x its rounding noise)
    per output, the noise being the larger distance of the two other
    evaluations (rounding-probe modes 3 and 32) from it."""
    a = O.solve(p, k, tangent=direction)
    L = O.lib()
    noise = {key: 0.0 for key in KEYS}
    for mode in (3, 32):
        L.sehf_oracle_set_inverse_mode(mode)
        try:
            b = O.solve(p, k, tangent=direction)
        finally:
            L.sehf_oracle_set_inverse_mode(0)
        for key, v in _errs(b, a).items():
            noise[key] = max(noise[key], v)
    bound = {key: max(1e-8, 4.0 * noise[key]) for key in KEYS}
    return a, bound


def _check(got, ref, bound, label):
    errs = _errs(got, ref)
    print(f"  {label}: " + " ".join(f"{k}={v:.1e}(<{bound[k]:.1e})" for k, v in errs.items()))
    for key in KEYS:
        assert errs[key] < bound[key], (label, key, errs[key], bound[key])


def test_tangent_directions_pol(ctx):
    import oracle_py as O
    from sehf import scan
    p = _pol_problem()
    ks = np.array([0.7, 0.9, 1.1, 0.7, 0.9])
    dirs = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 0.0], [0.0, 1.0], [0.5, -2.0]])
    S = scan.Scan(p, ctx)
    got = S.run(ks, tangent=dirs)
    # tangent runs back-propagate the dipoles: their forward outputs are those
    # of a plain back-propagated run
    S.set_dipole_path(scan.Scan.DIPOLES_BACKWARD)
    plain = S.run(ks)
    for e, (k, d) in enumerate(zip(ks, dirs)):
        ref, bound = _oracle_with_floor(O, p, float(k), tuple(d))
        _check(got[e], ref, bound, f"k={k} dir={tuple(d)}")
        for key in ("K", "A", "B", "K_fit", "b_overlap", "d_raw", "D"):
            assert np.array_equal(got[e][key], plain[e][key]), (key, k)


def test_tangent_batch_split_bitwise(ctx):
    """Energies are independent in the tangent too: the batch split and the
    concurrent groups change no bit."""
    from sehf import scan
    p = _pol_problem()
    ks = np.array([0.6, 0.8, 1.0, 1.2])
    S = scan.Scan(p, ctx)
    one = S.run(ks, batch=4, groups=1, tangent=(1.0, 0.5))
    many = S.run(ks, batch=1, groups=2, tangent=(1.0, 0.5))
    for a, b in zip(one, many):
        for key in KEYS:
            assert np.array_equal(a[key], b[key]), (key, a["k"])


def test_tangent_exchange_cutoff_inside_grid(ctx):
    """n_transition < Nr: beyond it W^{-1} = blockdiag(Sinv, D^{-1}) and the
    tangent's G = W^{-1}[:, psi] loses its f block."""
    import oracle_py as O
    from sehf import scan, synth
    p = synth.h2_like(l_cont=2, dr=0.05, r_max=45.0, seed=0)
    S = scan.Scan(p, ctx)
    assert S.info["n_transition"] < p.Nr, S.info
    ks = np.array([0.5, 0.9])
    got = S.run(ks, tangent=(1.0, 0.0))
    for e, k in enumerate(ks):
        ref, bound = _oracle_with_floor(O, p, float(k), (1.0, 0.0))
        _check(got[e], ref, bound, f"h2 cutoff k={k}")


def test_tangent_h2_l8_full_energy(ctx):
    """The plain CPU-runnable configuration (l_cont = 8, dr = 0.01, 3001
    points) along E at one of its energies (the oracle's tangent takes ~20 s
    per rounding order here)."""
    import oracle_py as O
    from sehf import scan, synth
    p, ks_all = synth.config("h2_l8")
    ks = ks_all[[4]]
    got = scan.Scan(p, ctx).run(ks, tangent=(1.0, 0.0))
    for e, k in enumerate(ks):
        ref, bound = _oracle_with_floor(O, p, float(k), (1.0, 0.0))
        _check(got[e], ref, bound, f"h2_l8 k={k:.2f}")


def test_tangent_h2o_l8_midsize(ctx):
    """The derivative along E and alpha on the mid-size system (N_tot = 324:
    128 x 128 GEMM tiles in the ratio-matrix tangents and the
    back-propagation products), one energy, mixed direction."""
    import oracle_py as O
    from sehf import scan, synth
    p = synth.h2o_like(l_cont=8, dr=0.02, r_max=10.0, n_occ=3, seed=1, l_orb_sce=6, pol=(9.8, 1.5),
                       allow_closed=True)
    got = scan.Scan(p, ctx).run(np.array([0.85]), tangent=(1.0, 0.25))
    ref, bound = _oracle_with_floor(O, p, 0.85, (1.0, 0.25))
    _check(got[0], ref, bound, "h2o_l8 k=0.85")
