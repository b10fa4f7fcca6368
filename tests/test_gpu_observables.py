"""Observables of a GPU scan (torch on the device) against the numpy/scipy
restatement of cross_section_delay.py applied to the CPU oracle's dipoles:
sigma, beta and Q_sigma within 1e-8 relative, tau within 1e-8 of its scale."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_observables_h2o_l4():
    import oracle_py as O
    import observables_ref as R
    from sehf import scan, synth, observables
    from sehf.steppers import GpuContext
    p = synth.h2o_like(l_cont=4, dr=0.02, r_max=14.0, n_occ=3, seed=1)
    ks = np.linspace(0.5, 1.3, 6)
    got = scan.Scan(p, GpuContext(True)).run(ks, batch=6)
    E0 = float(p.orb_energies[p.n_occ - 1])
    obs = observables.from_scan(got, E0, device="cuda")
    assert obs["sigma_L"].device.type == "cuda"
    ref = O.solve_scan(p, ks)
    om = 0.5 * ks * ks - E0
    D = np.stack([r["D"] for r in ref])
    DL, DV = D[:, 0:3].reshape(len(ks), -1), D[:, 3:6].reshape(len(ks), -1)
    want = {"sigma_L": R.sigma_from_dipole(ks, DL, om, "length"),
            "sigma_V": R.sigma_from_dipole(ks, DV, om, "velocity"),
            "q_sigma": R.q_sigma(DL, DV, om),
            "beta_L": R.compute_beta(D[:, 0], D[:, 1], D[:, 2], list(range(p.n_psi))),
            "beta_V": R.compute_beta(D[:, 3], D[:, 4], D[:, 5], list(range(p.n_psi))),
            "tau_L": R.time_delay(ks, DL), "tau_V": R.time_delay(ks, DV)}
    for key, w in want.items():
        g = obs[key].cpu().numpy()
        err = np.abs(g - w).max() / max(np.abs(w).max(), 1e-300)
        print(key, f"{err:.2e}")
        assert err < 1e-8, (key, err)


def test_analytic_delay_from_tangent_run():
    """from_scan on a tangent run along E gives the delay from the analytic
    dd/dE; against the same formula on the oracle's tangent, within 1e-6 of
    max(1, |tau|) (the derivative's parity bound carried through tau)."""
    import oracle_py as O
    from sehf import scan, synth, observables
    from sehf.steppers import GpuContext
    p = synth.h2o_like(l_cont=4, dr=0.02, r_max=14.0, n_occ=3, seed=1)
    ks = np.array([0.6, 0.95, 1.25])
    got = scan.Scan(p, GpuContext(True)).run(ks, batch=3, tangent=(1.0, 0.0))
    obs = observables.from_scan(got, float(p.orb_energies[p.n_occ - 1]), device="cuda")
    for e, k in enumerate(ks):
        ref = O.solve(p, float(k), tangent=(1.0, 0.0))
        for gauge, sl in (("L", slice(0, 3)), ("V", slice(3, 6))):
            D, dD = ref["D"][sl].ravel(), ref["dD"][sl].ravel()
            want = (np.conj(D) * dD).sum().imag / (np.abs(D) ** 2).sum()
            g = float(obs[f"tau_{gauge}_analytic"][e])
            print(gauge, k, g, want)
            assert abs(g - want) < 1e-6 * max(1.0, abs(want)), (gauge, k, g, want)
