"""The 16-centre ring generator (the large-molecule configuration's
inputs), on a reduced size: its Hartree potential from Gaunt-contracted
component pairs equals the dense angular-quadrature build of the other
configurations, the orbitals are orthonormal, the monopole carries the
electron count, and the geometry is the specified one."""
import math

import numpy as np

from sehf import synth


def test_ring_hartree_matches_dense_build():
    p = synth.ring_like(l_cont=4, dr=0.02, r_max=16.0, n_occ=5, seed=4, l_sce=4)
    dense = synth.hartree_from_orbitals(p.psi_lm, p.n_occ, 4, 2 * p.l_cont, p.dr)
    err = np.max(np.abs(dense - p.V_H)) / np.max(np.abs(dense))
    print("V_H ring vs dense", err)
    assert err < 1e-12


def test_ring_orbitals_and_geometry():
    p = synth.ring_like(l_cont=8, dr=0.02, r_max=20.0, n_occ=6, seed=4, l_sce=8)
    Nr = p.Nr
    r = np.arange(Nr) * p.dr
    w = r * r * synth._simpson_weights(Nr, p.dr)
    orbs = np.concatenate([p.psi_lm, p.init_psi_lm[None]])
    S = np.einsum("ilr,jlr,r->ij", orbs, orbs, w)
    assert np.max(np.abs(S - np.eye(len(orbs)))) < 1e-12
    # the largest l present is l_sce
    lam = np.nonzero(np.any(orbs != 0.0, axis=(0, 2)))[0]
    assert int(math.isqrt(int(lam.max()))) == 8
    # V_H monopole -> N_el sqrt(4 pi) / r far out
    assert abs(p.V_H[0, -1] * r[-1] - 2 * p.n_occ * math.sqrt(4 * math.pi)) < 1e-6
    rad = np.hypot(p.xyz[:, 0], p.xyz[:, 1])
    assert np.allclose(rad[:8], 2.7) and np.allclose(rad[8:], 5.2) and np.all(p.xyz[:, 2] == 0.0)
    assert list(p.Z) == [6] * 8 + [9] * 8
    assert np.allclose((p.Z[:, None] * p.xyz).sum(axis=0), 0.0, atol=1e-12)
