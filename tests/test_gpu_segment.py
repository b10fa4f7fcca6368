"""Segment cycles (sehf_scan_run_segment), the benchmark's timed step.

A segment cycle runs one radial segment of each of the scan's three sweeps
on the scan handle's batch buffers.  It must leave every later whole scan
on the same handle bitwise unchanged (each run starts from its own
checkpoints, boundary and accumulators), report a timing for each part, and
reject segments outside the grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


def test_segment_cycles_leave_scans_bitwise_unchanged(ctx):
    from sehf import scan, synth
    from sehf._lib import SehfError
    p = synth.h2o_like(l_cont=4, dr=0.02, r_max=14.0, n_occ=3, seed=1)
    ks = np.array([0.4, 0.8, 1.2])
    S = scan.Scan(p, ctx)
    ref = S.run(ks, batch=3, groups=1)
    ref_t = S.run(ks, batch=3, groups=1, tangent=(1.0, 0.3))
    lay = S.layout()
    print("layout", lay)
    assert lay["L"] >= 2 and lay["n_seg"] * lay["L"] >= p.Nr - 1 > (lay["n_seg"] - 1) * lay["L"]
    assert lay["prepass_seg_lo"] == lay["n_fit_lo"] // lay["L"]
    for seg in (0, 1, lay["n_seg"] // 2, lay["n_seg"] - 1):
        t = S.run_segment(ks, seg)
        assert all(v >= 0.0 for v in t.values()) and t["forward"] + t["recompute"] + t["back"] > 0.0
        t = S.run_segment(ks, seg, tangent=(1.0, 0.3))
        assert t["back"] > 0.0
        print(f"segment {seg}: {t}")
    with pytest.raises(SehfError, match="outside"):
        S.run_segment(ks, lay["n_seg"])
    got = S.run(ks, batch=3, groups=1)
    got_t = S.run(ks, batch=3, groups=1, tangent=(1.0, 0.3))
    for a, b in zip(got, ref):
        for key in ("K", "A", "B", "K_fit", "b_overlap", "d_raw", "D", "diag"):
            assert np.array_equal(a[key], b[key]), key
    for a, b in zip(got_t, ref_t):
        for key in ("K", "D", "dK", "dD", "dd_raw"):
            assert np.array_equal(a[key], b[key]), key
    S.close()
