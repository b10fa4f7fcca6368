"""Device storage modes of the scan (sehf_plan_storage's decisions).

* Streamed tables (SEHF_TABLES_STREAMED): S0_n and B_n D_n^{-1} rebuilt per
  radial step on the side stream give every output of plain, back-propagated
  and tangent runs bitwise equal to the resident tables, on a case with the
  exchange cut-off inside the grid and on a mid-size closed-channel case.
* Host checkpoints (SEHF_CKPT_HOST): ratio-matrix checkpoints in pinned host
  memory give bitwise the same outputs as checkpoints in HBM (the tangent and
  back-propagated runs read every checkpoint).
* The handle's device bytes: streamed tables hold a fraction of the resident
  ones, and the batch buffers equal the planner's count for the same batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("K", "A", "B", "K_fit", "b_overlap", "d_raw", "D", "diag")
TKEYS = ("dK", "dA", "dB", "db_overlap", "dd_raw", "dD")


@pytest.fixture(scope="module")
def ctx():
    from sehf.steppers import GpuContext
    return GpuContext(True)


def _cases():
    from sehf import synth
    return {
        "h2_cutoff": (synth.h2_like(l_cont=2, dr=0.05, r_max=45.0, seed=0), np.array([0.5, 0.9])),
        "h2o_l8": (synth.h2o_like(l_cont=8, dr=0.02, r_max=10.0, n_occ=3, seed=1, l_orb_sce=6, allow_closed=True),
                   np.array([0.55, 1.05])),
        "h2o_l4_pol": (synth.h2o_like(l_cont=4, dr=0.02, r_max=14.0, n_occ=3, seed=1, pol=(9.8, 1.5)),
                       np.array([0.5, 1.3])),
    }


def _same(a_list, b_list, keys, label):
    for a, b in zip(a_list, b_list):
        for key in keys:
            assert np.array_equal(a[key], b[key]), (label, key, a["k"])


@pytest.mark.parametrize("case", ["h2_cutoff", "h2o_l8", "h2o_l4_pol"])
def test_streamed_tables_bitwise(ctx, case):
    from sehf import scan
    p, ks = _cases()[case]
    R = scan.Scan(p, ctx, tables="resident")
    S = scan.Scan(p, ctx, tables="streamed")
    mr, ms = R.memory(), S.memory()
    print(case, "shared bytes resident", mr["shared"], "streamed", ms["shared"])
    assert ms["shared"] < mr["shared"]
    for path in (scan.Scan.DIPOLES_FORWARD, scan.Scan.DIPOLES_BACKWARD):
        R.set_dipole_path(path)
        S.set_dipole_path(path)
        _same(R.run(ks), S.run(ks), KEYS, f"{case} path {path}")
    tr = R.run(ks, tangent=(1.0, 0.5))
    ts = S.run(ks, tangent=(1.0, 0.5))
    _same(tr, ts, KEYS + TKEYS, f"{case} tangent")
    # a segment cycle reads the segment's rebuilt tables too
    lay = S.layout()
    R.run_segment(ks, lay["n_seg"] // 2)
    S.run_segment(ks, lay["n_seg"] // 2)
    R.close()
    S.close()


def test_host_checkpoints_bitwise(ctx):
    from sehf import scan
    p, ks = _cases()["h2o_l4_pol"]
    S = scan.Scan(p, ctx)
    S.set_dipole_path(scan.Scan.DIPOLES_BACKWARD)
    dev = S.run(ks)
    dev_t = S.run(ks, tangent=(1.0, 0.0))
    assert S.memory()["host"] == 0
    S.set_checkpoints("host")
    host = S.run(ks)
    host_t = S.run(ks, tangent=(1.0, 0.0))
    m = S.memory()
    print("host checkpoint bytes", m["host"])
    assert m["host"] > 0
    _same(dev, host, KEYS, "host checkpoints")
    _same(dev_t, host_t, KEYS + TKEYS, "host checkpoints, tangent")
    S.close()


def test_batch_bytes_match_the_plan(ctx):
    """The planner counts the scan's batch buffers through the same buffer
    list the scan allocates from: for the batch and spacing it plans, the
    handle holds exactly the planned bytes, up to the analytic-start table,
    which the planner sizes for n_start = 1 ((n_start - 1) x batch x N_tot
    doubles more in the scan)."""
    from sehf import scan
    p, ks = _cases()["h2o_l8"]
    plan = scan.plan_storage(p, len(ks), device_bytes=200 << 30)
    print(plan["report"])
    assert plan["tables"] == "resident" and plan["batch"] == len(ks)
    S = scan.Scan(p, ctx, tables=plan["tables"])
    S.run(ks, batch=plan["batch"], seg_len=plan["seg_len"])
    held = S.memory()["batch"]
    ns = S.info["n_start"]
    S.close()
    B = plan["batch"]
    expect = B * plan["per_energy_bytes"] + 8 * (ns - 1) * B * p.n_tot
    print("held", held, "expected", expect, "n_start", ns)
    assert held == expect
