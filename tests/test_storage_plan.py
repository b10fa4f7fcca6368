"""The device storage planner (sehf_plan_storage), on the CPU: the policy
the reference's planner tests hold for its RAM/disk placement
(src/tests/test_storage_planner.cpp), restated for HBM.

* Everything fits: resident tables, device checkpoints, the default
  spacing ceil(sqrt(Nr - 1)), batch = min(energies, cap).
* Resident tables are kept while they cost no batch width; when streamed
  tables widen the batch (the l_cont = 30 tangent run), they are chosen.
* When the default spacing leaves no batch at all, the spacing is shortened
  (forward dipoles keep checkpoints only over the top segments) and then the
  checkpoints move to host memory (the l_cont = 60 configuration).
* The budget is (device - runtime) x (1 - reserve); a budget that holds no
  energy fails with the report; bad inputs fail.
"""
import types

import pytest

from sehf import scan

GB = 10 ** 9


def sizes(l_cont, n_occ, Nr):
    le = min(l_cont, 10)
    return types.SimpleNamespace(Nr=Nr, n_psi=(l_cont + 1) ** 2, n_f=n_occ * (le + 1) ** 2, l_cont=l_cont,
                                 n_occ=n_occ, n_lambda=(l_cont + le + 1) ** 2)


def test_everything_fits():
    p = sizes(8, 1, 3001)
    pl = scan.plan_storage(p, 10, device_bytes=288 * GB)
    assert pl["tables"] == "resident" and pl["checkpoints"] == "device"
    assert pl["seg_len"] == 55 and pl["batch"] == 10
    pl = scan.plan_storage(p, 1000, device_bytes=288 * GB)
    assert pl["batch"] == 64                       # the default cap
    pl = scan.plan_storage(p, 1000, device_bytes=288 * GB, batch_cap=200)
    assert pl["batch"] == 200


def test_metric_configuration_resident():
    pl = scan.plan_storage(sizes(30, 5, 6001), 256, device_bytes=288 * GB)
    print(pl["report"])
    assert pl["tables"] == "resident" and pl["checkpoints"] == "device" and pl["batch"] == 64
    assert pl["seg_len"] == 78 and pl["n_ckpt"] == 3   # forward dipoles: the top segments only
    assert pl["shared_bytes"] > 70 * GB


def test_streamed_tables_when_they_widen_the_batch():
    pl = scan.plan_storage(sizes(30, 5, 6001), 256, tangent=True, device_bytes=288 * GB)
    print(pl["report"])
    assert pl["tables"] == "streamed"
    assert pl["n_ckpt"] == 77                           # tangent runs back-propagate: every segment
    res = scan.plan_storage(sizes(30, 5, 6001), 256, tangent=True, device_bytes=288 * GB, batch_cap=25)
    assert res["tables"] == "resident" and res["batch"] == 25   # resident already reaches the wanted width


def test_large_configuration_shortens_spacing_then_uses_host():
    p = sizes(60, 60, 5001)
    dev_only = scan.plan_storage(p, 32, device_bytes=288 * GB)
    print(dev_only["report"])
    assert dev_only["tables"] == "streamed" and dev_only["checkpoints"] == "device"
    assert dev_only["seg_len"] < 71 and dev_only["batch"] >= 1
    with_host = scan.plan_storage(p, 32, device_bytes=288 * GB, host_bytes=200 * GB)
    print(with_host["report"])
    assert with_host["checkpoints"] == "host" and with_host["batch"] > dev_only["batch"]
    assert with_host["host_per_energy_bytes"] * with_host["batch"] <= 200 * GB
    tan = scan.plan_storage(p, 32, tangent=True, device_bytes=288 * GB, host_bytes=200 * GB)
    assert tan["checkpoints"] == "host" and tan["batch"] == 1


def test_budget_arithmetic():
    p = sizes(8, 1, 3001)
    pl = scan.plan_storage(p, 4, device_bytes=100 * GB, runtime_bytes=10 * GB, reserve_fraction=0.2)
    assert pl["budget_bytes"] == int(90 * GB * 0.8)
    pl0 = scan.plan_storage(p, 4, device_bytes=100 * GB, runtime_bytes=10 * GB, reserve_fraction=0.0)
    assert pl0["budget_bytes"] == 90 * GB


def test_no_fit_and_bad_inputs_fail():
    with pytest.raises(RuntimeError, match="no placement fits one energy"):
        scan.plan_storage(sizes(30, 5, 6001), 8, device_bytes=2 * GB, runtime_bytes=0)
    with pytest.raises(RuntimeError, match="reserve_fraction"):
        scan.plan_storage(sizes(8, 1, 3001), 4, device_bytes=GB, reserve_fraction=0.95)
    with pytest.raises(RuntimeError, match="sizes out of range"):
        scan.plan_storage(sizes(8, 1, 3001), 0, device_bytes=GB)
