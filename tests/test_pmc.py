"""The rocprofv3 summary tool (ib_parallel_amd/pmc.py) on small hand-made
CSV runs whose answers are worked out by hand here: the copy calibration of
FETCH_SIZE and WRITE_SIZE, the restriction to the dispatches after the
calibration copy, the sum over counter dimensions within one dispatch, the
FP64 operation count and the kernel-trace averages."""
import argparse
import csv
import json
import os

import pytest

from ib_parallel_amd import pmc

KA = "void ibg::k_alpha(ibg::Args)"
KB = "void ibg::k_beta<2, true>(int, double const*)"
COPY = "__amd_rocclr_copyBuffer"
COPY_MB = 1.0  # 1024 KiB


def write_counters(folder, rows):
    os.makedirs(folder)
    with open(os.path.join(folder, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                    "End_Timestamp"])
        for d, name, counter, value, t0, t1 in rows:
            w.writerow([d, name, counter, value, t0, t1])


@pytest.fixture
def runs(tmp_path):
    # dispatch 2: a small status copy; 3: k_alpha during set-up (excluded);
    # 5: the calibration copy; 6-9: the timed dispatches.  Dispatch 7's
    # FETCH_SIZE comes in two dimensions (60 + 40).
    fetch = [
        (2, COPY, "FETCH_SIZE", 1.0, 0, 10),
        (3, KA, "FETCH_SIZE", 999.0, 20, 30),
        (5, COPY, "FETCH_SIZE", 512.0, 40, 1040),
        (7, KA, "FETCH_SIZE", 60.0, 1100, 1200),
        (7, KA, "FETCH_SIZE", 40.0, 1100, 1200),
        (9, KA, "FETCH_SIZE", 300.0, 1300, 1400),
        (8, KB, "FETCH_SIZE", 10.0, 1250, 1260),
    ]
    write = [
        (2, COPY, "WRITE_SIZE", 1.0, 0, 10),
        (3, KA, "WRITE_SIZE", 777.0, 20, 30),
        (5, COPY, "WRITE_SIZE", 1024.0, 40, 1040),
        (7, KA, "WRITE_SIZE", 50.0, 1100, 1200),
        (9, KA, "WRITE_SIZE", 50.0, 1300, 1400),
        (8, KB, "WRITE_SIZE", 4.0, 1250, 1260),
    ]
    f64 = []
    for counter, a7, a9 in (("SQ_INSTS_VALU_ADD_F64", 1, 3), ("SQ_INSTS_VALU_MUL_F64", 2, 2),
                            ("SQ_INSTS_VALU_TRANS_F64", 0, 1), ("SQ_INSTS_VALU_FMA_F64", 3, 5),
                            ("SQ_INSTS_VALU_MFMA_MOPS_F64", 4, 0)):
        f64 += [(3, KA, counter, 1000, 20, 30), (5, COPY, counter, 0, 40, 1040),
                (7, KA, counter, a7, 1100, 1200), (9, KA, counter, a9, 1300, 1400)]
    write_counters(str(tmp_path / "fetch"), fetch)
    write_counters(str(tmp_path / "write"), write)
    write_counters(str(tmp_path / "f64"), f64)
    return tmp_path


def test_traffic_is_calibrated_by_the_copy_and_restricted_to_the_timed_dispatches(runs, capsys):
    out = str(runs / "traffic.json")
    pmc.traffic(argparse.Namespace(fetch_dir=str(runs / "fetch"), write_dir=str(runs / "write"), copy_mb=COPY_MB,
                                   f64=str(runs / "f64"), json=out))
    with open(out) as fh:
        tr = json.load(fh)
    # FETCH_SIZE of the 1 MiB copy reads 512 KiB: factor 2; WRITE_SIZE 1024: factor 1
    assert tr["copy"]["scale_fetch"] == pytest.approx(2.0)
    assert tr["copy"]["scale_write"] == pytest.approx(1.0)
    a = tr["kernels"]["k_alpha"]
    # dispatches 7 (60 + 40) and 9 (300); dispatch 3 before the copy is not counted
    assert a["dispatches"] == 2
    assert a["fetch_kib"] == pytest.approx(200.0)
    assert a["write_kib"] == pytest.approx(50.0)
    assert a["bytes_per_launch"] == pytest.approx((200.0 * 2 + 50.0) * 1024)
    # 64 (ADD + MUL + TRANS + 2 FMA) + 512 MFMA_MOPS, averaged over 7 and 9
    ops7 = 64 * (1 + 2 + 0 + 2 * 3) + 512 * 4
    ops9 = 64 * (3 + 2 + 1 + 2 * 5) + 512 * 0
    assert a["fp64_ops_per_launch"] == pytest.approx((ops7 + ops9) / 2)
    b = tr["kernels"]["k_beta<2, true>"]
    assert b["bytes_per_launch"] == pytest.approx((10.0 * 2 + 4.0) * 1024)
    assert "k_alpha" in capsys.readouterr().out


def write_trace(folder, rows):
    os.makedirs(folder)
    with open(os.path.join(folder, "run_kernel_trace.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in rows:
            w.writerow(r)


def test_stats_after_the_copy(tmp_path, capsys):
    write_trace(str(tmp_path / "trace"), [
        (1, KA, 0, 9000),             # set-up: excluded
        (2, COPY, 10000, 10100),      # a small copy
        (3, COPY, 20000, 90000),      # the calibration copy (the longest)
        (4, KA, 100000, 102000),
        (5, KB, 103000, 103500),
        (6, KA, 104000, 108000),
    ])
    pmc.stats(argparse.Namespace(trace_dir=str(tmp_path / "trace"), after_copy=True, first=None, until=None, top=10))
    lines = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("| k_")]
    # k_alpha: 2 and 4 us -> 2 calls, 6 us total, 3 us mean, share 6 / 6.5
    assert lines[0].split("|")[1:8] == [" k_alpha ", " 2 ", " 0.006 ", " 3.00 ", " 2.00 ", " 4.00 ", " 0.923 "]
    assert lines[1].split("|")[1:4] == [" k_beta<2, true> ", " 1 ", " 0.001 "]


def test_stats_between_two_kernels(tmp_path, capsys):
    write_trace(str(tmp_path / "trace"), [
        (1, KB, 0, 1000),
        (2, KA, 2000, 3000),          # the window opens at the first k_alpha
        (3, KA, 4000, 7000),
        (4, KB, 8000, 8500),          # ... and closes before this k_beta
        (5, KA, 9000, 19000),
    ])
    pmc.stats(argparse.Namespace(trace_dir=str(tmp_path / "trace"), after_copy=False, first="k_alpha",
                                 until="k_beta<2, true>", top=10))
    lines = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("| k_")]
    assert len(lines) == 1
    assert lines[0].split("|")[1:5] == [" k_alpha ", " 2 ", " 0.004 ", " 2.00 "]
