"""Host logic of bench.py, on the CPU: the one-GPU guard and the HBM-traffic arithmetic
that turns two rocprofv3 --pmc counter files into bytes per Heff application."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def write_counters(path, counter, rows):
    # rocprofv3 counter_collection.csv columns that the parser reads
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for d, (name, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": d + 1, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": v})
            # another counter in the same file is ignored
            w.writerow({"Dispatch_Id": d + 1, "Kernel_Name": name, "Counter_Name": "GRBM_COUNT", "Counter_Value": 7})


@pytest.mark.parametrize("fetch_per_gib,write_per_gib", [(2.0, 1.0), (1.0, 1.0), (0.5, 1.25)])
def test_traffic_numbers(tmp_path, fetch_per_gib, write_per_gib):
    # a counter that reads fetch_per_gib GiB for the 1 GiB copy mark is scaled by its
    # inverse; only the dispatches between the two largest marks count, per application
    gib_kib = bench.COPY_MARK_DOUBLES * 8 / 1024.0
    steps = 3
    kread = [100.0, 250.0, 50.0] * steps       # KiB as the counter reads them
    kwrite = [40.0, 10.0, 5.0] * steps
    small_mark = ("elementwise_kernel small", 12.0)  # a torch copy before the marks: not a bracket
    fetch = [small_mark, ("copyBuffer", gib_kib * fetch_per_gib)] + [("k_tiles_fit", v) for v in kread] + \
            [("elementwise_kernel", gib_kib * fetch_per_gib), ("k_after", 1e9)]
    write = [small_mark, ("copyBuffer", gib_kib * write_per_gib)] + [("k_tiles_fit", v) for v in kwrite] + \
            [("elementwise_kernel", gib_kib * write_per_gib), ("k_after", 1e9)]
    fp, wp = str(tmp_path / "fetch.csv"), str(tmp_path / "write.csv")
    write_counters(fp, "FETCH_SIZE", fetch)
    write_counters(wp, "WRITE_SIZE", write)
    out = bench.traffic_numbers(fp, wp, steps)
    rd = sum(kread) / fetch_per_gib * 1024.0 / steps
    wr = sum(kwrite) / write_per_gib * 1024.0 / steps
    assert out["read_bytes"] == pytest.approx(rd, rel=1e-12)
    assert out["write_bytes"] == pytest.approx(wr, rel=1e-12)
    assert out["bytes_per_launch"] == pytest.approx(rd + wr, rel=1e-12)
    assert out["fetch_scale"] == pytest.approx(1.0 / fetch_per_gib, abs=1e-4)
    assert out["dispatches_per_application"] == 3
    assert out["applications"] == steps


def test_roofline_record_matrix_core_bound():
    # the middle-site product of the L=32 D=2000 benchmark as measured on one MI355X:
    # 16.7922 ms of kernels for F_H = 2.2925e11 flops over B_H = 7.2768e8 algorithmic bytes
    r = bench.roofline_record(229251582970.0, 727681104.0, 16.7922, 54519203899.42752, "profiles/x.json")
    assert r["bound"] == "mfma"           # 315 flop/B x 8 TB/s is far above 78.6 TFLOP/s
    assert r["peak"] == 78.6
    assert r["achieved"] == 13.6523
    assert r["frac"] == 0.1737
    assert r["traffic_over_B_H"] == 74.92
    assert r["unit"] == "TFLOP/s" and r["traffic_source"] == "profiles/x.json"


def test_roofline_record_hbm_bound():
    # 1 flop per byte: the bound is 8 TFLOP/s of HBM, and 1 TFLOP/s is 1/8 of it
    r = bench.roofline_record(1e9, 1e9, 1.0, None, None)
    assert r["bound"] == "hbm"
    assert r["peak"] == 8.0
    assert r["achieved"] == 1.0
    assert r["frac"] == 0.125
    assert r["traffic_over_B_H"] is None
