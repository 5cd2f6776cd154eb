"""Host logic of bench.py that needs no GPU: the one-GPU rule and the
profile summaries that `roofline.traffic` and the kernel-time tables come
from (rocprofv3 CSV in, scaled bytes and tables out)."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "single MI355X" in r.stderr and r.stdout == ""


def _write_pmc(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                                          "End_Timestamp"])
        w.writeheader()
        for name, value in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": value,
                        "Start_Timestamp": 0, "End_Timestamp": 1000})


def test_traffic_summary_scales_by_the_copy_kernel(tmp_path):
    """FETCH_SIZE / WRITE_SIZE are KiB per dispatch; each is scaled by what
    the plain copy of known size reads or writes, so a copy that reports
    half its bytes (fetch) doubles the traversal's fetch."""
    copy = 1 << 30
    trace = "void dl::k_trace_queue<false>(dl::BvhView, dl::RenderBufs, int*)"
    _write_pmc(str(tmp_path / "f"), "FETCH_SIZE", [(trace, 1000.0), (trace, 3000.0), ("k_copy", copy / 2048.0)])
    _write_pmc(str(tmp_path / "w"), "WRITE_SIZE", [(trace, 500.0), ("k_copy", copy / 1024.0)])
    out = tmp_path / "traffic.json"
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"), str(out), copy_bytes=copy, commands="cmd")
    t = json.loads(out.read_text())
    assert t["dispatches"] == 2
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    # mean fetch 2000 KiB x 2, write 500 KiB x 1
    assert t["hbm_bytes_per_launch"] == pytest.approx(2000 * 1024 * 2.0 + 500 * 1024 * 1.0)


def test_kernel_time_table(tmp_path):
    d = tmp_path / "prof"
    d.mkdir()
    with open(d / "bench_kernel_stats.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs", "MinNs", "MaxNs",
                                          "Percentage"])
        w.writeheader()
        w.writerow({"Name": "void dl::k_trace_queue<false>(dl::BvhView, dl::RenderBufs, int*)", "Calls": 4,
                    "TotalDurationNs": 8e6, "AverageNs": 2e6, "MinNs": 1e6, "MaxNs": 3e6, "Percentage": 80.0})
    out = tmp_path / "stats.md"
    bench.summarize_stats(str(d), str(out), "cmd", "title")
    text = out.read_text()
    assert "# Kernel time: title" in text and "```\ncmd\n```" in text
    assert "| k_trace_queue<false> | 4 | 8.00 | 2000.0 | 1000.0 | 3000.0 | 80.00 |" in text
