"""CPU checks of the measurement tools: bench.py's one-GPU contract, and the rocprofv3 summary
arithmetic that turns counter readings into the bytes bench.py reports as roofline.traffic."""
import csv
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single MI355X" in r.stderr
    assert r.stdout == ""  # no JSON line


def _write_csv(path: Path, rows, fields):
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_profile_summary_calibration(tmp_path):
    """The record-split copy moves 32 bytes per particle each way; a kernel whose FETCH_SIZE reads
    twice the copy's fetched twice its bytes, and one whose WRITE_SIZE reads a quarter of the
    copy's wrote a quarter.  Durations come from each kernel's last N dispatches only."""
    from watercube_amd import profile_summary as ps

    n = 64000
    copy = 32.0 * n
    trace = [{"Kernel_Name": "k_split", "Start_Timestamp": 0, "End_Timestamp": 1000}]
    t = 2000
    for d in (100, 200, 300):  # the last two average 250 ns
        trace.append({"Kernel_Name": "k_density", "Start_Timestamp": t, "End_Timestamp": t + d})
        t += 1000
    _write_csv(tmp_path / "trace" / "p_kernel_trace.csv", trace, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    _write_csv(tmp_path / "fetch" / "p_counter_collection.csv",
               [{"Kernel_Name": "k_split", "Counter_Name": "FETCH_SIZE", "Counter_Value": 500},
                {"Kernel_Name": "k_density", "Counter_Name": "FETCH_SIZE", "Counter_Value": 900},
                {"Kernel_Name": "k_density", "Counter_Name": "FETCH_SIZE", "Counter_Value": 1100}], fields)
    _write_csv(tmp_path / "write" / "p_counter_collection.csv",
               [{"Kernel_Name": "k_split", "Counter_Name": "WRITE_SIZE", "Counter_Value": 1000},
                {"Kernel_Name": "k_density", "Counter_Name": "WRITE_SIZE", "Counter_Value": 250}], fields)
    out = tmp_path / "out"
    t = ps.summarise(tmp_path / "trace", tmp_path / "fetch", tmp_path / "write", None, "synthetic", n, "cmd", out,
                     last=2)
    k = t["kernels"]["k_density"]
    assert k["read_bytes"] == pytest.approx(2 * copy)
    assert k["write_bytes"] == pytest.approx(0.25 * copy)
    assert k["hbm_bytes_per_launch"] == pytest.approx(2.25 * copy)
    assert json.loads((out / "traffic_synthetic.json").read_text())["kernels"]["k_density"]["read_bytes"] == \
        pytest.approx(2 * copy)
    table = (out / "kernels_synthetic.md").read_text()
    assert "| k_density | 2 | 0.25 |" in table  # two dispatches, 250 ns = 0.25 us


def test_algorithmic_bytes_are_the_per_pass_minimum():
    """Each array touched once per pass, each neighbour gather counted once: 176 N + 36 C bytes per
    step over the five passes; 24 N + 8 C for the density pass (positions, cell ranges, density
    and pressure out).  At the 1M block (C = 62^3) the density pass's figure is 25.9 MB, and a
    282 us launch is 0.0115 of the 8 TB/s peak."""
    sys.path.insert(0, str(ROOT))
    import bench

    n, c = 1_000_000, 62 ** 3
    a = bench.algorithmic_bytes(n, c)
    assert sum(a.values()) == 176 * n + 36 * c
    assert a["density"] == 24 * n + 8 * c
    assert a["update"] == 64 * n + 8 * c
    assert a["density"] / 282e-6 / 1e9 / bench.HBM_PEAK_GBS == pytest.approx(0.0115, abs=5e-4)
    o = bench.overhead_bytes(n, 29.5)
    assert o["density"] == 4 * n + 118 * n + 4 * n
    r = bench._rounded({"value": 1.23456789, "x": {"y": 1.23456789}, "z": [2.0 / 3.0]})
    assert r == {"value": 1.23456789, "x": {"y": 1.235}, "z": [0.6667]}
