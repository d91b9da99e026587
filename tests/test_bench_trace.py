"""bench.py's kernel-trace summary (--trace): the device-idle fraction is one
minus the union of the kernel intervals over the span, overlapping kernels
count once, and --trace-match / --trace-last select the dispatches."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def write_trace(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for name, a, e in rows:
            w.writerow({"Kernel_Name": name, "Start_Timestamp": a, "End_Timestamp": e})


def test_idle_fraction_and_overlap(bench, tmp_path):
    p = str(tmp_path / "t.csv")
    # [0, 100) and [50, 150) overlap: busy 150 of a 400 ns span, then
    # [300, 400): busy 250, idle 150 / 400
    write_trace(p, [("void fbk::k_a<double>(int)", 0, 100),
                    ("void fbk::k_b<double>(int)", 50, 150),
                    ("void fbk::k_a<double>(int)", 300, 400)])
    s = bench.trace_summary(p)
    assert s["kernels"] == 3
    assert s["span_ms"] == pytest.approx(400e-6)
    assert s["busy_ms"] == pytest.approx(250e-6)
    assert s["idle_frac"] == pytest.approx(150 / 400)
    by = {k["name"]: k for k in s["by_kernel"]}
    assert by["void fbk::k_a<double>"]["count"] == 2
    assert by["void fbk::k_a<double>"]["total_ms"] == pytest.approx(200e-6)


def test_match_and_last(bench, tmp_path):
    p = str(tmp_path / "t.csv")
    write_trace(p, [("void fbk::k_geometry<double>(x)", 0, 10),
                    ("void fbk::kb_scattering_density<double, false, false, 2>(y)", 20, 50),
                    ("void fbk::kb_multiple_scattering<double, false, 1, false>(z)", 60, 70)])
    s = bench.trace_summary(p, match="::kb_")
    assert s["kernels"] == 2
    assert s["span_ms"] == pytest.approx(50e-6)
    assert s["idle_frac"] == pytest.approx(10 / 50)
    assert all("::kb_" in k["name"] for k in s["by_kernel"])
    last = bench.trace_summary(p, last=1)
    assert last["kernels"] == 1 and last["by_kernel"][0]["name"].endswith("multiple_scattering"
                                                                          "<double, false, 1, false>")
    assert bench.trace_summary(p, match="nothing")["kernels"] == 0
