"""bench.py's JSON line, assembled from fixed timings on the CPU: the keys
the benchmark contract names, the roofline's byte accounting (28 B of state
per pass, the max-wave's 16 B reported separately), the committed traffic
summaries it reads, and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

CELLS = 4096 * 4096


def test_fused_roofline_counts_28_bytes():
    r = bench.roofline(CELLS, 70.0, 0.2345, 0.0076, False, "open_ocean_4096")
    assert r["kernel"] == "step_fused" and r["bytes_per_cell"] == 28
    assert r["achieved"] == pytest.approx(28 * CELLS / 0.2345e-3 / 1e9, abs=0.1)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, abs=1e-4)
    assert r["with_maxwave"]["bytes_per_cell"] == 44
    assert r["with_maxwave"]["achieved"] == pytest.approx(44 * CELLS / 0.2345e-3 / 1e9, abs=0.1)
    assert r["step_frac"] == pytest.approx(28 * 70.0 / 8000.0, abs=1e-4)
    assert r["step_frac_with_maxwave"] == pytest.approx(44 * 70.0 / 8000.0, abs=1e-4)
    assert r["other_kernel"] == {"step_fixup": {"ms_per_launch": 0.0076}}
    # committed rocprofv3 summary: HBM bytes per launch, max-wave included,
    # between the algorithmic bytes without and with the max-wave update
    assert r["traffic"] == 701131469
    assert r["algorithmic_bytes"] == 28 * CELLS
    assert r["algorithmic_bytes"] < r["traffic"] <= r["with_maxwave"]["algorithmic_bytes"]
    assert r["peak"] == 8000.0 and r["unit"] == "GB/s" and r["bound"] == "hbm"
    assert r["copy_gbs"] is None and r["frac_of_copy"] is None
    r = bench.roofline(CELLS, 70.0, 0.2345, 0.0076, False, "open_ocean_4096", copy_gbs=5000.0)
    assert r["frac_of_copy"] == pytest.approx(r["achieved"] / 5000.0, abs=1e-4)


def test_unfused_roofline_picks_the_slower_sweep():
    r = bench.roofline(CELLS, 60.0, 0.10, 0.17, True, "open_ocean_4096")
    assert r["kernel"] == "sweep_y" and r["ms_per_launch"] == 0.17
    assert r["with_maxwave"]["bytes_per_cell"] == 44          # the y-sweep carries the max-wave
    assert r["other_kernel"]["sweep_x"]["bytes_per_cell"] == 28
    assert r["step_bytes_per_cell"] == 56
    r = bench.roofline(CELLS, 60.0, 0.20, 0.17, True, "open_ocean_4096")
    assert r["kernel"] == "sweep_x" and r["with_maxwave"]["bytes_per_cell"] == 28


def test_traffic_summaries():
    assert bench.pmc_traffic("step_fused", "coastal_8192x6144") == 2071869235
    assert bench.pmc_traffic("sweep_x", "open_ocean_4096") == 484130816
    assert bench.pmc_traffic("step_fused", "open_ocean_2048") is None


def test_line_has_the_contract_keys():
    roof = bench.roofline(CELLS, 70.0, 0.2345, 0.0076, False, "open_ocean_4096")
    cpu = {"value": 0.15, "unit": "Gcell-updates/s", "cores": 16, "kind": "port", "sample": "x"}
    line = bench.bench_line(70.0, 10000, 20, 2.39, "2D Cartesian open ocean 4096x4096", 4096, 4096,
                            2.02, True, roof, None, cpu)
    json.dumps(line)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in line, k
    assert line["config"]["workload"] == "2D Cartesian open ocean 4096x4096"
    assert line["ms_per_step"] == pytest.approx(0.239)
    assert line["n_gpus"] == 1 and line["vs_baseline"] is None and line["higher_is_better"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in line["roofline"], k
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in line["cpu_baseline"], k


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 2 and "one GPU" in out.stderr
