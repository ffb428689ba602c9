"""bench.py's output line on a short run: one JSON line with the fields the
measurement contract names, the roofline object consistent with itself, and
the timed value consistent with the step time."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
@pytest.mark.timeout(300)
def test_bench_json_line():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "5",
                        "--warmup", "2", "--timing-steps", "4", "--no-cpu", "--no-batch",
                        "--no-strong", "--no-tune"],
                       capture_output=True, text=True, timeout=280)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2
    assert d["unit"] == "cell-steps/s" and d["dtype"] == "f64" and d["higher_is_better"]
    assert d["vs_baseline"] is None
    cells = d["config"]["cells"]
    assert abs(d["value"] - cells / (d["ms_per_step"] * 1e-3)) <= 1e-6 * d["value"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0.0 < rf["frac"] < 1.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    k = ("sweep_x", "sweep_y", "sweep_z").index(rf["kernel"])
    assert abs(rf["achieved"] - rf["bytes_per_launch"] / (rf["kernel_ms"][k] * 1e-3) / 1e9) \
        <= 1e-3 * rf["achieved"]
