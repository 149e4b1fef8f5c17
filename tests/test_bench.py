"""bench.py's command-line contract: one JSON line with the metric, the
roofline and the CPU baseline; more than one GPU is refused."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args, timeout=600):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    res = _bench("--gpus", "2", timeout=120)
    assert res.returncode == 2
    msg = json.loads(res.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


@pytest.mark.gpu
def test_json_line_contract():
    res = _bench("--n", "20000", "--steps", "40", "--warmup", "5", "--timing-steps", "10", "--cpu-seconds", "1")
    assert res.returncode == 0, res.stderr[-2000:]
    out = json.loads(res.stdout.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 40 and out["warmup"] == 5
    assert out["unit"] == "particle-steps/s" and out["value"] > 0 and out["higher_is_better"] is True
    reps = out["repeat_values"]
    assert out["repeats"] == len(reps) == 3 and sorted(reps)[1] == out["value"]
    assert out["contacts_per_s"] >= 0 and out["contacts_per_step"] >= 0
    assert out["dtype"] == "f64" and out["data"] == "synthetic" and "workload" in out["config"]
    rf = out["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-12
    cb = out["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] >= 1 and cb["value"] > 0 and cb["unit"] == "particle-steps/s"
    assert out["value"] > cb["value"]
