"""bench.py's output line on the GPU, on a small run: one JSON object with the metric,
the workload, the roofline of the dominant kernel and the CPU baseline, in the fields
and vocabulary the benchmark contract names."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_bench_line_fields():
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--n", "1000", "--steps", "40", "--warmup", "5",
           "--prof-steps", "20", "--cpu-seconds", "1", "--no-surrogate", "--no-ensemble", "--no-collision",
           "--no-e2e"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 40 and d["warmup"] == 5
    assert d["value"] > 0 and d["higher_is_better"] is True
    assert "workload" in d["config"]
    roof = d["roofline"]
    assert roof["bound"] in ("hbm", "mfma")
    assert roof["unit"] in ("GB/s", "TFLOP/s")
    assert 0 < roof["achieved"] < roof["peak"]
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-12
    cpu = d["cpu_baseline"]
    assert cpu["kind"] in ("reference", "port") and cpu["cores"] >= 1 and cpu["value"] > 0
    assert cpu["unit"] and cpu["sample"]
