"""The JSON line of bench.py on the GPU: the fields the headline measurement
is read from, value consistent with ms_per_step, the roofline fraction, and the
CPU leg's bit-exact re-check.  Collected after the other GPU tests, so the
bench subprocess starts on a warm box."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_bench_json_line():
    print("bench.py --steps 2 (headline workload, short)", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--steps", "2",
                        "--warmup", "1", "--no-serve", "--no-rmt", "--cpu-seconds", "2",
                        "--no-cpu-threads"],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
                "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["unit"] == "pairs/s" and d["higher_is_better"] is True
    assert d["dtype"] == "fp32" and d["data"] == "synthetic"
    assert "workload" in d["config"]
    assert d["value"] > 0
    # value is the pairs of the timed blocks over the timed seconds
    assert abs(d["value"] * d["ms_per_step"] * 1e-3 - 32768) < 0.01 * 32768
    roof = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in roof, key
    assert 0 < roof["frac"] < 1
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-9
    cpu = d["cpu_baseline"]
    assert cpu["kind"] == "port" and cpu["cores"] == 1 and cpu["value"] > 0
    # the CPU leg re-checks bit-exact parity on its sample
    assert "bit-exact on" in cpu["sample"]
    checked = cpu["sample"].split("bit-exact on ")[1].split("/")
    assert checked[0] == checked[1].split()[0]
