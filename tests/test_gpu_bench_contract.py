"""bench.py's JSON line on a GPU, at a small size with the long sections
off: one line on stdout carrying every key the line's contract names, with
the types and relations it states (value = cells x steps / wall, the HBM
fraction = achieved / peak, the CPU baseline timed on this host)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_line_contract():
    n, steps, warmup = 256, 4, 2
    args = [sys.executable, os.path.join(ROOT, "bench.py"), "--n", str(n), "--steps", str(steps),
            "--warmup", str(warmup), "--forms=mixed", "--gate-n", "0", "--ensemble", "0",
            "--large-n", "0", "--slab-rows", "0", "--selector-gain", "0", "--cpu-n", "64",
            "--cpu-t-end", "0.01", "--cpu-parallel", "2"]
    print("bench.py at 256^2, sections off", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout
    d = json.loads(lines[0])
    for key, typ in (("metric", str), ("value", float), ("unit", str), ("n_gpus", int),
                     ("steps", int), ("warmup", int), ("ms_per_step", float),
                     ("higher_is_better", bool), ("scaling", str), ("dtype", str),
                     ("data", str), ("config", dict), ("roofline", dict),
                     ("cpu_baseline", dict)):
        assert isinstance(d[key], typ), (key, d.get(key))
    assert d["vs_baseline"] is None
    assert d["n_gpus"] == 1 and d["steps"] == steps and d["warmup"] == warmup
    assert d["unit"] == "cell-updates/s" and d["higher_is_better"] is True
    assert d["scaling"] == "weak" and d["dtype"] == "f64"
    assert "workload" in d["config"] and "model" not in d["config"]
    assert d["value"] > 0
    assert abs(d["value"] - n * n * steps / (d["ms_per_step"] * steps * 1e-3)) <= 1e-6 * d["value"]
    roof = d["roofline"]
    assert roof["bound"] in ("hbm", "mfma") and roof["unit"] == "GB/s"
    assert roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) <= 1e-3
    assert roof["traffic"] is None or roof["traffic"] > 0
    cpu = d["cpu_baseline"]
    assert cpu["value"] > 0 and cpu["unit"] == "cell-updates/s" and cpu["cores"] == 1
    assert cpu["kind"] in ("reference", "port") and cpu["sample"]
    assert cpu["parallel"]["runs"] == 2
    assert d["forms"]["mixed"] > 0 and d["forms"]["fp64"] == pytest.approx(d["value"], rel=1e-6)
