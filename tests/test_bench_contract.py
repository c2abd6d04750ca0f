"""bench.py's contract, on the CPU.

  * one GPU only: --gpus above 1, or a WORLD_SIZE above 1, exits with code
    2 and a message, before anything touches a device;
  * with no GPU visible it exits with code 2 and says so;
  * its CPU-baseline leg times the reference CPU build (or the oracle when
    that build is absent) on a bounded sample and reports the fields the
    bench line carries: value, unit, cores, kind, sample, and with
    parallel > 1 the aggregate of that many single-thread runs at once.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=300,
                          env=dict(os.environ, **(env or {})))


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], None), (["--gpus", "8"], None),
                                      ([], {"WORLD_SIZE": "2"})])
def test_more_than_one_gpu_is_refused(args, env):
    r = run(args, env)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU may be visible here")
def test_no_gpu_visible_exits_with_a_message():
    r = run(["--steps", "1", "--warmup", "0"])
    assert r.returncode == 2, r.stderr
    assert "no GPU visible" in r.stderr
    assert r.stdout == ""


def test_cpu_baseline_leg():
    sys.path.insert(0, ROOT)
    import bench

    out = bench.cpu_baseline(32, 0.002, parallel=2)
    assert out is not None, "neither oracle/_ref nor oracle/build holds the FP64 CPU program"
    assert out["unit"] == "cell-updates/s"
    assert out["cores"] == 1 and out["threading"] == "1 thread per run"
    assert out["kind"] in ("reference", "port")
    assert out["value"] > 0
    assert out["sample"].startswith("Orszag-Tang 32x32 FP64 MUSCL+HLLD to t=0.002")
    par = out["parallel"]
    assert par["runs"] == 2 and par["value"] > 0
    assert par["per_run_min_max"][0] <= par["per_run_min_max"][1]
