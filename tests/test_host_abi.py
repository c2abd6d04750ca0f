"""CPU tests of the product's host side: C ABI surface, host-evaluated
initial conditions, CLI argument handling.  No device call is made."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhd_glm.h")
UINT = {np.float64: np.uint64, np.float32: np.uint32}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(mhd):
    lib = mhd.load()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(mhd.EXPORTED)


def test_header_is_plain_c():
    text = open(HEADER).read()
    for banned in ("torch", "hipStream_t", "std::", "template"):
        assert banned not in re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", HEADER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("problem", ["ot", "kh", "khb"])
@pytest.mark.parametrize("dtype,prec", [(np.float64, 0), (np.float32, 1)])
def test_initial_state_matches_oracle(mhd, oracle, problem, dtype, prec):
    nx, ny = 48, 40
    u = mhd.problem_state(problem, nx, ny, prec)
    ref = np.moveaxis(oracle.init(problem, nx, ny, dtype=dtype), -1, 0)
    assert u.dtype == dtype
    assert np.array_equal(u.view(UINT[dtype]), np.ascontiguousarray(ref).view(UINT[dtype]))


def test_problem_state_rejects_unknown(mhd):
    with pytest.raises(mhd.MhdError):
        mhd.problem_state("sod", 8, 8)


def run_cli(mhd, args, tmp_path, precision="fp64"):
    return mhd.run_cli(args, str(tmp_path), precision=precision)


@pytest.mark.parametrize("args,msg", [
    (["0", "8"], "invalid grid or final time"),
    (["8", "8", "-1"], "invalid grid or final time"),
    (["8", "8", "0.5", "1.5"], "CFL must be in (0, 1)"),
    (["8", "8", "0.5", "0.25", "weno"], "order must be 'muscl' or 'first'"),
    (["8", "8", "0.5", "0.25", "muscl", "roe"], "solver must be 'hll' or 'hlld'"),
    (["8", "8", "0.5", "0.25", "muscl", "hlld", "0"], "alpha must be positive"),
    (["8", "8", "0.5", "0.25", "muscl", "hlld", "0.18", "sod"], "problem must be"),
    (["8", "8", "0.5", "0.25", "muscl", "hlld", "0.18", "otv"], "problem must be"),
    (["8", "-3"], "invalid grid or final time"),
])
def test_cli_usage_errors(mhd, tmp_path, args, msg):
    r = run_cli(mhd, args, tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr


def test_cli_rejects_bad_precision(mhd, tmp_path):
    r = run_cli(mhd, ["8", "8"], tmp_path, precision="fp16")
    assert r.returncode == 1


def test_otv_problem_reduces_to_ot(mhd):
    for prec in (0, 1):
        a = mhd.problem_state("ot", 40, 36, prec)
        b = mhd.problem_state_ex("otv", (1.0, 0.0, 0.0, 0.0, 0.0), 40, 36, prec)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_otv_problem_closed_form(mhd):
    """Amplitude and phases enter as documented (independent numpy form)."""
    amp, p1, p2, p3, p4 = 0.9, 0.3, 1.1, 2.0, 5.5
    n = 32
    u = mhd.problem_state_ex("otv", (amp, p1, p2, p3, p4), n, n)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    rho = 25.0 / (36.0 * np.pi)
    c = 1.0 / np.sqrt(4.0 * np.pi)
    assert np.allclose(u[0], rho, rtol=1e-15)
    assert np.allclose(u[1] / rho, -np.sin(2 * np.pi * Y + p1), atol=1e-14)
    assert np.allclose(u[2] / rho, np.sin(2 * np.pi * X + p2), atol=1e-14)
    assert np.allclose(u[5], -amp * np.sin(2 * np.pi * Y + p3) * c, atol=1e-14)
    assert np.allclose(u[6], amp * np.sin(4 * np.pi * X + p4) * c, atol=1e-14)
    with pytest.raises(mhd.MhdError):
        mhd.problem_state_ex("otv", (1.0, 0.0), n, n)


@pytest.mark.parametrize("dtype,prec", [(np.float64, 0), (np.float32, 1)])
def test_otv_initial_state_matches_oracle(mhd, oracle, dtype, prec):
    """The oracle's own otv (closed form pinned by the numpy test above and
    by its reduction to ot) equals the product host's bit for bit."""
    params = (1.067437115364036, 0.8064116573994691, 1.759240013263218, 0.20557563316160354,
              0.5272182398647961)
    nx, ny = 40, 36
    u = mhd.problem_state_ex("otv", params, nx, ny, prec)
    ref = np.moveaxis(oracle.init("otv", nx, ny, dtype=dtype, params=params), -1, 0)
    assert np.array_equal(u.view(UINT[dtype]), np.ascontiguousarray(ref).view(UINT[dtype]))
    ot = oracle.init("ot", nx, ny, dtype=dtype)
    one = oracle.init("otv", nx, ny, dtype=dtype, params=(1.0, 0.0, 0.0, 0.0, 0.0))
    assert np.array_equal(ot.view(UINT[dtype]), one.view(UINT[dtype]))


def test_autotune_table_lookup(tmp_path):
    """The library picks the first matching row of the autotuner's table."""
    import subprocess
    import sys
    from mhdglm import autotune

    table = tmp_path / "autotune.txt"
    autotune.write_table([(0, 256 * 256, 1), (0, 1024 * 1024, 0), (0, 4096 * 4096, 4),
                          (2, 512 * 512, 5), (2, 2048 * 2048, 3, 2)], str(table), "test")
    # A table of the earlier three-column format reads with one stage per launch.
    with open(table, "a") as f:
        f.write("1 4096 2\n")
    cases = ((0, 128), (0, 256), (0, 512), (0, 1024), (0, 8192), (2, 512), (2, 4096),
             (1, 64), (1, 128))
    code = ("import sys; sys.path.insert(0, sys.argv[1]); import mhdglm; l = mhdglm.load(); "
            f"cases = {cases!r}; "
            "print([l.mhd_autotuned_tile(p, n, n) for p, n in cases], "
            "[l.mhd_autotuned_fuse(p, n, n) for p, n in cases])")
    env = dict(os.environ, MHD_AUTOTUNE_FILE=str(table))
    r = subprocess.run([sys.executable, "-c", code, os.path.join(ROOT, "precision-study_amd")],
                       env=env, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "[1, 1, 0, 0, 4, 5, 3, 2, 3] [1, 1, 1, 1, 1, 1, 2, 1, 1]"
