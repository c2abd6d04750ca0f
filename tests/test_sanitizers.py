"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

  * the CPU oracle: every C entry point on small grids and special states
    (zeros, NaN, infinities, extreme magnitudes), every problem, both
    precisions, the region masks, a failing run; and the oracle's CLI on a
    whole small run;
  * the product's host side (mhd_api.cpp and the CLI, rebuilt with the
    sanitizers against the same device objects): initial conditions,
    argument validation, the autotuner's table, the CLI's usage errors and
    its reading of a selector model file -- everything that runs without a
    device.
A sanitizer stops the program at the first error it finds; the tests require
a clean exit and no report.  Leak checking is off for programs that load the
HIP runtime (its own allocations live until exit)."""
import os
import subprocess
import zipfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PKG = os.path.join(ROOT, "precision-study_amd")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def make(directory):
    r = subprocess.run(["make", "-C", directory, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


def run(args, env_extra, **kw):
    env = dict(os.environ, **env_extra)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env, **kw)
    for rep in REPORTS:
        assert rep not in r.stderr, r.stderr[-3000:]
    return r


@pytest.fixture(scope="module")
def oracle_san():
    make(ORACLE)
    return os.path.join(ORACLE, "build", "san")


@pytest.fixture(scope="module")
def host_san():
    make(PKG)
    return os.path.join(PKG, "build", "san")


def test_oracle_entry_points(oracle_san):
    r = run([os.path.join(oracle_san, "sanitize_checks")], {"ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_oracle_cli_whole_run(oracle_san, tmp_path):
    r = run([os.path.join(oracle_san, "mhd2d_glm_oracle_fp64"), "32", "32", "0.05", "0.25",
             "muscl", "hlld", "0.18"], {"ASAN_OPTIONS": "detect_leaks=1", "OUTDIR": str(tmp_path)})
    assert r.returncode == 0 and "Done. Steps:" in r.stdout
    assert len(list(tmp_path.iterdir())) == 1


def test_host_entry_points(host_san):
    r = run([os.path.join(host_san, "host_checks")],
            {"ASAN_OPTIONS": "detect_leaks=0",
             "MHD_AUTOTUNE_FILE": os.path.join(PKG, "autotune.txt")})
    assert r.returncode == 0 and "0 failures" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("args", [
    ["0", "8"], ["8", "8", "0.5", "1.5"], ["8", "8", "0.5", "0.25", "weno"],
    ["8", "8", "0.5", "0.25", "muscl", "roe"], ["8", "8", "0.5", "0.25", "muscl", "hlld", "0"],
    ["8", "8", "0.5", "0.25", "muscl", "hlld", "0.18", "otv"], ["8", "-3"]])
def test_host_cli_usage_errors(host_san, tmp_path, args):
    r = run([os.path.join(host_san, "mhd2d_glm_amd")] + args,
            {"ASAN_OPTIONS": "detect_leaks=0", "OUTDIR": str(tmp_path)})
    assert r.returncode == 1 and "Error" in r.stderr


def test_host_cli_selector_model(host_san, tmp_path):
    """MHD_SELECTOR: the model file is read before any device call -- refused
    outside the mixed form, for a missing, non-zip, compressed or incomplete
    file and for a member whose shape overflows its byte count; the shipped
    model is read (the run then needs a device)."""
    import numpy as np

    cli = os.path.join(host_san, "mhd2d_glm_amd")
    model = os.path.join(PKG, "selector_model.npz")
    args = [cli, "16", "16", "0.01", "0.25", "muscl", "hlld", "0.18"]
    garbage = tmp_path / "garbage.npz"
    garbage.write_bytes(bytes(range(256)) * 4)
    d = dict(np.load(model))
    np.savez_compressed(tmp_path / "compressed.npz", **d)
    np.savez(tmp_path / "partial.npz", params=d["params"], mean=d["mean"])
    # Stored members over 16 bytes of data whose headers claim 2^62 + 1
    # floats (the byte count wraps to 4 in 64 bits), or 100^10.
    for name, shape in (("huge", "(4611686018427387905,)"), ("many", "(" + "100, " * 10 + ")")):
        hdr = "{'descr': '<f4', 'fortran_order': False, 'shape': %s, }" % shape
        hdr += " " * ((64 - (10 + len(hdr) + 1) % 64) % 64) + "\n"
        blob = b"\x93NUMPY\x01\x00" + len(hdr).to_bytes(2, "little") + hdr.encode() + bytes(16)
        with zipfile.ZipFile(tmp_path / f"{name}.npz", "w", zipfile.ZIP_STORED) as z:
            z.writestr("mean.npy", blob)
    base = {"ASAN_OPTIONS": "detect_leaks=0", "OUTDIR": str(tmp_path)}
    cases = [({"MHD_SELECTOR": model}, "needs MHD_PRECISION=mixed"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(tmp_path / "none.npz")},
              "cannot open"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(garbage)}, "not a zip archive"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(tmp_path / "compressed.npz")},
              "compressed"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(tmp_path / "partial.npz")},
              "needs params, mean, std, threshold"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(tmp_path / "huge.npz")},
              "mean.npy: bad shape in .npy header"),
             ({"MHD_PRECISION": "mixed", "MHD_SELECTOR": str(tmp_path / "many.npz")},
              "mean.npy: truncated .npy data")]
    for env, msg in cases:
        r = run(args, dict(base, **env))
        assert r.returncode == 1 and msg in r.stderr, (env, r.stderr)
    r = run(args, dict(base, MHD_PRECISION="mixed", MHD_SELECTOR=model))
    assert "Learned tile selector: " + model in r.stdout, r.stdout + r.stderr
