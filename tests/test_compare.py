"""CPU tests of the .dat comparator and manifest (mhdglm/compare.py).

The norms are pinned by their closed forms (compare_precision.py:52-57 of
the reference: L1 = mean|d|, L2 = sqrt(mean d^2), Linf = max|d|,
relL2 = L2/RMS(ref)) on synthetic data where each has a known value, and
the .dat reader/md5 against oracle output.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from mhdglm import compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_norms_closed_form():
    ref = np.array([3.0, -4.0, 0.0, 0.0])
    test = ref - np.array([1.0, -1.0, 2.0, 0.0])
    n = compare.norms(ref, test)
    assert n["L1"] == pytest.approx(1.0)
    assert n["L2"] == pytest.approx(np.sqrt(6.0 / 4.0))
    assert n["Linf"] == pytest.approx(2.0)
    assert n["relL2"] == pytest.approx(np.sqrt(6.0 / 4.0) / np.sqrt(25.0 / 4.0))
    assert compare.norms(np.zeros(3), np.ones(3))["relL2"] == 0.0


def test_fields_from_state_matches_oracle(oracle):
    u = oracle.init("ot", 16, 12)
    W = oracle.cons_to_prim(u.reshape(-1, 9))
    f = compare.fields_from_state(np.moveaxis(u, -1, 0))
    for k, name in enumerate(["rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz"]):
        assert np.allclose(f[name].ravel(), W[:, k], rtol=1e-14, atol=1e-15)
    assert np.allclose(f["|B|"] ** 2, f["Bx"] ** 2 + f["By"] ** 2 + f["Bz"] ** 2)


def test_dat_round_trip_and_manifest(tmp_path):
    """Oracle .dat files: loader, self-comparison, FP64-vs-FP32 table, md5s."""
    outs = {}
    for prec in ("fp64", "fp32"):
        d = tmp_path / prec
        r = subprocess.run([pyoracle.ORACLE_BIN[prec], "24", "16", "0.05"],
                           env=dict(os.environ, OUTDIR=str(d)), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        (outs[prec],) = list(d.iterdir())
    a = compare.load_dat(str(outs["fp64"]))
    assert a.shape == (24 * 16, 13)
    same = compare.compare_fields(compare.fields_from_dat(a), compare.fields_from_dat(a))
    assert all(v["Linf"] == 0.0 for v in same.values())
    table = compare.compare_files(str(outs["fp64"]), str(outs["fp32"]))
    rho = [l for l in table.splitlines() if l.strip().startswith("rho")][0]
    assert 0 < float(rho.split()[-1]) < 1e-5
    assert "|B|" in table and "psi" in table
    man = compare.manifest(str(tmp_path)).splitlines()
    assert len(man) == 2
    for line in man:
        full, data, rel = line.split()
        p = tmp_path / rel
        assert full == compare.file_md5(str(p))
        rows = b"".join(l for l in p.read_bytes().splitlines(keepends=True) if not l.startswith(b"#"))
        import hashlib
        assert data == hashlib.md5(rows).hexdigest()


def test_cli_entry(tmp_path):
    r = subprocess.run(["python", "-m", "mhdglm.compare"], cwd=os.path.join(ROOT, "precision-study_amd"),
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr


def test_noise_floor_columns(tmp_path):
    """--floor: relL2(REF, FLOOR) and TEST's multiple of it, per variable."""
    rows = []
    rng = np.random.default_rng(1)
    base = rng.standard_normal((64, 13)) + 2.0
    for name, scale in (("ref", 0.0), ("test", 4e-6), ("floor", 2e-6)):
        a = base.copy()
        a[:, 2:11] += scale * rng.standard_normal((64, 9))
        p = tmp_path / f"{name}.dat"
        np.savetxt(p, a, header="layout test", fmt="%.17e")
        rows.append(str(p))
    table = compare.compare_files(rows[0], rows[1], rows[2])
    assert "# floor = floor.dat" in table
    rho = [l for l in table.splitlines() if l.strip().startswith("rho")][0].split()
    ref_rows = compare.compare_fields(compare.fields_from_dat(compare.load_dat(rows[0])),
                                      compare.fields_from_dat(compare.load_dat(rows[2])))
    assert float(rho[5]) == pytest.approx(ref_rows["rho"]["relL2"], rel=1e-2)
    assert float(rho[6]) == pytest.approx(float(rho[4]) / float(rho[5]), rel=2e-2)
    assert 1.0 < float(rho[6]) < 4.0
