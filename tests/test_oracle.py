"""CPU tests: pin the oracle, check the host side and the C ABI surface.

The oracle (oracle/mhd_oracle.hpp) is pinned against
  * the reference's whole-run goldens (md5 of .dat data rows, step and
    fallback counts, divB/psi diagnostics) -- tests/golden/goldens.json;
  * the reference's own physics functions and CPU driver compiled from their
    sources into oracle/_ref (bitwise, random and special states, every
    order/solver combination);
  * invariants a plausible slip would break: HLL/HLLD(U,U) = physical flux,
    y flux = rotated x flux, GLM normal state of equal sides, the closed form
    of dt, prim->cons->prim round trip.
"""
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle
import states

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAVE_REF = os.path.exists(os.path.join(pyoracle.REF_DIR, "libmhdref_fp64.so"))
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")

DTYPES = {"fp64": np.float64, "fp32": np.float32}
UINT = {np.float64: np.uint64, np.float32: np.uint32}


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(UINT[a.dtype.type])


def assert_bitwise(a, b):
    assert a.shape == b.shape
    ba, bb = bits(a), bits(b)
    bad = np.argwhere(ba != bb)
    assert bad.size == 0, f"{len(bad)} mismatches, first at {bad[0]}: {a[tuple(bad[0])]!r} vs {b[tuple(bad[0])]!r}"


def swap_xy(U):
    V = np.array(U, copy=True)
    V[:, [1, 2]] = U[:, [2, 1]]
    V[:, [5, 6]] = U[:, [6, 5]]
    return V


# --------------------------------------------------------------------------
# oracle vs the reference's own header functions
# --------------------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("hlld", [True, False])
def test_flux_matches_reference_headers(oracle, prec, direction, hlld):
    ref = pyoracle.RefHeaders(prec)
    dt = DTYPES[prec]
    UL, UR = states.face_pairs(seed=7 + direction, n=4096)
    for ch in (0.7, 2.5, 1e-13):
        Fo, so = oracle.glm_flux(UL.astype(dt), UR.astype(dt), ch, direction, hlld, dtype=dt)
        Fr, sr = ref.glm_flux(UL.astype(dt), UR.astype(dt), ch, direction, hlld)
        assert_bitwise(Fo, Fr)
        assert np.array_equal(so, sr)
    if hlld:
        # The sample reaches every status.
        assert set(np.unique(so)) >= {0, 1, 2} or set(np.unique(sr)) >= {0, 1}


@needs_ref
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pointwise_matches_reference_headers(oracle, prec):
    ref = pyoracle.RefHeaders(prec)
    dt = DTYPES[prec]
    U = states.cell_states(n=4096).astype(dt)
    assert_bitwise(oracle.cons_to_prim(U, dtype=dt), ref.cons_to_prim(U))
    W = ref.cons_to_prim(U)
    assert_bitwise(oracle.prim_to_cons(W, dtype=dt), ref.prim_to_cons(W))
    for d in (0, 1):
        assert_bitwise(oracle.fast_speed(W, d, dtype=dt), ref.fast_speed(W, d))


@needs_ref
@pytest.mark.parametrize("order,solver", [("muscl", "hlld"), ("muscl", "hll"), ("first", "hlld"), ("first", "hll")])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_small_runs_match_reference_binary(tmp_path, prec, order, solver):
    """Byte-identical .dat files (headers included) on a non-square grid."""
    args = ["40", "24", "0.15", "0.3", order, solver, "0.18"]
    outs = []
    for name, exe in (("orc", pyoracle.ORACLE_BIN[prec]), ("ref", pyoracle.REF_BIN[prec])):
        d = tmp_path / name
        r = subprocess.run([exe] + args, env=dict(os.environ, OUTDIR=str(d)), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        (f,) = list(d.iterdir())
        outs.append((f.read_bytes(), re.sub(r"Time loop wall time = .*\n", "", r.stdout.split("=====\n", 1)[1])))
    assert outs[0][0] == outs[1][0]
    assert outs[0][1].replace(str(tmp_path / "orc"), "X") == outs[1][1].replace(str(tmp_path / "ref"), "X")


@needs_ref
@pytest.mark.parametrize("nx,ny", [(1, 8), (7, 1), (1, 1), (2, 3), (3, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_degenerate_grids_match_reference_binary(tmp_path, prec, nx, ny):
    """Grids narrower than the 2-cell halo.  With a dimension of 1 the
    reference's periodic fill copies the far ghost cell from the near one
    before that is refreshed, so it holds an older state; the oracle's
    (and the GPU's) neighbour is the cell itself.  Every other cell of
    that stencil is the cell itself, so the slope minmod(stale - U, 0) is
    0 and the stale value cannot reach the result: the .dat files are
    byte-identical."""
    args = [str(nx), str(ny), "0.05", "0.3", "muscl", "hlld", "0.18"]
    outs = []
    for name, exe in (("orc", pyoracle.ORACLE_BIN[prec]), ("ref", pyoracle.REF_BIN[prec])):
        d = tmp_path / name
        r = subprocess.run([exe] + args, env=dict(os.environ, OUTDIR=str(d)), capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        (f,) = list(d.iterdir())
        outs.append(f.read_bytes())
    assert outs[0] == outs[1]


# --------------------------------------------------------------------------
# invariants
# --------------------------------------------------------------------------
def physical_flux_x(U):
    """Ideal-MHD x flux from conserved states (float64, numpy)."""
    g = 5.0 / 3.0
    rho = U[:, 0]
    vx, vy, vz = U[:, 1] / rho, U[:, 2] / rho, U[:, 3] / rho
    bx, by, bz = U[:, 5], U[:, 6], U[:, 7]
    b2 = bx * bx + by * by + bz * bz
    p = (g - 1.0) * (U[:, 4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * b2)
    pt = p + 0.5 * b2
    vb = vx * bx + vy * by + vz * bz
    F = np.zeros((U.shape[0], 8))
    F[:, 0] = rho * vx
    F[:, 1] = rho * vx * vx + pt - bx * bx
    F[:, 2] = rho * vx * vy - bx * by
    F[:, 3] = rho * vx * vz - bx * bz
    F[:, 4] = (U[:, 4] + pt) * vx - bx * vb
    F[:, 6] = vx * by - bx * vy
    F[:, 7] = vx * bz - bx * vz
    return F


@pytest.mark.parametrize("hlld", [True, False])
def test_consistency_equal_states(oracle, hlld):
    rng = np.random.default_rng(5)
    U = states.prim_to_cons(states.random_prims(rng, 2000))
    ch = 1.3
    F, st = oracle.glm_flux(U, U, ch, 0, hlld)
    Fp = physical_flux_x(U)
    scale = np.abs(Fp).max(axis=1, keepdims=True) + np.abs(U).max(axis=1, keepdims=True)
    mask = [0, 1, 2, 3, 4, 6, 7]
    err = np.abs(F[:, mask] - Fp[:, mask]) / scale
    assert err.max() < 1e-13
    # GLM subsystem with equal sides: Bn* = Bn, psi* = psi.
    assert np.allclose(F[:, 5], U[:, 8], rtol=0, atol=1e-15)
    assert np.allclose(F[:, 8], ch * ch * U[:, 5], rtol=1e-15, atol=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_y_flux_is_rotated_x_flux(oracle, dtype):
    UL, UR = states.face_pairs(seed=11, n=2048)
    UL, UR = UL.astype(dtype), UR.astype(dtype)
    for hlld in (True, False):
        Fy, sy = oracle.glm_flux(UL, UR, 0.9, 1, hlld, dtype=dtype)
        Fx, sx = oracle.glm_flux(swap_xy(UL), swap_xy(UR), 0.9, 0, hlld, dtype=dtype)
        assert_bitwise(Fy, swap_xy(Fx))
        assert np.array_equal(sy, sx)


def test_glm_normal_state_linear_system(oracle):
    """F[Bn] and F[psi] are the exact upwind flux of the linear (Bn, psi) system."""
    rng = np.random.default_rng(3)
    W = states.random_prims(rng, 1000)
    UL = states.prim_to_cons(W)
    UR = states.prim_to_cons(states.random_prims(rng, 1000))
    ch = 1.7
    F, _ = oracle.glm_flux(UL, UR, ch, 0, True)
    bL, bR, pL, pR = UL[:, 5], UR[:, 5], UL[:, 8], UR[:, 8]
    # Characteristic variables w+- = psi +- ch*Bn travel with speeds +-ch.
    bn = 0.5 * (bL + bR) - 0.5 * (pR - pL) / ch
    ps = 0.5 * (pL + pR) - 0.5 * ch * (bR - bL)
    assert np.allclose(F[:, 5], ps, rtol=1e-14, atol=1e-14)
    assert np.allclose(F[:, 8], ch * ch * bn, rtol=1e-14, atol=1e-14)
    wplus = ps + ch * bn
    assert np.allclose(wplus, pL + ch * bL, rtol=1e-13, atol=1e-13)


def test_prim_cons_round_trip(oracle):
    rng = np.random.default_rng(8)
    W = states.random_prims(rng, 5000)
    W2 = oracle.cons_to_prim(oracle.prim_to_cons(W))
    rel = np.abs(W2 - W) / np.maximum(np.abs(W), 1e-300)
    assert rel[:, [0, 5, 6, 7, 8]].max() == 0.0
    assert rel[:, 1:4].max() < 4e-16
    assert np.abs(W2[:, 4] - W[:, 4]).max() < 1e-13


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dt_closed_form(oracle, dtype):
    """The first dt equals cfl / (c_h/dx + c_h/dy) with c_h the max signal speed."""
    nx, ny = 24, 16
    u = oracle.init("ot", nx, ny, dtype=dtype)
    ch = dtype(max(oracle.max_signal_speed(u), 1e-12))
    dx, dy = dtype(1.0) / dtype(nx), dtype(1.0) / dtype(ny)
    dt = dtype(0.25) / (ch / dx + ch / dy)
    _, info = oracle.advance(u, 1, t_end=0.5, cfl=0.25)
    assert info["code"] == 0 and info["steps"] == 1
    assert bits(np.array([info["t"]], dtype=dtype))[0] == bits(np.array([dt], dtype=dtype))[0]


# --------------------------------------------------------------------------
# whole-run goldens
# --------------------------------------------------------------------------
def data_md5(path):
    h = hashlib.md5()
    with open(path, "rb") as f:
        for line in f:
            if not line.startswith(b"#"):
                h.update(line)
    return h.hexdigest()


def parse_summary(stdout):
    out = {}
    m = re.search(r"Done\. Steps: (\d+)", stdout)
    out["steps"] = int(m.group(1))
    m = re.search(r"fallbacks = (\d+) \(degenerate=(\d+), nonphysical=(\d+)\)", stdout)
    if m:
        out["fallbacks"], out["degenerate"], out["nonphysical"] = map(int, m.groups())
    m = re.search(r"divB L1 = (\S+), divB Linf = (\S+), psi Linf = (\S+)", stdout)
    out["divB_L1"], out["divB_Linf"], out["psi_Linf"] = m.groups()
    return out


def test_oracle_whole_run_goldens(tmp_path, goldens):
    """OT 128^2 FP64 and FP32 to t=0.5 reproduce the reference goldens."""
    procs = {}
    for prec in ("fp64", "fp32"):
        d = tmp_path / prec
        procs[prec] = (d, subprocess.Popen(
            [pyoracle.ORACLE_BIN[prec], "128", "128", "0.5", "0.25", "muscl", "hlld", "0.18"],
            env=dict(os.environ, OUTDIR=str(d)), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    for prec, (d, p) in procs.items():
        out, err = p.communicate(timeout=600)
        assert p.returncode == 0, err
        g = goldens["ot"][f"128_{prec}"]
        s = parse_summary(out)
        for k in ("steps", "fallbacks", "degenerate", "nonphysical", "divB_L1", "divB_Linf", "psi_Linf"):
            assert s[k] == g[k], (prec, k, s[k], g[k])
        (f,) = list(d.iterdir())
        assert data_md5(f) == g["md5"]
