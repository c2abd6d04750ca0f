"""GPU parity tests: the HIP path through the C ABI against the oracle.

Bar: bitwise equality (FP contraction off) for FP64 and FP32, on
  * the operator-level functions (flux, recovery, wave speed) over random
    and special states;
  * whole SSPRK2 steps on small grids (every order/solver, odd sizes, every
    tile shape, random perturbed states, all three problems);
  * whole runs to t=0.5 against the reference goldens (md5 of .dat data
    rows, steps, fallback counts);
and, at the full 4096^2 size, size-independent properties (conservation,
determinism, graph/stream equivalence).
"""
import hashlib
import os
import re
import sys

import numpy as np
import pytest

import pyoracle
import states

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UINT = {np.float64: np.uint64, np.float32: np.uint32}
PREC = {"fp64": (0, np.float64), "fp32": (1, np.float32)}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_bitwise(a, b, what=""):
    """Bitwise equality, except that NaNs compare equal whatever their sign
    and payload: IEEE 754 leaves those unspecified, x86 SSE produces a
    negative default NaN and CDNA a positive one.  NaN positions must match."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), f"{what}: NaN positions differ"
    ba, bb = a.view(UINT[a.dtype.type]), b.view(UINT[b.dtype.type])
    bad = np.argwhere((ba != bb) & ~na)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}: {a[tuple(bad[0])]!r} vs {b[tuple(bad[0])]!r}"


# --------------------------------------------------------------------------
# operator level
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("hlld", [True, False])
def test_glm_flux_bitwise(mhd, oracle, prec, direction, hlld):
    p, dt = PREC[prec]
    say(f"flux {prec} dir={direction} hlld={hlld}")
    UL, UR = states.face_pairs(seed=21 + direction, n=8192)
    UL, UR = UL.astype(dt), UR.astype(dt)
    for ch in (0.7, 2.5, 1e-13):
        Fg, sg = mhd.glm_flux(UL, UR, ch, direction, hlld, precision=p)
        Fo, so = oracle.glm_flux(UL, UR, ch, direction, hlld, dtype=dt)
        assert_bitwise(Fg, Fo, f"flux ch={ch}")
        assert np.array_equal(sg, so)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_recovery_and_wave_speed_bitwise(mhd, oracle, prec):
    p, dt = PREC[prec]
    say(f"c2p/speed {prec}")
    U = states.cell_states(n=8192).astype(dt)
    assert_bitwise(mhd.cons_to_prim(U, precision=p), oracle.cons_to_prim(U, dtype=dt), "c2p")
    W = oracle.cons_to_prim(U, dtype=dt)
    for d in (0, 1):
        assert_bitwise(mhd.fast_speed(W, d, precision=p), oracle.fast_speed(W, d, dtype=dt), f"cf{d}")


def test_fp64_sqrt_div_correctly_rounded(mhd, oracle):
    """Wave speeds on widely scaled inputs exercise sqrt and division."""
    rng = np.random.default_rng(4)
    W = states.random_prims(rng, 20000)
    W[:, 0] *= 10.0 ** rng.uniform(-10, 10, W.shape[0])
    W[:, 4] *= 10.0 ** rng.uniform(-10, 10, W.shape[0])
    for d in (0, 1):
        assert_bitwise(mhd.fast_speed(W, d), oracle.fast_speed(W, d), f"cf{d}")


# --------------------------------------------------------------------------
# whole steps on small grids
# --------------------------------------------------------------------------
def gpu_run(mhd, u_aos, steps, prec="fp64", muscl=True, hlld=True, tile=-1, t_end=0.5,
            cfl=0.25, batch=0):
    p, dt = PREC[prec]
    ny, nx, _ = u_aos.shape
    s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, cfl=cfl, muscl=muscl, hlld=hlld, precision=p,
                                    tile=tile, batch_steps=batch))
    s.set_state(np.moveaxis(u_aos, -1, 0))
    st = s.advance(t_end, steps)
    out = np.moveaxis(s.get_state(0), 0, -1)
    u1 = np.moveaxis(s.get_state(1), 0, -1)
    s.close()
    return out, u1, st


def check_steps(mhd, oracle, u, steps, prec="fp64", **kw):
    ug, _, st = gpu_run(mhd, u, steps, prec, **kw)
    okw = {k: v for k, v in kw.items() if k in ("muscl", "hlld", "t_end", "cfl")}
    uo, info = oracle.advance(u, steps, **okw)
    assert st.code == info["code"] == 0
    assert st.steps == info["steps"]
    assert_bitwise(np.array([st.t], dtype=u.dtype), np.array([info["t"]], dtype=u.dtype), "t")
    assert_bitwise(ug, uo, "state")
    if kw.get("hlld", True):
        cnt = info["counters"]
        assert (st.x_evaluations, st.y_evaluations, st.degenerate, st.nonphysical) == tuple(int(c) for c in cnt)
    return st


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("muscl,hlld", [(True, True), (True, False), (False, True), (False, False)])
def test_steps_ot_all_schemes(mhd, oracle, prec, muscl, hlld):
    _, dt = PREC[prec]
    say(f"steps OT 40x24 {prec} muscl={muscl} hlld={hlld}")
    u = oracle.init("ot", 40, 24, dtype=dt)
    check_steps(mhd, oracle, u, 25, prec, muscl=muscl, hlld=hlld)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("nx,ny", [(64, 64), (50, 37), (33, 70), (2, 5), (1, 9), (7, 1), (1, 1),
                                   (3, 2)])
def test_steps_tiles_and_odd_sizes(mhd, oracle, tile, nx, ny):
    say(f"steps OT {nx}x{ny} tile={tile}")
    u = oracle.init("ot", nx, ny)
    check_steps(mhd, oracle, u, 12, "fp64", tile=tile)


@pytest.mark.parametrize("problem", ["kh", "khb"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_steps_kelvin_helmholtz(mhd, oracle, problem, prec):
    _, dt = PREC[prec]
    say(f"steps {problem} 64x64 {prec}")
    u = oracle.init(problem, 64, 64, dtype=dt)
    check_steps(mhd, oracle, u, 30, prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_steps_random_perturbed_state(mhd, oracle, prec):
    """Noisy states reach the HLLD fallbacks and first-order reconstruction."""
    _, dt = PREC[prec]
    rng = np.random.default_rng(17)
    u = oracle.init("ot", 48, 48).astype(np.float64)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W *= 1.0 + 0.3 * rng.standard_normal(W.shape)
    W[..., 0] = np.abs(W[..., 0]) + 0.01
    W[..., 4] = np.abs(W[..., 4]) + 0.01
    W[..., 8] = 0.05 * rng.standard_normal(W.shape[:2])
    W[5:9, 10:14, 5] = 0.0     # a patch with B_n = 0
    W[5:9, 10:14, 6] = 0.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape).astype(dt)
    say(f"steps random 48x48 {prec}")
    st = check_steps(mhd, oracle, u, 15, prec, cfl=0.2)
    assert st.degenerate > 0


def test_graph_and_batch_sizes_agree(mhd, oracle):
    u = oracle.init("ot", 64, 48)
    ref, _, st0 = gpu_run(mhd, u, 40, batch=1)
    for batch in (7, 32):
        ug, _, st = gpu_run(mhd, u, 40, batch=batch)
        assert st.steps == st0.steps == 40
        assert_bitwise(ug, ref, f"batch={batch}")


def test_t_end_reached_exactly(mhd, oracle):
    """The last dt is clamped to t_end; steps stop there (mhd2d_glm_cpu.cpp:523-524)."""
    u = oracle.init("ot", 32, 32)
    ug, _, st = gpu_run(mhd, u, 10000, t_end=0.03)
    uo, info = oracle.advance(u, 10000, t_end=0.03)
    assert st.done == 1 and st.steps == info["steps"]
    assert st.t == 0.03 and info["t"] == 0.03
    assert_bitwise(ug, uo)


# --------------------------------------------------------------------------
# failures: exit code, step and first bad cell (mhd2d_glm_cpu.cpp:538-566)
# --------------------------------------------------------------------------
def test_failure_stage1_reports_first_bad_cell(mhd, oracle):
    u = oracle.init("ot", 32, 32)
    for (j, i) in ((20, 7), (9, 30)):
        u[j, i, 4] = 1e-8   # E far below kinetic + magnetic energy
    ug, u1g, st = gpu_run(mhd, u, 100)
    _, info = oracle.advance(u, 100)
    assert info["code"] == 3 and st.code == 3
    assert (st.bad_i, st.bad_j) == info["bad"]
    assert st.fail_step == info["steps"] == 0
    # The device state is left as the reference leaves it (U untouched).
    assert_bitwise(ug, u)


def test_failure_matches_oracle_when_later(mhd, oracle):
    """A marginal state that survives some steps, then fails."""
    u = oracle.init("ot", 32, 32)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W[..., 4] *= 0.02
    W[10:20, 10:20, 1] *= 8.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape)
    ug, u1g, st = gpu_run(mhd, u, 400, cfl=0.9)
    uo, info = oracle.advance(u, 400, cfl=0.9)
    assert st.code == info["code"]
    assert st.steps == info["steps"]
    if info["code"] in (3, 4):
        assert (st.bad_i, st.bad_j) == info["bad"]
    assert_bitwise(ug, uo)


# --------------------------------------------------------------------------
# whole runs against the reference goldens (through the drop-in CLI)
# --------------------------------------------------------------------------
def data_md5(path):
    h = hashlib.md5()
    with open(path, "rb") as f:
        for line in f:
            if not line.startswith(b"#"):
                h.update(line)
    return h.hexdigest()


@pytest.mark.parametrize("n", [128, 192])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_whole_run_goldens(mhd, goldens, tmp_path, n, prec):
    say(f"cli OT {n} {prec}")
    r = mhd.run_cli([n, n, 0.5, 0.25, "muscl", "hlld", 0.18], str(tmp_path), precision=prec)
    assert r.returncode == 0, r.stderr
    g = goldens["ot"][f"{n}_{prec}"]
    m = re.search(r"Done\. Steps: (\d+)", r.stdout)
    assert int(m.group(1)) == g["steps"]
    m = re.search(r"HLLD->HLL fallbacks = (\d+) \(degenerate=(\d+), nonphysical=(\d+)\)", r.stdout)
    assert tuple(map(int, m.groups())) == (g["fallbacks"], g["degenerate"], g["nonphysical"])
    m = re.search(r"divB L1 = (\S+), divB Linf = (\S+), psi Linf = (\S+)", r.stdout)
    assert m.groups() == (g["divB_L1"], g["divB_Linf"], g["psi_Linf"])
    (f,) = list(tmp_path.iterdir())
    assert f.name == f"orszag_tang_glm_muscl_hlld_{prec}_N{n}x{n}.dat"
    assert data_md5(f) == g["md5"]


def test_cli_fp32_vs_fp64_table_192(mhd, tmp_path):
    """The comparator over the FP64 and FP32 .dat files of OT 192^2 (bitwise
    the reference CPU builds' files) prints the reference's own table
    (logs/mhd/raptor_scan/ot2d_cmp_nativefp32.txt) digit for digit: L1, L2,
    Linf, relL2 per variable and the divB lines."""
    import json

    from mhdglm import compare

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cmp_ot192.json")))
    n = g["n"]
    files = {}
    for prec in ("fp64", "fp32"):
        d = tmp_path / prec
        d.mkdir()
        say(f"cli OT {n} {prec} for the comparison table")
        r = mhd.run_cli([n, n, g["t_end"], 0.25, "muscl", "hlld", 0.18], str(d), precision=prec)
        assert r.returncode == 0, r.stderr
        (files[prec],) = list(d.iterdir())
    table = compare.compare_files(str(files["fp64"]), str(files["fp32"]))
    rows = {}
    for line in table.splitlines():
        f = line.split()
        if len(f) == 5 and f[0] in g["native_fp32"]["rows"]:
            rows[f[0]] = f[1:]
    for var, want in g["native_fp32"]["rows"].items():
        assert rows[var] == want["text"], (var, rows[var], want["text"])
    divb = re.findall(r"L1=(\S+)\s+Linf=(\S+)", table)
    assert divb == [(g["native_fp32"]["divB"][k]["L1"], g["native_fp32"]["divB"][k]["Linf"])
                    for k in ("fp64", "fp32")]


# --------------------------------------------------------------------------
# full size: size-independent properties
# --------------------------------------------------------------------------
def test_full_size_conservation_and_determinism(mhd):
    """OT 4096^2, 6 steps: mass, momentum and energy sums are conserved by
    the flux form (periodic), and two runs are bitwise identical."""
    n = 4096
    say("OT 4096^2 conservation")
    outs = []
    for rep in range(2):
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n))
        s.init_problem("ot")
        u0 = s.get_state(0)
        st = s.advance(0.5, 6)
        assert st.code == 0 and st.steps == 6
        outs.append(s.get_state(0))
        s.close()
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    u = outs[0]
    for k in (0, 1, 2, 3, 4, 5, 6, 7):
        tot0, tot1 = u0[k].sum(dtype=np.float64), u[k].sum(dtype=np.float64)
        scale = np.abs(u0[k]).sum(dtype=np.float64)
        assert abs(tot1 - tot0) <= 1e-12 * scale, (k, tot0, tot1)


@pytest.mark.parametrize("nx,ny", [(1, 8), (7, 1), (1, 1), (3, 2)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_degenerate_grids_match_oracle(mhd, tmp_path, prec, nx, ny):
    """Grids narrower than the 2-cell halo through the drop-in CLI: the
    .dat data rows, the step count, the fallback counters and the divB
    line equal the oracle CLI's (whose .dat equals the reference CPU
    binary's byte for byte at these sizes)."""
    import subprocess

    args = [str(nx), str(ny), "0.05", "0.3", "muscl", "hlld", "0.18"]
    gdir, odir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    odir.mkdir()
    say(f"cli degenerate {nx}x{ny} {prec}")
    r = mhd.run_cli(args, str(gdir), precision=prec)
    assert r.returncode == 0, r.stderr
    exe = os.path.join(ROOT, "oracle", "build", f"mhd2d_glm_oracle_{prec}")
    ro = subprocess.run([exe] + args, env=dict(os.environ, OUTDIR=str(odir)), capture_output=True,
                        text=True, timeout=300)
    assert ro.returncode == 0, ro.stderr
    (g,) = list(gdir.iterdir())
    (o,) = list(odir.iterdir())
    assert g.name == o.name
    assert data_md5(g) == data_md5(o)
    for pat in (r"Done\. Steps: (\d+)", r"HLLD->HLL fallbacks = (.*)", r"divB L1 = (.*)"):
        assert re.search(pat, r.stdout).group(1) == re.search(pat, ro.stdout).group(1), pat


@pytest.mark.parametrize("problem", ["kh", "khb"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_kelvin_helmholtz_matches_oracle(mhd, tmp_path, prec, problem):
    """The drop-in CLI on the Kelvin-Helmholtz problems: the file is named
    and titled as the reference GPU driver names and titles its KH output
    (mhd2d_glm_gpu.cu:592-611: base "kh", "# Kelvin-Helmholtz ideal MHD
    with mixed GLM (GPU)"; `khb`, this build's variant, takes its own
    name), and its data rows, step count, fallback counters and divB line
    equal the oracle CLI's on the same problem."""
    import subprocess

    args = ["48", "32", "0.05", "0.3", "muscl", "hlld", "0.18", problem]
    gdir, odir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    odir.mkdir()
    say(f"cli {problem} 48x32 {prec}")
    r = mhd.run_cli(args, str(gdir), precision=prec)
    assert r.returncode == 0, r.stderr
    exe = os.path.join(ROOT, "oracle", "build", f"mhd2d_glm_oracle_{prec}")
    ro = subprocess.run([exe] + args, env=dict(os.environ, OUTDIR=str(odir)), capture_output=True,
                        text=True, timeout=300)
    assert ro.returncode == 0, ro.stderr
    (g,) = list(gdir.iterdir())
    (o,) = list(odir.iterdir())
    assert g.name == o.name == f"{problem}_glm_muscl_hlld_{prec}_N48x32.dat"
    assert open(g).readline() == "# Kelvin-Helmholtz ideal MHD with mixed GLM (GPU)\n"
    assert data_md5(g) == data_md5(o)
    for pat in (r"Done\. Steps: (\d+)", r"HLLD->HLL fallbacks = (.*)", r"divB L1 = (.*)"):
        assert re.search(pat, r.stdout).group(1) == re.search(pat, ro.stdout).group(1), pat


@pytest.mark.parametrize("order,solver", [("muscl", "hlld"), ("muscl", "hll"), ("first", "hlld"),
                                          ("first", "hll")])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_dat_header_matches_reference_cpu_build(mhd, tmp_path, prec, order, solver):
    """The .dat of the drop-in CLI for every order and solver: the reference
    GPU build writes the CPU build's '#' lines with ' (GPU)' appended to the
    first (mhd2d_glm_gpu.cu:610-625 against mhd2d_glm_cpu.cpp:628-645), the
    HLLD statistics line for the HLLD solver only.  The CPU side is the
    oracle's CLI, whose .dat equals the reference CPU binary's byte for byte
    for all four schemes; the run is bitwise, so t, ch_last, the counters
    and the data rows agree too."""
    import subprocess

    args = ["64", "48", "0.05", "0.25", order, solver, "0.18"]
    gdir, odir = tmp_path / "gpu", tmp_path / "cpu"
    gdir.mkdir()
    odir.mkdir()
    say(f"cli header {prec} {order} {solver}")
    r = mhd.run_cli(args, str(gdir), precision=prec)
    assert r.returncode == 0, r.stderr
    exe = os.path.join(ROOT, "oracle", "build", f"mhd2d_glm_oracle_{prec}")
    ro = subprocess.run([exe] + args, env=dict(os.environ, OUTDIR=str(odir)), capture_output=True,
                        text=True, timeout=300)
    assert ro.returncode == 0, ro.stderr
    (g,) = list(gdir.iterdir())
    (o,) = list(odir.iterdir())
    assert g.name == o.name
    gh = [ln for ln in open(g) if ln.startswith("#")]
    oh = [ln for ln in open(o) if ln.startswith("#")]
    # The HLLD statistics line is written for the HLLD solver only.
    assert len(gh) == len(oh) == (4 if solver == "hlld" else 3)
    assert gh[0] == oh[0].rstrip("\n") + " (GPU)\n"
    assert gh[1:] == oh[1:]
    assert data_md5(g) == data_md5(o)
