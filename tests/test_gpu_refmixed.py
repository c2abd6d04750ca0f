"""GPU tests of the reference's stage-level mixed form (precision ref_mixed,
after src/mhd/mhd2d_glm_gpu_mixed.cu): FP64 state, Riemann fluxes and SSPRK2
updates; FP32 recovery, MUSCL reconstruction, face prim->cons, rhs flux
differences, psi damping and CFL speeds.

Bar: bitwise equality with the oracle's restatement (oracle/mhd_refmixed.hpp,
itself pinned against the reference's FP32 header functions) on whole steps
-- state, step count, t and HLLD fallback counters -- over both solvers,
every tile shape, odd sizes, noisy states and a positivity failure; an
ensemble instance is bitwise the single-grid run; and at 512^2 to t=0.5 its
distance from FP64 is of the size the reference logged for its own mixed
build against its FP64 build (tests/golden/refmixed_ot512.json).
"""
import json
import os
import re
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
UINT = {np.float64: np.uint64, np.float32: np.uint32}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_bitwise(a, b, what=""):
    """Bitwise equality; NaNs compare equal whatever their sign and payload."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), f"{what}: NaN positions differ"
    ba, bb = a.view(UINT[a.dtype.type]), b.view(UINT[b.dtype.type])
    bad = np.argwhere((ba != bb) & ~na)
    assert bad.size == 0, f"{what}: {len(bad)} mismatches, first {tuple(bad[0])}: {a[tuple(bad[0])]!r} vs {b[tuple(bad[0])]!r}"


def gpu_run(mhd, u_aos, steps, hlld=True, tile=-1, t_end=0.5, cfl=0.25, batch=0):
    ny, nx, _ = u_aos.shape
    s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, cfl=cfl, hlld=hlld, tile=tile,
                                    batch_steps=batch, precision=mhd.PREC_REF_MIXED))
    s.set_state(np.moveaxis(u_aos, -1, 0))
    st = s.advance(t_end, steps)
    out = np.moveaxis(s.get_state(0), 0, -1)
    s.close()
    return out, st


def check_steps(mhd, oracle, u, steps, hlld=True, tile=-1, cfl=0.25, t_end=0.5):
    ug, st = gpu_run(mhd, u, steps, hlld=hlld, tile=tile, cfl=cfl, t_end=t_end)
    uo, info = oracle.refmixed_advance(u, steps, hlld=hlld, cfl=cfl, t_end=t_end)
    assert st.code == info["code"]
    assert st.steps == info["steps"]
    if info["code"] in (3, 4):
        assert (st.bad_i, st.bad_j) == info["bad"]
        return st, info
    assert_bitwise(np.array([st.t]), np.array([info["t"]]), "t")
    assert_bitwise(ug, uo, "state")
    if hlld:
        cnt = tuple(int(c) for c in info["counters"])
        assert (st.x_evaluations, st.y_evaluations, st.degenerate, st.nonphysical) == cnt
    return st, info


@pytest.mark.parametrize("hlld", [True, False])
@pytest.mark.parametrize("problem", ["ot", "kh", "khb"])
def test_steps_bitwise(mhd, oracle, problem, hlld):
    say(f"ref_mixed steps {problem} 64x48 hlld={hlld}")
    u = oracle.init(problem, 64, 48)
    st, _ = check_steps(mhd, oracle, u, 30, hlld=hlld)
    assert st.code == 0 and st.steps == 30


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("nx,ny", [(50, 37), (33, 70), (2, 5)])
def test_steps_tiles_and_odd_sizes(mhd, oracle, tile, nx, ny):
    say(f"ref_mixed steps OT {nx}x{ny} tile={tile}")
    check_steps(mhd, oracle, oracle.init("ot", nx, ny), 20, tile=tile)


def test_steps_noisy_state(mhd, oracle):
    """Noisy primitives reach the HLLD fallbacks and the first-order
    reconstruction fallback of the FP32 recon."""
    rng = np.random.default_rng(17)
    u = oracle.init("ot", 48, 48)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W *= 1.0 + 0.3 * rng.standard_normal(W.shape)
    W[..., 0] = np.abs(W[..., 0]) + 0.01
    W[..., 4] = np.abs(W[..., 4]) + 0.01
    W[..., 8] = 0.05 * rng.standard_normal(W.shape[:2])
    W[5:9, 10:14, 5] = 0.0
    W[5:9, 10:14, 6] = 0.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape)
    say("ref_mixed steps noisy 48x48")
    st, _ = check_steps(mhd, oracle, u, 15, cfl=0.2)
    assert st.code == 0 and st.degenerate > 0


def test_positivity_failure_matches_oracle(mhd, oracle):
    """A low-pressure fast patch at CFL 0.9: both halt on the same step,
    stage and first bad cell."""
    u = oracle.init("ot", 32, 32)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W[..., 4] *= 0.02
    W[10:20, 10:20, 1] *= 8.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape)
    say("ref_mixed positivity failure 32x32 cfl 0.9")
    st, info = check_steps(mhd, oracle, u, 400, cfl=0.9)
    assert info["code"] in (3, 4)


def test_t_end_and_graph_batches(mhd, oracle):
    """Whole short run to t_end (last dt clipped), and step batches of 1 and
    the default give the same bits."""
    u = oracle.init("ot", 64, 48)
    say("ref_mixed to t=0.03")
    st, info = check_steps(mhd, oracle, u, 10000, t_end=0.03)
    assert st.done == 1 and info["code"] == 0
    a, _ = gpu_run(mhd, u, 40, batch=1)
    b, _ = gpu_run(mhd, u, 40, batch=0)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_first_order_rejected(mhd):
    with pytest.raises(mhd.MhdError):
        mhd.Solver(mhd.SolverConfig(nx=16, ny=16, muscl=False, precision=mhd.PREC_REF_MIXED))


def test_ensemble_matches_single(mhd):
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PREC_REF_MIXED)
    sizes = [(64, 64), (48, 80), (33, 57)]
    params = [(1.0, 0, 0, 0, 0), (0.9, 0.3, 1.1, 2.0, 5.5), (1.05, 0.5, 0.5, 0.5, 0.5)]
    say(f"ref_mixed ensemble {sizes}")
    ens = mhd.Ensemble(cfg, sizes)
    for i, p in enumerate(params):
        ens.init_problem(i, "otv", p)
    sts = ens.advance(0.05)
    for i, ((nx, ny), p) in enumerate(zip(sizes, params)):
        s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=mhd.PREC_REF_MIXED))
        s.set_state(mhd.problem_state_ex("otv", p, nx, ny))
        st = s.advance(0.05)
        ref = s.get_state(0)
        s.close()
        assert sts[i].code == 0 and sts[i].done == 1
        assert sts[i].steps == st.steps and sts[i].t == st.t
        assert np.array_equal(ens.get_state(i).view(np.uint64), ref.view(np.uint64)), i
    ens.close()


def test_ot512_distance_from_fp64_as_reference_logged(mhd):
    """512^2 Orszag-Tang to t=0.5 (5239 steps in the reference's log): relL2
    of the form against FP64 per field, within a factor 4 of the reference's
    mixed-vs-FP64 figures (the run is chaotic, so only the size is fixed)."""
    g = json.load(open(os.path.join(HERE, "golden", "refmixed_ot512.json")))
    n = g["n"]
    say("ref_mixed vs fp64 OT 512^2 to t=0.5")
    out = {}
    for name, prec in (("fp64", mhd.PREC_FP64), ("ref_mixed", mhd.PREC_REF_MIXED)):
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec))
        s.init_problem("ot")
        st = s.advance(g["t_end"])
        out[name] = (s.get_state(0), st)
        s.close()
        say(f"  {name}: steps {st.steps}, code {st.code}")
        assert st.code == 0 and st.done == 1
    assert out["fp64"][1].steps == g["steps"]
    assert abs(out["ref_mixed"][1].steps - g["steps"]) <= 2
    from mhdglm import compare

    names = ("rho", "vx", "vy", "p", "Bx", "By", "E")
    errs = compare.rel_l2_state(out["fp64"][0], out["ref_mixed"][0], names)
    for k in names:
        ref = g["rows"][k]["relL2"]
        say(f"  {k}: relL2 {errs[k]:.3e} (reference {ref:.3e})")
        assert ref / 4 <= errs[k] <= 4 * ref, k


def test_ot192_distance_from_fp64_as_truncation_scan(mhd):
    """192^2 Orszag-Tang to t=0.5: the form against FP64, per field, within
    25% of the reference's scan that truncates the same regions (recovery,
    CFL, reconstruction, rhs, GLM source) to FP32 mantissas while the flux
    and update stay FP64 (ot2d_table.txt row keepFU); and far below the
    native FP32 build's distance at the same size."""
    g = json.load(open(os.path.join(HERE, "golden", "cmp_ot192.json")))
    n = g["n"]
    say("ref_mixed vs fp64 OT 192^2 to t=0.5")
    out = {}
    for name, prec in (("fp64", mhd.PREC_FP64), ("ref_mixed", mhd.PREC_REF_MIXED)):
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec))
        s.init_problem("ot")
        st = s.advance(g["t_end"])
        out[name] = s.get_state(0)
        s.close()
        assert st.code == 0 and st.done == 1
    from mhdglm import compare

    names = ("rho", "vx", "vy", "p", "Bx", "By", "E")
    errs = compare.rel_l2_state(out["fp64"], out["ref_mixed"], names)
    for k in names:
        keep, native = g["keepFU"]["rows"][k]["relL2"], g["native_fp32"]["rows"][k]["relL2"]
        say(f"  {k}: relL2 {errs[k]:.3e} (truncation scan {keep:.3e}, native FP32 {native:.3e})")
        assert keep / 1.25 <= errs[k] <= 1.25 * keep, k
        assert errs[k] < native / 10, k


def test_cli_program_name_and_env(mhd, tmp_path):
    """A program named as the reference's mixed build runs this form and
    writes the reference's file tag; MHD_PRECISION=ref_mixed selects it too
    and the first-order scheme is refused."""
    link = tmp_path / "mhd2d_glm_gpu_mixed"
    os.symlink(mhd.CLI_PATH, link)
    outdir = tmp_path / "out"
    outdir.mkdir()
    env = {k: v for k, v in os.environ.items() if k != "MHD_PRECISION"}
    env["OUTDIR"] = str(outdir)
    import subprocess

    say("cli _gpu_mixed 64^2")
    r = subprocess.run([str(link), "64", "64", "0.05", "0.25", "muscl", "hlld", "0.18"], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    (f,) = list(outdir.iterdir())
    assert f.name == "orszag_tang_glm_muscl_hlld_mixed_N64x64.dat"
    steps = int(re.search(r"Done\. Steps: (\d+)", r.stdout).group(1))
    (tmp_path / "o2").mkdir()
    (tmp_path / "o3").mkdir()
    r2 = mhd.run_cli([64, 64, 0.05, 0.25, "muscl", "hlld", 0.18], str(tmp_path / "o2"),
                     precision="ref_mixed")
    assert r2.returncode == 0, r2.stderr
    assert int(re.search(r"Done\. Steps: (\d+)", r2.stdout).group(1)) == steps
    r3 = mhd.run_cli([64, 64, 0.05, 0.25, "first", "hlld", 0.18], str(tmp_path / "o3"),
                     precision="ref_mixed")
    assert r3.returncode == 1 and "muscl" in r3.stderr
