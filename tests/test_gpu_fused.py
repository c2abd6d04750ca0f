"""GPU tests of the two-stage step (fusion depth 2: both SSPRK2 stages in one
kernel launch, stage 1 recomputed on a 2-cell ring around each tile).

  * exact forms (FP64, FP32, the reference-mixed split): bitwise the CPU
    oracle and the one-stage-per-launch path, HLLD counters included, on odd
    grid sizes (partial tiles, grids smaller than a tile), noisy states,
    graph batches, a t_end inside a batch, and odd step counts (the state
    then ends in the alternate buffer);
  * positivity failures: the exit code, step, first bad cell and the state
    left behind are the oracle's;
  * mixed forms: within round-off of the one-stage path; a step in which a
    tile meets an unusable cell is recomputed by the one-stage kernels with
    their FP64 fallback;
  * checkpoint/restart across the buffer alternation, and the
    configuration rules.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PREC = {"fp64": (0, np.float64), "fp32": (1, np.float32)}
TILE_16x15 = 3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_bitwise(a, b, what=""):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, what
    ui = np.uint64 if a.dtype == np.float64 else np.uint32
    diff = np.nonzero(a.view(ui) != b.view(ui))
    assert diff[0].size == 0, f"{what}: {diff[0].size} elements differ, first at {[d[0] for d in diff]}"


def solver(mhd, nx, ny, precision, fuse, **kw):
    return mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=precision, tile=TILE_16x15,
                                       fuse=fuse, **kw))


def gpu_run(mhd, u_aos, steps, precision, fuse, t_end=0.5, **kw):
    ny, nx, _ = u_aos.shape
    s = solver(mhd, nx, ny, precision, fuse, **kw)
    s.set_state(np.moveaxis(u_aos, -1, 0))
    st = s.advance(t_end, steps)
    out = np.moveaxis(s.get_state(0), 0, -1)
    s.close()
    return out, st


def check_oracle(mhd, oracle, u, steps, prec, **kw):
    p, _ = PREC[prec]
    ug, st = gpu_run(mhd, u, steps, p, 2, **kw)
    okw = {k: v for k, v in kw.items() if k in ("t_end", "cfl")}
    uo, info = oracle.advance(u, steps, **okw)
    assert st.code == info["code"] == 0
    assert st.steps == info["steps"]
    assert_bitwise(np.array([st.t], dtype=u.dtype), np.array([info["t"]], dtype=u.dtype), "t")
    assert_bitwise(ug, uo, "state")
    cnt = info["counters"]
    assert (st.x_evaluations, st.y_evaluations, st.degenerate, st.nonphysical) == \
        tuple(int(c) for c in cnt)
    return st


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,ny", [(64, 48), (50, 37), (10, 9), (33, 70), (1, 9), (3, 2)])
def test_fused_steps_match_oracle(mhd, oracle, prec, nx, ny):
    _, dt = PREC[prec]
    say(f"fused steps OT {nx}x{ny} {prec}, 13 steps")
    u = oracle.init("ot", nx, ny, dtype=dt)
    check_oracle(mhd, oracle, u, 13, prec)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_noisy_state_and_counters(mhd, oracle, prec):
    """Noisy states reach the HLLD fallbacks (counted once per grid face)."""
    _, dt = PREC[prec]
    rng = np.random.default_rng(17)
    u = oracle.init("ot", 48, 48).astype(np.float64)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W *= 1.0 + 0.3 * rng.standard_normal(W.shape)
    W[..., 0] = np.abs(W[..., 0]) + 0.01
    W[..., 4] = np.abs(W[..., 4]) + 0.01
    W[..., 8] = 0.05 * rng.standard_normal(W.shape[:2])
    W[5:9, 10:14, 5] = 0.0
    W[5:9, 10:14, 6] = 0.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape).astype(dt)
    say(f"fused noisy 48x48 {prec}")
    st = check_oracle(mhd, oracle, u, 15, prec, cfl=0.2)
    assert st.degenerate > 0


@pytest.mark.parametrize("problem", ["kh", "khb"])
def test_fused_kelvin_helmholtz(mhd, oracle, problem):
    say(f"fused {problem} 64x64 fp64")
    u = oracle.init(problem, 64, 64)
    check_oracle(mhd, oracle, u, 31, "fp64")


def test_fused_graph_batches_and_t_end(mhd, oracle):
    u = oracle.init("ot", 64, 48)
    uo, info = oracle.advance(u, 41)
    for batch in (1, 7, 32):
        say(f"fused 64x48, 41 steps, batch {batch}")
        ug, st = gpu_run(mhd, u, 41, 0, 2, batch_steps=batch)
        assert st.steps == 41
        assert_bitwise(ug, uo, f"batch={batch}")
    say("fused 32x32 to t=0.03")
    u = oracle.init("ot", 32, 32)
    ug, st = gpu_run(mhd, u, 10000, 0, 2, t_end=0.03)
    uo, info = oracle.advance(u, 10000, t_end=0.03)
    assert st.done == 1 and st.steps == info["steps"] and st.t == 0.03
    assert_bitwise(ug, uo)


def test_fused_matches_one_stage_whole_runs(mhd):
    """Every exact form bitwise the one-stage path; graph-batched to t_end."""
    for name, prec in (("fp64", mhd.PREC_FP64), ("fp32", mhd.PREC_FP32),
                       ("ref_mixed", mhd.PREC_REF_MIXED)):
        say(f"fused vs one-stage OT 96x80 {name}, 61 steps")
        out = {}
        for fuse in (1, 2):
            s = solver(mhd, 96, 80, prec, fuse)
            s.init_problem("ot")
            st = s.advance(0.5, 61)
            out[fuse] = (s.get_state(0), st)
            s.close()
        (a, sa), (b, sb) = out[1], out[2]
        assert sa.steps == sb.steps == 61
        assert (sa.degenerate, sa.nonphysical, sa.t) == (sb.degenerate, sb.nonphysical, sb.t)
        assert_bitwise(a, b, name)


def test_fused_failure_stage1(mhd, oracle):
    u = oracle.init("ot", 32, 32)
    for (j, i) in ((20, 7), (9, 30)):
        u[j, i, 4] = 1e-8
    say("fused stage-1 failure 32x32")
    ug, st = gpu_run(mhd, u, 100, 0, 2)
    _, info = oracle.advance(u, 100)
    assert info["code"] == 3 and st.code == 3
    assert (st.bad_i, st.bad_j) == info["bad"]
    assert st.fail_step == info["steps"] == 0
    assert_bitwise(ug, u)


def test_fused_failure_later(mhd, oracle):
    """A marginal state that survives some steps, then fails: code, step, cell,
    the state left behind and the counters are the oracle's."""
    u = oracle.init("ot", 32, 32)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W[..., 4] *= 0.02
    W[10:20, 10:20, 1] *= 8.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape)
    say("fused marginal state 32x32, cfl 0.9")
    ug, st = gpu_run(mhd, u, 400, 0, 2, cfl=0.9)
    uo, info = oracle.advance(u, 400, cfl=0.9)
    assert st.code == info["code"]
    assert st.steps == info["steps"]
    if info["code"] in (3, 4):
        assert (st.bad_i, st.bad_j) == info["bad"]
    assert_bitwise(ug, uo)
    cnt = info["counters"]
    assert (st.degenerate, st.nonphysical) == (int(cnt[2]), int(cnt[3]))


def rel_l2(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("form,tol", [("mixed", 1e-6), ("mixed_kahan", 1e-6),
                                      ("fp64_fma", 1e-12)])
def test_fused_forms_within_tolerance_of_fp64(mhd, form, tol):
    """The contraction-on forms are not bitwise across kernels (each
    instantiation may contract its FMAs differently): both depths stay within
    the form's tolerance of FP64, as every tile shape does."""
    say(f"fused {form} OT 128x128, 25 steps")
    u = {}
    for name, prec, fuse in (("fp64", mhd.PREC_FP64, 1), ("one", mhd.PRECISIONS[form], 1),
                             ("two", mhd.PRECISIONS[form], 2)):
        s = solver(mhd, 128, 128, prec, fuse)
        s.init_problem("ot")
        st = s.advance(0.5, 25)
        u[name] = s.get_state(0).astype(np.float64)
        s.close()
        assert st.code == 0 and st.steps == 25 and st.fallback_tiles == 0, name
        assert st.fused_redo_steps == 0, name
    for k in range(8):
        scale = np.sqrt(np.mean(u["fp64"][k] ** 2)) + 1e-300
        for name in ("one", "two"):
            err = np.sqrt(np.mean((u[name][k] - u["fp64"][k]) ** 2)) / scale
            assert err < tol, (name, k, err)


def low_pressure_contact(n):
    """Density wave advected at Mach ~ 4e4 through a pressure of 1e-9: the FP32
    round-off of the energy flux exceeds the internal energy."""
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.5 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    vx, p, bx = 1.0, 1e-9, 1e-3
    u = np.zeros((9, n, n))
    u[0] = rho
    u[1] = rho * vx
    u[5] = bx
    u[4] = p / (5.0 / 3.0 - 1.0) + 0.5 * rho * vx * vx + 0.5 * bx * bx
    return u


def test_fused_mixed_recomputes_voted_steps(mhd):
    """A step in which a tile meets an unusable cell is recomputed by the
    one-stage kernels (tile-level FP64 fallback): the run goes on like the
    one-stage mixed run; without the fallback it halts like it."""
    n, steps = 64, 20
    s = solver(mhd, n, n, mhd.PREC_FP64, 1)
    s.set_state(low_pressure_contact(n))
    s64 = s.advance(10.0, steps)
    u64 = s.get_state(0)
    s.close()
    assert s64.code == 0 and s64.steps == steps
    runs = {}
    for fuse in (1, 2):
        for fb in (0, 1):
            s = solver(mhd, n, n, mhd.PREC_MIXED, fuse, fallback=fb)
            s.set_state(low_pressure_contact(n))
            st = s.advance(10.0, steps)
            runs[(fuse, fb)] = (s.get_state(0), st)
            s.close()
            say(f"  fuse {fuse} fallback {fb}: code {st.code}, steps {st.steps}, "
                f"fallback tiles {st.fallback_tiles}, recomputed steps {st.fused_redo_steps}")
    u1, s1 = runs[(1, 0)]
    u2, s2 = runs[(2, 0)]
    assert s1.code == s2.code == 0 and s1.steps == s2.steps == steps
    assert s2.fused_redo_steps > 0 and s2.fallback_tiles > 0
    assert s1.fused_redo_steps == 0
    assert rel_l2(u1[0], u64[0]) < 1e-6 and rel_l2(u2[0], u64[0]) < 1e-6
    assert runs[(1, 1)][1].code in (3, 4)
    assert runs[(2, 1)][1].code == runs[(1, 1)][1].code
    assert runs[(2, 1)][1].steps == runs[(1, 1)][1].steps


def test_fused_checkpoint_restart_and_state_access(mhd, tmp_path):
    """Odd step counts leave the state in the alternate buffer; every reader
    sees the current state, and a restart continues bitwise."""
    s = solver(mhd, 64, 48, mhd.PREC_FP64, 2)
    s.init_problem("ot")
    s.advance(0.5, 11)
    ref = s.get_state(0)
    with pytest.raises(mhd.MhdError):
        s.get_state(1)
    s.close()
    s = solver(mhd, 64, 48, mhd.PREC_FP64, 2)
    s.init_problem("ot")
    s.advance(0.5, 5)
    path = str(tmp_path / "ck.bin")
    s.save(path)
    s.close()
    s = solver(mhd, 64, 48, mhd.PREC_FP64, 2)
    s.load(path)
    st = s.advance(0.5, 6)
    assert st.steps == 11
    assert_bitwise(s.get_state(0), ref, "restart")
    s.close()
    s = solver(mhd, 64, 48, mhd.PREC_FP64, 1)
    s.init_problem("ot")
    s.advance(0.5, 11)
    assert_bitwise(s.get_state(0), ref, "one-stage")
    s.close()


def test_fused_configuration_rules(mhd):
    with pytest.raises(mhd.MhdError):
        mhd.Solver(mhd.SolverConfig(nx=32, ny=32, tile=0, fuse=2))
    with pytest.raises(mhd.MhdError):
        mhd.Solver(mhd.SolverConfig(nx=32, ny=32, muscl=False, fuse=2))
    with pytest.raises(mhd.MhdError):
        mhd.Solver(mhd.SolverConfig(nx=32, ny=32, precision=mhd.PREC_MIXED, fallback=2, fuse=2))
    with pytest.raises(mhd.MhdError):
        mhd.Solver(mhd.SolverConfig(nx=32, ny=32, fuse=3))
    s = mhd.Solver(mhd.SolverConfig(nx=32, ny=32, tile=TILE_16x15, fuse=2))
    s.close()
