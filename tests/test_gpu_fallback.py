"""GPU tests of the mixed form's FP64 fallback.

A mixed stage tile whose FP32-flux update fails the positivity check writes
nothing and is re-evaluated by the FP64 kernel; only an FP64 failure halts.
  * check mode (every tile re-evaluated) is bitwise the FP64 form, for a
    single grid and for an ensemble;
  * when nothing fails, the fallback changes no bit;
  * a low-pressure, high-Mach contact on which the FP32 flux breaks
    positivity within a few steps: with the fallback the mixed form runs on,
    close to FP64;
  * a mixed run that halts with an FP64-confirmed unusable cell (injected
    at a given step through the MHD_FAULT_STEP test hook, as a trajectory
    that drifted into a state the FP64 tile fallback cannot advance would):
    the rollback re-runs an earlier window in FP64 and the run finishes, for
    a single grid and inside an ensemble alike (without it the run halts, as
    the other forms do);
  * the Orszag-Tang variant on which such a drift was first seen finishes in
    the mixed form, close to FP64;
  * a state that the FP64 form cannot advance either: the rollback ends in a
    restart that is the FP64 run, bitwise, halt included.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# Ensemble-workload instance (seed 2026, index 16) whose mixed trajectory
# drifted (with the build of that time) into a state that halted in FP64.
OTV_HARD = (1.067437115364036, 0.8064116573994691, 1.759240013263218, 0.20557563316160354,
            0.5272182398647961)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(mhd, n, precision, steps=-1, t_end=0.5, fallback=0, problem="ot", params=()):
    # 16x15 tiles whatever the tuned table picks: the tests count tiles.
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=precision, fallback=fallback, tile=3))
    if params:
        s.set_state(mhd.problem_state_ex(problem, params, n, n))
    else:
        s.init_problem(problem)
    st = s.advance(t_end, steps)
    u = s.get_state(0)
    s.close()
    return u, st


@pytest.mark.parametrize("n", [96, 128])
def test_check_mode_is_fp64(mhd, n):
    say(f"fallback check mode OT {n}^2, 40 steps")
    u64, s64 = run(mhd, n, mhd.PREC_FP64, steps=40)
    uc, sc = run(mhd, n, mhd.PREC_MIXED, steps=40, fallback=2)
    assert sc.code == 0 and sc.steps == 40
    assert np.array_equal(u64.view(np.uint64), uc.view(np.uint64))
    tiles = -(-n // 16) * -(-n // 15)
    assert sc.fallback_tiles == 2 * 40 * tiles
    assert s64.fallback_tiles == 0


def test_check_mode_ensemble_is_fp64(mhd):
    say("fallback check mode, ensemble of 3")
    sizes = [(64, 64), (96, 80), (130, 70)]
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PREC_MIXED, fallback=2)
    ens = mhd.Ensemble(cfg, sizes)
    for i in range(3):
        ens.init_problem(i, "ot")
    sts = ens.advance(10.0, 30)
    got = [ens.get_state(i) for i in range(3)]
    ens.close()
    for i, (nx, ny) in enumerate(sizes):
        s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=mhd.PREC_FP64))
        s.init_problem("ot")
        s.advance(10.0, 30)
        ref = s.get_state(0)
        s.close()
        assert sts[i].code == 0 and sts[i].fallback_tiles > 0
        assert np.array_equal(ref.view(np.uint64), got[i].view(np.uint64)), i


def test_untriggered_fallback_changes_nothing(mhd):
    say("OT 128^2 mixed with and without fallback")
    ua, sa = run(mhd, 128, mhd.PREC_MIXED, fallback=0)
    ub, sb = run(mhd, 128, mhd.PREC_MIXED, fallback=1)
    assert sa.code == sb.code == 0 and sa.done == 1
    assert sa.fallback_tiles == 0
    assert np.array_equal(ua.view(np.uint64), ub.view(np.uint64))


def low_pressure_contact(n):
    """Density wave advected at Mach ~ 4e4 through a pressure of 1e-9: the
    FP32 round-off of the energy flux (~ 6e-8 rho v^3) exceeds the internal
    energy, the FP64 flux's does not."""
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


def test_fallback_rescues_fp32_flux_failure(mhd):
    n, steps = 64, 20
    say("low-pressure contact 64^2")
    runs = {}
    for name, prec, fb in (("fp64", mhd.PREC_FP64, 0), ("off", mhd.PREC_MIXED, 1),
                           ("on", mhd.PREC_MIXED, 0)):
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec, fallback=fb))
        s.set_state(low_pressure_contact(n))
        st = s.advance(10.0, steps)
        runs[name] = (s.get_state(0), st)
        s.close()
        say(f"  {name}: code {st.code}, steps {st.steps}, fallback tiles {st.fallback_tiles}")
    u64, s64 = runs["fp64"]
    assert s64.code == 0 and s64.steps == steps
    assert runs["off"][1].code in (3, 4)          # the FP32 flux breaks positivity
    um, sm = runs["on"]
    assert sm.code == 0 and sm.steps == steps and sm.fallback_tiles > 0
    from mhdglm import compare

    errs = compare.rel_l2_state(u64, um, ("rho", "p", "|B|"))
    say(f"  relL2 vs fp64 {errs}")
    p = compare.fields_from_state(um)["p"]
    assert p.min() > 0
    assert errs["rho"] < 1e-6 and errs["p"] < 1e-2


def run_rb(mhd, n, precision, rollback_steps=0, fallback=0, t_end=0.5, steps=-1, u0=None):
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=precision, fallback=fallback,
                                    rollback_steps=rollback_steps))
    s.set_state(mhd.problem_state_ex("otv", OTV_HARD, n, n) if u0 is None else u0)
    st = s.advance(t_end, steps)
    u = s.get_state(0)
    s.close()
    return u, st


def test_rollback_rescues_injected_failure(mhd, monkeypatch):
    """Orszag-Tang 256^2 in the mixed form, with stage 1 of step 1000 made to
    report cell 0 unusable as the FP64 fallback would (MHD_FAULT_STEP):
    without the rollback the run halts there with code 3 (the reference's
    contract); with it the run returns to the newest snapshot at least 256
    steps earlier, re-runs from there in FP64 until 256 steps past the
    failure (the injection does not fire inside the window), then finishes
    mixed, at the mixed form's distance from FP64.  Inside an ensemble both
    instances (each injected at its step 1000) do the same, bitwise."""
    from mhdglm import compare

    n, k = 256, 1000

    def ot(prec, rollback_steps=0):
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec, rollback_steps=rollback_steps))
        s.init_problem("ot")
        st = s.advance(0.5)
        u = s.get_state(0)
        s.close()
        return u, st

    u64, s64 = ot(mhd.PREC_FP64)
    assert s64.code == 0 and s64.done == 1
    monkeypatch.setenv("MHD_FAULT_STEP", str(k))
    _, f64 = ot(mhd.PREC_FP64)
    assert f64.code == 0 and f64.steps == s64.steps  # the hook is the mixed form's only
    _, off = ot(mhd.PREC_MIXED, rollback_steps=-1)
    say(f"  rollback off: code {off.code} step {off.fail_step} cell ({off.bad_i},{off.bad_j})")
    assert (off.code, off.fail_step, off.bad_i, off.bad_j, off.steps) == (3, k, 0, 0, k)
    um, sm = ot(mhd.PREC_MIXED)
    errs = compare.rel_l2_state(u64, um, ("rho", "p", "|B|"))
    say(f"  rollback on: code {sm.code} done {sm.done} steps {sm.steps} rollbacks {sm.rollbacks} "
        f"restarts {sm.restarts} FP64 steps {sm.exact_steps} relL2 {errs}")
    assert sm.code == 0 and sm.done == 1
    assert sm.rollbacks == 1 and sm.restarts == 0
    assert 256 < sm.exact_steps <= 3 * 256
    assert max(errs.values()) < 1e-3
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PREC_MIXED)
    ens = mhd.Ensemble(cfg, [(128, 128), (n, n)])
    for i in range(2):
        ens.init_problem(i, "ot")
    sts = ens.advance(0.5)
    ue = ens.get_state(1)
    u128 = ens.get_state(0)
    ens.close()
    s = mhd.Solver(mhd.SolverConfig(nx=128, ny=128, precision=mhd.PREC_MIXED))
    s.init_problem("ot")
    st128 = s.advance(0.5)
    u128s = s.get_state(0)
    s.close()
    assert st128.rollbacks == 1 and st128.code == 0 and st128.done == 1
    for got, ref, st in ((sts[0], st128, u128), (sts[1], sm, ue)):
        assert got.code == 0 and got.done == 1
        assert (got.rollbacks, got.exact_steps, got.steps) == (ref.rollbacks, ref.exact_steps,
                                                              ref.steps)
    assert np.array_equal(ue.view(np.uint64), um.view(np.uint64))
    assert np.array_equal(u128.view(np.uint64), u128s.view(np.uint64))


def test_hard_variant_finishes_mixed(mhd):
    """The Orszag-Tang variant of the ensemble workload (seed 2026, index
    16) at 1024^2, on which a mixed trajectory was seen to drift into a
    state the FP64 tile fallback could not advance: the mixed form finishes
    it (through a rollback window if it drifts so again) at the mixed form's
    distance from its FP64 run."""
    from mhdglm import compare

    n = 1024
    say("hard OT variant 1024^2")
    u64, s64 = run_rb(mhd, n, mhd.PREC_FP64)
    assert s64.code == 0 and s64.done == 1
    um, sm = run_rb(mhd, n, mhd.PREC_MIXED)
    errs = compare.rel_l2_state(u64, um, ("rho", "p", "|B|"))
    say(f"  mixed: code {sm.code} done {sm.done} rollbacks {sm.rollbacks} restarts {sm.restarts} "
        f"relL2 {errs}")
    assert sm.code == 0 and sm.done == 1 and sm.restarts == 0
    assert max(errs.values()) < 1e-3


def marginal_state(mhd, oracle, n=32):
    """Low pressure and a fast square: FP64 itself fails after some steps."""
    u = oracle.init("ot", n, n)
    W = oracle.cons_to_prim(u.reshape(-1, 9)).reshape(u.shape)
    W[..., 4] *= 0.02
    W[10:20, 10:20, 1] *= 8.0
    u = oracle.prim_to_cons(W.reshape(-1, 9)).reshape(W.shape)
    return np.ascontiguousarray(np.moveaxis(u, -1, 0))


def test_rollback_restart_ends_as_fp64(mhd, oracle):
    """A state the FP64 form itself cannot advance past some step: the mixed
    run's rollback window fails in FP64 too, so the run restarts from the
    given state in FP64 throughout, and halts exactly as the FP64 form does
    (step, code, bad cell and state, bitwise)."""
    u0 = marginal_state(mhd, oracle)
    runs = {}
    for name, prec in (("fp64", mhd.PREC_FP64), ("mixed", mhd.PREC_MIXED)):
        s = mhd.Solver(mhd.SolverConfig(nx=32, ny=32, precision=prec, cfl=0.9, rollback_steps=16))
        s.set_state(u0)
        st = s.advance(0.5, 400)
        runs[name] = (s.get_state(0), st)
        s.close()
        say(f"  {name}: code {st.code} steps {st.steps} fail {st.fail_step} bad ({st.bad_i},{st.bad_j}) "
            f"rollbacks {st.rollbacks} restarts {st.restarts}")
    (u64, s64), (um, sm) = runs["fp64"], runs["mixed"]
    assert s64.code in (3, 4)
    assert sm.restarts == 1
    assert (sm.code, sm.fail_step, sm.bad_i, sm.bad_j, sm.steps) == (
        s64.code, s64.fail_step, s64.bad_i, s64.bad_j, s64.steps)
    assert sm.exact_steps == s64.steps
    assert np.array_equal(um.view(np.uint64), u64.view(np.uint64))


def test_init_problem_rejects_parameterised_problem(mhd):
    """"otv" needs five parameters: mhd_init_problem refuses it instead of
    uploading an empty state."""
    s = mhd.Solver(mhd.SolverConfig(nx=32, ny=32))
    with pytest.raises(mhd.MhdError) as ei:
        s.init_problem("otv")
    assert ei.value.code == 1 and "parameters" in str(ei.value)
    s.close()
