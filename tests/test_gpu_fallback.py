"""GPU tests of the mixed form's FP64 fallback.

A mixed stage tile whose FP32-flux update fails the positivity check writes
nothing and is re-evaluated by the FP64 kernel; only an FP64 failure halts.
  * check mode (every tile re-evaluated) is bitwise the FP64 form, for a
    single grid and for an ensemble;
  * when nothing fails, the fallback changes no bit;
  * a low-pressure, high-Mach contact on which the FP32 flux breaks
    positivity within a few steps: with the fallback the mixed form runs on,
    close to FP64;
  * the limit: an Orszag-Tang variant whose mixed trajectory drifts (by
    t ~ 0.42) into a state from which even the FP64 step fails halts with the
    FP64 check's code, for a single grid and inside an ensemble alike.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# Ensemble-workload instance (seed 2026, index 16) that halts in FP32 flux.
OTV_HARD = (1.067437115364036, 0.8064116573994691, 1.759240013263218, 0.20557563316160354,
            0.5272182398647961)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(mhd, n, precision, steps=-1, t_end=0.5, fallback=0, problem="ot", params=()):
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=precision, fallback=fallback))
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


def test_fallback_limit_trajectory_divergence(mhd):
    """An Orszag-Tang variant (ensemble instance 16 of seed 2026) at 1024^2:
    FP64 finishes; the mixed (and FP32) trajectory reaches a state at t ~ 0.42
    from which even the FP64 step is not physical, so the fallback fires and
    the FP64 check halts the run (code 3), as the FP64 form itself would from
    that state.  An ensemble holding the instance halts identically."""
    n = 1024
    say("hard OT variant 1024^2")
    u64, s64 = run(mhd, n, mhd.PREC_FP64, problem="otv", params=OTV_HARD)
    assert s64.code == 0 and s64.done == 1
    _, off = run(mhd, n, mhd.PREC_MIXED, fallback=1, problem="otv", params=OTV_HARD)
    um, sm = run(mhd, n, mhd.PREC_MIXED, fallback=0, problem="otv", params=OTV_HARD)
    say(f"  fallback off: code {off.code} step {off.fail_step}; on: code {sm.code} step "
        f"{sm.fail_step}, fallback tiles {sm.fallback_tiles}")
    assert off.code == 3 and sm.code == 3
    assert sm.fallback_tiles >= 1 and sm.fail_step == off.fail_step
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PREC_MIXED)
    ens = mhd.Ensemble(cfg, [(128, 128), (n, n)])
    ens.init_problem(0, "ot")
    ens.init_problem(1, "otv", OTV_HARD)
    sts = ens.advance(0.5)
    ue = ens.get_state(1)
    ens.close()
    assert sts[0].code == 0 and sts[0].done == 1
    assert sts[1].code == 3 and sts[1].fail_step == sm.fail_step
    assert sts[1].fallback_tiles == sm.fallback_tiles
    assert np.array_equal(ue.view(np.uint64), um.view(np.uint64))
