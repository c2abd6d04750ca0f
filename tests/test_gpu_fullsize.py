"""GPU tests of the chosen form at the sizes the headline is quoted on:
Orszag-Tang 4096^2 in the mixed form with the learned tile selector (HIP
graphs of 32 steps), 8192^2 (4.8 GB per FP64 state) in FP64, mixed and
mixed with the selector, and 16384^2 (19 GB per FP64 state; the largest
power-of-two square below the 2^31-cell index limit) in FP64 and mixed.
Size-independent properties only (the CPU oracle cannot run these sizes):
  * determinism: two runs are bitwise identical;
  * conservation: the periodic flux form keeps the sums of rho, momentum,
    energy and B to rounding;
  * routing: with every tile flagged the selected run is the FP64 run, bit
    for bit, at 4096^2;
  * audits: the sample is 1% of the mixed tile-steps.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def conserved(u0, u, rel=1e-12):
    for k in (0, 1, 2, 3, 4, 5, 6, 7):
        tot0, tot1 = u0[k].sum(dtype=np.float64), u[k].sum(dtype=np.float64)
        scale = np.abs(u0[k]).sum(dtype=np.float64) + np.abs(u[k]).sum(dtype=np.float64)
        assert abs(tot1 - tot0) <= rel * scale, (k, tot0, tot1)


def selected_run(mhd, n, steps, **sel):
    from mhdglm import selector

    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=mhd.PREC_MIXED))
    s.init_problem("ot")
    u0 = s.get_state(0)
    s.configure_selector(selector.load(selector.DEFAULT_MODEL), **sel)
    st = s.advance(0.5, steps)
    u = s.get_state(0)
    stats = s.selector_stats()
    tiles = s.selector_tiles()
    s.close()
    return u0, u, st, stats, tiles


def test_selector_4096_deterministic_conservative_audited(mhd):
    n, steps = 4096, 40
    say(f"mixed + selector OT {n}^2, {steps} steps, twice")
    u0, a, sa, stats_a, tiles = selected_run(mhd, n, steps)
    _, b, sb, stats_b, _ = selected_run(mhd, n, steps)
    say(f"  status {sa.code} steps {sa.steps}; selector {stats_a}")
    assert sa.code == sb.code == 0 and sa.steps == sb.steps == steps
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert stats_a == stats_b and stats_a["steps"] == steps
    conserved(u0, a)
    mixed_tile_steps = steps * tiles - stats_a["fp64_tile_steps"]
    assert abs(stats_a["audits"] - 0.01 * mixed_tile_steps) <= 2 * steps


def test_selector_4096_all_flagged_is_fp64(mhd):
    n, steps = 4096, 6
    say(f"selector threshold -1 at {n}^2 vs FP64, {steps} steps")
    _, us, st, stats, tiles = selected_run(mhd, n, steps, threshold=-1.0, audit_frac=0.0)
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n))
    s.init_problem("ot")
    s.advance(0.5, steps)
    u64 = s.get_state(0)
    s.close()
    assert st.code == 0 and stats["fp64_tile_steps"] == steps * tiles and stats["audits"] == 0
    assert np.array_equal(us.view(np.uint64), u64.view(np.uint64))


def test_8192_fp64_and_mixed(mhd):
    from mhdglm import compare

    n, steps = 8192, 4
    out = {}
    for name, prec in (("fp64", mhd.PREC_FP64), ("mixed", mhd.PREC_MIXED)):
        say(f"OT {n}^2 {name}, {steps} steps")
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec))
        s.init_problem("ot")
        if name == "fp64":
            u0 = s.get_state(0)
        st = s.advance(0.5, steps)
        assert st.code == 0 and st.steps == steps
        out[name] = s.get_state(0)
        s.close()
        conserved(u0, out[name])
    errs = compare.rel_l2_state(out["fp64"], out["mixed"], ("rho", "p", "|B|"))
    say(f"  mixed vs FP64 after {steps} steps: {errs}")
    assert max(errs.values()) < 1e-6
    # The chosen form at this size: mixed + learned selector (graph-batched).
    say(f"OT {n}^2 mixed + selector, {steps} steps")
    u0s, us, st, stats, tiles = selected_run(mhd, n, steps)
    assert st.code == 0 and st.steps == steps and stats["steps"] == steps
    assert np.array_equal(u0s.view(np.uint64), u0.view(np.uint64))
    conserved(u0, us)
    mixed_tile_steps = steps * tiles - stats["fp64_tile_steps"]
    assert abs(stats["audits"] - 0.01 * mixed_tile_steps) <= 2 * steps
    errs = compare.rel_l2_state(out["fp64"], us, ("rho", "p", "|B|"))
    say(f"  selector vs FP64 after {steps} steps: {errs}; selector {stats}")
    assert max(errs.values()) < 1e-6


def test_16384_fp64_and_mixed(mhd):
    """268M cells: about 39 GB of HBM per context (state and U1), one
    context at a time; the cell index, plane offsets and the first-bad-cell
    index stay within their types."""
    from mhdglm import compare

    n, steps = 16384, 2
    out = {}
    for name, prec in (("fp64", mhd.PREC_FP64), ("mixed", mhd.PREC_MIXED)):
        say(f"OT {n}^2 {name}, {steps} steps")
        s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=prec))
        s.init_problem("ot")
        if name == "fp64":
            u0 = s.get_state(0)
        st = s.advance(0.5, steps)
        assert st.code == 0 and st.steps == steps
        out[name] = s.get_state(0)
        s.close()
        conserved(u0, out[name])
    errs = compare.rel_l2_state(out["fp64"], out["mixed"], ("rho", "p", "|B|"))
    say(f"  mixed vs FP64 after {steps} steps: {errs}")
    assert max(errs.values()) < 1e-6
