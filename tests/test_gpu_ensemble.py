"""GPU tests of ensembles (stage kernels over a per-stage live list of work
items: one block per item by default, or a persistent work-queue grid with
MHD_ENSEMBLE_KERNEL=queue).

Contract: an instance advanced inside an ensemble is bitwise the same
instance advanced alone by a single context (same form, same tile shape),
for uneven sizes, instances that finish at different step counts, and an
instance that halts (which must not disturb the others).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def single(mhd, cfg, nx, ny, problem, params, t_end, steps=-1, state=None):
    s = mhd.Solver(mhd.SolverConfig(**{**cfg.__dict__, "nx": nx, "ny": ny}))
    if state is not None:
        s.set_state(state)
    elif params:
        s.set_state(mhd.problem_state_ex(problem, params, nx, ny, cfg.precision))
    else:
        s.init_problem(problem)
    st = s.advance(t_end, steps)
    u = s.get_state(0)
    s.close()
    return u, st


def bits(u):
    return np.ascontiguousarray(u).view(np.uint8)


@pytest.mark.parametrize("form", ["fp64", "fp32", "mixed", "mixed_kahan"])
def test_ensemble_matches_single(mhd, form):
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PRECISIONS[form])
    # The last four are narrower than the 2-cell halo.
    sizes = [(64, 64), (48, 80), (96, 32), (40, 40), (33, 57), (1, 9), (7, 1), (1, 1), (3, 2)]
    params = [(1.0, 0, 0, 0, 0), (0.9, 0.3, 1.1, 2.0, 5.5), (1.2, 6.0, 0.1, 3.3, 0.7),
              (0.8, 1.0, 2.0, 3.0, 4.0), (1.05, 0.5, 0.5, 0.5, 0.5), (1.1, 0.2, 0.4, 0.6, 0.8),
              (0.95, 1.5, 2.5, 3.5, 4.5), (1.0, 0.7, 0.7, 0.7, 0.7), (0.85, 5.0, 4.0, 3.0, 2.0)]
    say(f"ensemble {form} {sizes}")
    ens = mhd.Ensemble(cfg, sizes)
    for i, p in enumerate(params):
        ens.init_problem(i, "otv", p)
    sts = ens.advance(0.05)
    for i, ((nx, ny), p) in enumerate(zip(sizes, params)):
        u_e = ens.get_state(i)
        u_s, st_s = single(mhd, cfg, nx, ny, "otv", p, 0.05)
        assert sts[i].code == 0 and sts[i].done == 1
        assert sts[i].steps == st_s.steps and sts[i].t == st_s.t
        assert (sts[i].degenerate, sts[i].nonphysical) == (st_s.degenerate, st_s.nonphysical)
        assert np.array_equal(bits(u_e), bits(u_s)), f"instance {i} {nx}x{ny}"
    # Different sizes take different step counts to t_end.
    assert len({s.steps for s in sts}) > 1
    ens.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_queue_kernel_matches_flat(mhd, form, monkeypatch):
    """The persistent work-queue stage kernels give the flat kernels' states:
    bitwise in FP64, and in the mixed form within rounding of the flat run
    (the two kernels are separate instantiations of the stage, whose FMA
    contraction may fall differently)."""
    cfg = mhd.SolverConfig(nx=2, ny=2, precision=mhd.PRECISIONS[form])
    sizes = [(64, 64), (48, 80), (96, 32), (33, 57)]
    params = [(1.0, 0, 0, 0, 0), (0.9, 0.3, 1.1, 2.0, 5.5), (1.2, 6.0, 0.1, 3.3, 0.7),
              (1.05, 0.5, 0.5, 0.5, 0.5)]
    out = {}
    for kernel in ("flat", "queue"):
        monkeypatch.setenv("MHD_ENSEMBLE_KERNEL", kernel)
        say(f"ensemble {form}, {kernel} kernel")
        ens = mhd.Ensemble(cfg, sizes)
        for i, p in enumerate(params):
            ens.init_problem(i, "otv", p)
        sts = ens.advance(0.03)
        assert all(st.code == 0 and st.done == 1 for st in sts)
        out[kernel] = ([st.steps for st in sts], [ens.get_state(i) for i in range(len(sizes))])
        ens.close()
    assert out["flat"][0] == out["queue"][0]
    for a, b in zip(out["flat"][1], out["queue"][1]):
        if form == "fp64":
            assert np.array_equal(bits(a), bits(b))
        else:
            assert np.max(np.abs(a - b) / (np.abs(a) + 1e-3)) < 1e-9


def test_many_instances_live_list(mhd):
    """300 instances (more than one 256-instance round of the live-list
    kernel and more than 64 live slots per slot search), finishing at
    different step counts: each equals its single run, bitwise."""
    cfg = mhd.SolverConfig(nx=2, ny=2)
    shapes = [(16, 16), (20, 12), (33, 17), (8, 40)]
    sizes = [shapes[i % len(shapes)] for i in range(300)]
    ens = mhd.Ensemble(cfg, sizes)
    for i in range(len(sizes)):
        ens.init_problem(i, "otv", (0.8 + 0.001 * i, 0.01 * i, 0.02 * i, 0.03 * i, 0.04 * i))
    say("ensemble of 300 small instances, FP64, to t=0.02")
    sts = ens.advance(0.02)
    assert all(st.code == 0 and st.done == 1 for st in sts)
    assert len({st.steps for st in sts}) > 1
    for i in range(0, 300, 37):
        nx, ny = sizes[i]
        p = (0.8 + 0.001 * i, 0.01 * i, 0.02 * i, 0.03 * i, 0.04 * i)
        u_s, st_s = single(mhd, cfg, nx, ny, "otv", p, 0.02)
        assert sts[i].steps == st_s.steps
        assert np.array_equal(bits(ens.get_state(i)), bits(u_s)), i
    ens.close()


def test_ensemble_halted_instance_isolated(mhd, oracle):
    cfg = mhd.SolverConfig(nx=2, ny=2)
    sizes = [(32, 32), (40, 24), (32, 32)]
    ens = mhd.Ensemble(cfg, sizes)
    ens.init_problem(0, "ot")
    bad = mhd.problem_state("ot", 40, 24)
    bad[4, 10, 7] = 1e-8     # E far below kinetic + magnetic energy
    ens.set_state(1, bad)
    ens.init_problem(2, "ot")
    sts = ens.advance(0.02)
    assert sts[1].code == 3 and (sts[1].bad_i, sts[1].bad_j) == (7, 10) and sts[1].steps == 0
    for i in (0, 2):
        assert sts[i].code == 0 and sts[i].done == 1
        u_s, st_s = single(mhd, cfg, 32, 32, "ot", None, 0.02)
        assert np.array_equal(bits(ens.get_state(i)), bits(u_s))
    ens.close()


def test_ensemble_steps_limit_and_graph(mhd):
    """max_steps stops every instance at the same step count; batches replayed
    as graphs give the same states as single launches."""
    cfg = mhd.SolverConfig(nx=2, ny=2, batch_steps=8)
    sizes = [(64, 48), (48, 64)]
    ens = mhd.Ensemble(cfg, sizes)
    for i in range(2):
        ens.init_problem(i, "ot")
    sts = ens.advance(0.5, 20)
    assert [s.steps for s in sts] == [20, 20]
    for i, (nx, ny) in enumerate(sizes):
        u_s, _ = single(mhd, mhd.SolverConfig(nx=2, ny=2, batch_steps=1), nx, ny, "ot", None, 0.5,
                        steps=20)
        assert np.array_equal(bits(ens.get_state(i)), bits(u_s))
    ens.close()


# ---------------------------------------------------------------------------
# The batched-ensemble workload itself: 64 seeded Orszag-Tang variants of
# sizes 128^2-1024^2.
# ---------------------------------------------------------------------------
def workload_ensemble(mhd, form, n=64):
    from mhdglm.workloads import ensemble_workload

    sizes, params = ensemble_workload(n)
    ens = mhd.Ensemble(mhd.SolverConfig(nx=2, ny=2, precision=mhd.PRECISIONS[form]),
                       [(s, s) for s in sizes])
    for i, p in enumerate(params):
        ens.init_problem(i, "otv", p)
    return ens, sizes, params


def test_workload_fp64_instances_match_oracle(mhd, oracle):
    """FP64 ensemble of the 64-instance workload, 12 steps: every 128^2
    instance and the first 192^2 one equal the CPU oracle run on the oracle's
    own initial state, bitwise."""
    import pyoracle

    steps = 12
    say(f"64-instance workload, FP64, {steps} steps")
    ens, sizes, params = workload_ensemble(mhd, "fp64")
    sts = ens.advance(0.5, steps)
    picks = [i for i, s in enumerate(sizes) if s == 128] + [sizes.index(192)]
    got = {i: ens.get_state(i) for i in picks}
    ens.close()
    assert all(st.code == 0 and st.steps == steps for st in sts)
    for i in picks:
        n = sizes[i]
        u0 = oracle.init("otv", n, n, params=params[i])
        uo, info = oracle.advance(u0, steps)
        say(f"  instance {i} ({n}^2) vs oracle")
        assert info["steps"] == steps
        assert np.array_equal(pyoracle.soa_to_aos(got[i]).view(np.uint64), uo.view(np.uint64)), i
        assert (sts[i].degenerate, sts[i].nonphysical) == (int(info["counters"][2]),
                                                          int(info["counters"][3]))


def test_workload_mixed_finishes_every_instance(mhd):
    """The mixed ensemble of the 64-instance workload to t = 0.5: all 64
    instances finish (any whose trajectory drifts into a state the FP64
    tile fallback cannot advance through an FP64 rollback window, as
    instance 16 once did), none by a restart, and instance 16 ends at the
    mixed form's distance from its FP64 run."""
    from mhdglm import compare

    say("64-instance workload, mixed, to t=0.5")
    ens, sizes, params = workload_ensemble(mhd, "mixed")
    sts = ens.advance(0.5)
    rolled = [i for i, st in enumerate(sts) if st.rollbacks]
    say(f"  done {sum(st.done for st in sts)}/64, rolled back {rolled}, "
        f"restarts {sum(st.restarts for st in sts)}")
    assert all(st.code == 0 and st.done == 1 for st in sts)
    assert all(st.restarts == 0 for st in sts)
    u16 = ens.get_state(16)
    ens.close()
    n = sizes[16]
    u64, s64 = single(mhd, mhd.SolverConfig(nx=n, ny=n), n, n, "otv", params[16], 0.5)
    errs = compare.rel_l2_state(u64, u16, ("rho", "p", "|B|"))
    say(f"  instance 16 relL2 vs FP64 {errs}")
    assert s64.code == 0 and max(errs.values()) < 1e-3
