"""The no-allocation run (mhd_run_noalloc) against the context path
(mhd_create + mhd_set_state + mhd_advance + mhd_get_state) on the same
initial states, through the C ABI.

Two calls in a row, the second on a smaller grid (the scratch array keeps the
first call's U1, control block and fallback lists), must each equal the
context path bitwise; the FP64 runs also equal the CPU oracle.  The mixed
form's check mode (every tile re-evaluated by the FP64 kernel through the
fallback lists held in the scratch array) must equal the FP64 run.  A grid
beyond the scratch array is rejected before any launch.

The device state the run advances is a solver context's own state buffer
(mhd_device_state): device memory that a caller keeps on the device, with
planes padded to 64 elements, so `ld` exceeds nx * ny on these grids."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UINT = {np.dtype(np.float64): np.uint64, np.dtype(np.float32): np.uint32}
# Each call smaller than the one before; the last two are narrower than
# the 2-cell halo.
CASES = ((90, 70, 40), (50, 40, 25), (3, 2, 12), (1, 9, 12))


def say(msg):
    print(msg, flush=True)


def context_run(mhd, cfg, u0, steps):
    s = mhd.Solver(cfg)
    s.set_state(u0)
    st = s.advance(0.5, steps)
    u = s.get_state(0)
    s.close()
    return u, st


def noalloc_run(mhd, cfg, u0, steps):
    """Uploads u0 into a holder context's device state, runs the no-allocation
    entry point on that device memory, and downloads it again."""
    holder = mhd.Solver(cfg)
    holder.set_state(u0)
    ptr, ld = holder.device_state()
    assert ld > cfg.nx * cfg.ny
    st = mhd.run_noalloc(cfg, ptr, ld, 0.5, steps)
    u = holder.get_state(0)
    holder.close()
    return u, st


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(UINT[a.dtype]),
                          np.ascontiguousarray(b).view(UINT[b.dtype]))


def check_status(st, ref, steps):
    assert st.code == ref.code == 0
    assert st.steps == ref.steps == steps
    assert st.t == ref.t and st.dt == ref.dt and st.ch_last == ref.ch_last
    assert (st.degenerate, st.nonphysical) == (ref.degenerate, ref.nonphysical)
    assert st.fallback_tiles == ref.fallback_tiles


@pytest.mark.parametrize("form", ["fp64", "fp32", "mixed", "fp64_fma", "ref_mixed"])
def test_noalloc_equals_context_path(mhd, form):
    prec = mhd.PRECISIONS[form]
    for nx, ny, steps in CASES:
        say(f"noalloc {form} OT {nx}x{ny}, {steps} steps")
        cfg = mhd.SolverConfig(nx=nx, ny=ny, precision=prec)
        u0 = mhd.problem_state("ot", nx, ny, prec)
        ref, rst = context_run(mhd, cfg, u0, steps)
        got, st = noalloc_run(mhd, cfg, u0, steps)
        check_status(st, rst, steps)
        assert same_bits(got, ref), f"{form} {nx}x{ny}: differs from the context path"


def test_noalloc_fp64_equals_oracle(mhd, oracle):
    import pyoracle

    for nx, ny, steps in CASES:
        say(f"noalloc fp64 vs oracle OT {nx}x{ny}, {steps} steps")
        cfg = mhd.SolverConfig(nx=nx, ny=ny)
        u0a = oracle.init("ot", nx, ny)
        got, st = noalloc_run(mhd, cfg, pyoracle.aos_to_soa(u0a), steps)
        uo, info = oracle.advance(u0a, steps)
        assert st.code == 0 and st.steps == info["steps"] == steps
        assert same_bits(pyoracle.soa_to_aos(got), uo)
        assert (st.degenerate, st.nonphysical) == (int(info["counters"][2]),
                                                  int(info["counters"][3]))


def test_noalloc_mixed_check_mode_is_fp64(mhd):
    """fallback 2 queues every mixed tile on the scratch lists for the FP64
    kernel: the run is then the FP64 run, bitwise."""
    for nx, ny, steps in CASES:
        say(f"noalloc mixed check mode OT {nx}x{ny}, {steps} steps")
        u0 = mhd.problem_state("ot", nx, ny)
        ref, rst = context_run(mhd, mhd.SolverConfig(nx=nx, ny=ny), u0, steps)
        cfg = mhd.SolverConfig(nx=nx, ny=ny, precision=mhd.PREC_MIXED, fallback=2)
        got, st = noalloc_run(mhd, cfg, u0, steps)
        assert st.code == 0 and st.steps == steps and st.t == rst.t
        assert st.fallback_tiles > 0
        assert same_bits(got, ref)


def test_noalloc_rejects_oversize_grid(mhd):
    n = 1024  # FP64 U1 alone is 72 MiB > 24 MiB
    say(f"noalloc rejects {n}x{n}")
    cfg = mhd.SolverConfig(nx=n, ny=n)
    u0 = mhd.problem_state("ot", n, n)
    holder = mhd.Solver(cfg)
    holder.set_state(u0)
    ptr, ld = holder.device_state()
    with pytest.raises(mhd.MhdError) as ei:
        mhd.run_noalloc(cfg, ptr, ld, 0.5, 1)
    assert ei.value.code == 1
    kahan = mhd.SolverConfig(nx=n, ny=n, precision=mhd.PREC_MIXED_KAHAN)
    with pytest.raises(mhd.MhdError):
        mhd.run_noalloc(kahan, ptr, ld, 0.5, 1)
    assert same_bits(holder.get_state(0), u0)  # nothing ran
    holder.close()
