"""Throughput benchmark of the MI355X 2D MHD-GLM step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096] [--precision fp64]
                    [--tile T] [--fuse 1|2]

Workload: Orszag-Tang on an n x n periodic grid (default 4096^2, the size the
headline metric is quoted on), MUSCL + HLLD + mixed GLM, SSPRK2, CFL 0.25,
alpha 0.18, HLLD tolerance 1e-4.  One "step" is one full SSPRK2 step of the
whole grid (two fused stage kernels and one control kernel, the autotuned
choice; --fuse 2 runs both stages in one kernel); the state stays resident in
HBM.  value = n*n*K / (wall time of K steps) in cell-updates/s,
the reference's "Mcell-steps/s" unit (mhd2d_glm_cpu.cpp:589-591) times 1e6.

`value` is the plain FP64 path (bitwise the reference CPU build).  Also
reported on the same JSON line:
  roofline     -- the step against the HBM roofline as the metric defines
                  it: 144 B per FP64 cell-update (read U^n, write U^{n+1})
                  times the cell-updates one step performs, over the step's
                  kernel time measured here with HIP events on the solver
                  stream (stage 1 + stage 2 + finish), against 8 TB/s; the
                  same figure per stage kernel (its own compulsory bytes);
                  the rocprofv3-measured HBM traffic per step (committed
                  counter runs) and its ratio to the 144 B figure;
  gate         -- the headline's second half on its own configuration:
                  whole Orszag-Tang 4096^2 runs to t=0.5 in FP64, in the
                  mixed form with the learned tile selector, and in FP64 with
                  FMA contraction (the noise floor: two valid FP64 builds):
                  wall time, cell-updates/s, speedup over FP64, relative L2
                  against FP64 of rho, p, |B| and psi, and gate_met (rho, p
                  and |B| all <= 1e-5); it runs last, and leaves out the
                  floor run (saying so) when that would end past
                  --time-budget seconds (default 430);
  forms        -- throughput of the other arithmetic forms on the timed
                  workload (FP64 with FMA, mixed, mixed with the learned
                  tile selector, mixed + compensated, FP32, and the
                  reference's stage-level mixed form), the mixed-vs-FP64
                  speedup and the selector's FP64 tile share;
  ensemble     -- the batched workload: 64 Orszag-Tang instances (seeded
                  sizes in {128..1024}^2, B amplitude and sine phases varied)
                  advanced together to t=0.5 by the ensemble stage kernels
                  (one launch per stage over the device list of the running
                  instances' tiles), each with its own dt; aggregate cell-updates/s,
                  instances finished and FP64 rollbacks;
  large        -- FP64, mixed and mixed + selector steps at 8192^2 (4.8 GB
                  per FP64 state);
  selector_gain -- what the learned selector's routing buys: from one FP64
                  state at 512^2 (Orszag-Tang t=0.15, Kelvin-Helmholtz
                  t=2), k = 1..64 steps in FP64, in the mixed form, and in
                  the mixed form routed by the shipped network and by a
                  random network at the same FP64 tile share; relL2 against
                  FP64 and the share of the mixed form's error removed
                  (--selector-gain, 0 to skip);
  slab_share   -- one GPU's share of an 8-way row-slab split of the grid at
                  its real shape (n x 512 owned rows + 2-row halos): stage
                  kernels against the full grid (--slab-rows, 0 to skip);
  cpu_baseline -- the reference CPU build (oracle/_ref, serial, one core) on
                  a bounded sample of Orszag-Tang at 512^2 (the first ~50
                  steps); `parallel`: the aggregate rate of min(visible
                  CPUs, 16) such runs at once.
Optional sections (off by default; their results are kept in profiles/):
  --accuracy-n N        whole runs of every form at N^2 against FP64;
  --kh-n N              Kelvin-Helmholtz (the 1024^2 variant) to t=5;
  --selector-n N        Orszag-Tang N^2 whole runs, FP64 vs selector;
  --selector-breakdown N  the selector's cost per step, by part.
Profiling mode: --whole-run FORM runs only one whole Orszag-Tang --n^2 run
to t=0.5 in FORM ("selector" included) and prints its wall time and rate.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "precision-study_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "cell-updates/s, 2D MHD-GLM HLLD 4096²; mixed vs FP64 speedup at L2≤1e-5"
# Algorithmic HBM bytes per cell-update (9 variables): read U^n, write
# U^{n+1}; every intermediate stays on chip in an ideally fused step.
STEP_BYTES_PER_CELL = 2 * 9
# Compulsory HBM bytes per cell of each stage kernel of the two-launch step:
#   stage 1 reads U, writes U1;  stage 2 reads U1 and U, writes U.
STAGE_BYTES_PER_CELL = {1: 2 * 9, 2: 3 * 9}
GATE_L2 = 1e-5
T_START = time.perf_counter()


def elapsed() -> float:
    """Seconds since this process started timing anything."""
    return time.perf_counter() - T_START


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n: int, t_end: float, parallel: int = 0):
    """Reference CPU build (or the oracle if the reference build is absent),
    OT n^2 FP64 to t_end (a bounded sample of the whole run), one thread;
    then, with parallel > 1, that many independent single-thread runs at
    once (the host's aggregate throughput on the same build)."""
    ref = os.path.join(ROOT, "oracle", "_ref", "mhd2d_glm_cpu_fp64")
    kind = "reference"
    if not os.path.exists(ref):
        ref = os.path.join(ROOT, "oracle", "build", "mhd2d_glm_oracle_fp64")
        kind = "port"
    if not os.path.exists(ref):
        return None
    argv = [ref, str(n), str(n), str(t_end), "0.25", "muscl", "hlld", "0.18"]
    pat = re.compile(r"Time loop wall time = (\S+) s \((\S+) Mcell-steps/s\)")
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        r = subprocess.run(argv, env=dict(os.environ, OUTDIR=d), capture_output=True, text=True,
                           timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        return None
    m = pat.search(r.stdout)
    steps = int(re.search(r"Done\. Steps: (\d+)", r.stdout).group(1))
    out = {
        "value": float(m.group(2)) * 1e6,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": kind,
        "sample": f"Orszag-Tang {n}x{n} FP64 MUSCL+HLLD to t={t_end} ({steps} SSPRK2 steps, "
                  f"{float(m.group(1)):.1f} s loop, {wall:.1f} s process), serial g++ -O2 build",
        "threading": "1 thread per run",
        "host_cpu": host_cpu_model(),
        "host_cpus_visible": os.cpu_count(),
    }
    if parallel > 1:
        # k copies of the same run at once, each in its own output folder:
        # the aggregate of their loop rates (n^2 * steps / loop wall each).
        with tempfile.TemporaryDirectory() as d:
            procs = []
            for k in range(parallel):
                od = os.path.join(d, str(k))
                os.makedirs(od)
                procs.append(subprocess.Popen(argv, env=dict(os.environ, OUTDIR=od),
                                              stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                                              text=True))
            rates = []
            for p in procs:
                so, _ = p.communicate(timeout=900)
                mm = pat.search(so or "")
                if p.returncode == 0 and mm:
                    rates.append(float(mm.group(2)) * 1e6)
        if len(rates) == parallel:
            out["parallel"] = {"runs": parallel, "value": round(sum(rates), 1),
                               "per_run_min_max": [round(min(rates), 1), round(max(rates), 1)],
                               "threading": f"{parallel} independent single-thread runs at once"}
    return out


def host_cpu_model() -> str:
    """The host CPU's model name (Linux /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_solver(mhdglm, form, n, tile=-1, problem="ot", fuse=0):
    """n^2 solver of `problem` in an arithmetic form; "selector" is the mixed
    form with the committed selector model deciding FP64 tiles (which runs
    one stage per launch whatever `fuse` says), "selector@X" the same with
    the network's threshold X."""
    sel = form.startswith("selector")
    prec = mhdglm.PREC_MIXED if sel else mhdglm.PRECISIONS[form]
    solver = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n, precision=prec, tile=tile,
                                               fuse=fuse))
    solver.init_problem(problem)
    if sel:
        from mhdglm import selector
        thr = float(form.split("@", 1)[1]) if "@" in form else None
        solver.configure_selector(selector.load(selector.DEFAULT_MODEL), threshold=thr)
    return solver


def time_form(mhdglm, torch, form, n, steps, warmup, tile=-1, fuse=0):
    """Times `steps` SSPRK2 steps of OT n^2 in one arithmetic form."""
    log(f"bench: OT {n}x{n} {form}, warmup {warmup}, steps {steps}")
    solver = make_solver(mhdglm, form, n, tile, fuse=fuse)
    t_end = 0.5
    solver.enqueue_steps(t_end, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.enqueue_steps(t_end, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = solver.status()
    if st.code != 0 or st.steps != warmup + steps:
        print(f"bench: {form} run did not complete the timed steps: {st.as_dict()}",
              file=sys.stderr)
        sys.exit(1)
    wall = t1 - t0
    value = n * n * steps / wall
    extra = {}
    if form == "selector":
        sel = solver.selector_stats()
        tiles = solver.selector_tiles()
        extra = {"fp64_tile_fraction": sel["fp64_tile_steps"] / max(sel["steps"] * tiles, 1),
                 "audits": sel["audits"], "audit_misses": sel["misses"]}
        ms = None
    else:
        ms = solver.time_kernels(t_end, 20)
    solver.close()
    log(f"bench: {form}: {steps} steps in {wall:.3f} s -> {value:.4e} cell-updates/s {extra}")
    return {"value": value, "wall": wall, "kernel_ms": ms, **extra}


def whole_run(mhdglm, form, n, problem="ot", t_end=0.5):
    """One whole run from t=0; returns (state, status, wall seconds, selector stats)."""
    t0 = time.perf_counter()
    s = make_solver(mhdglm, form, n, problem=problem)
    while True:  # chunks, so that long runs report progress
        st = s.advance(t_end, 4000)
        if st.code != 0 or st.done:
            break
        log(f"bench: {problem} {n}^2 {form}: step {st.steps}, t={st.t:.4f}")
    wall = time.perf_counter() - t0
    u = s.get_state(0).astype("float64")
    sel = None
    if form.startswith("selector"):
        sel = s.selector_stats()
        sel["fp64_tile_fraction"] = round(sel["fp64_tile_steps"] / max(sel["steps"] * s.selector_tiles(), 1), 5)
    s.close()
    log(f"bench: {problem} {n}^2 {form}: {st.steps} steps, code {st.code}, {wall:.1f} s")
    return u, st, wall, sel


def accuracy_check(mhdglm, n, forms, problem="ot", t_end=0.5):
    """Whole runs of `problem` at n^2 to t_end in FP64 and each form:
    relative L2 against FP64 of rho, p, |B|, psi, steps, and the rate."""
    from mhdglm import compare

    u64, s64, w64, _ = whole_run(mhdglm, "fp64", n, problem, t_end)
    out = {"problem": problem, "n": n, "t_end": t_end, "fp64_steps": s64.steps,
           "fp64_cell_updates_per_s": round(n * n * s64.steps / w64, 1), "relL2_vs_fp64": {}}
    for form in forms:
        u, st, wall, sel = whole_run(mhdglm, form, n, problem, t_end)
        errs = compare.rel_l2_state(u64, u, ("rho", "p", "|B|", "psi"))
        row = {k: float(f"{v:.3e}") for k, v in errs.items()}
        row["steps"] = st.steps
        row["code"] = st.code
        row["cell_updates_per_s"] = round(n * n * st.steps / wall, 1)
        row["speedup_vs_fp64"] = round(row["cell_updates_per_s"] / out["fp64_cell_updates_per_s"], 3)
        if sel:
            row["selector"] = sel
        out["relL2_vs_fp64"][form] = row
    return out


def gate_check(mhdglm, n, with_floor=True, t_end=0.5, budget_s=None):
    """The metric's accuracy-gated speedup on its own configuration: whole
    Orszag-Tang n^2 runs to t_end in FP64, in the mixed form with the learned
    tile selector and (with_floor) in FP64 with FMA contraction, whose
    distance from FP64 is the noise floor of this chaotic flow.  The floor
    run takes about as long as the FP64 run; it is left out (and says so)
    when it would end past `budget_s` seconds of this process."""
    from mhdglm import compare

    fields = ("rho", "p", "|B|", "psi")
    u64, s64, w64, _ = whole_run(mhdglm, "fp64", n)
    out = {"problem": "ot", "n": n, "t_end": t_end, "gate_relL2": GATE_L2,
           "fp64": {"steps": s64.steps, "code": s64.code, "wall_s": round(w64, 2),
                    "cell_updates_per_s": round(n * n * s64.steps / w64, 1)}}
    forms = ["selector"] + (["fp64_fma"] if with_floor else [])
    for form in forms:
        if form == "fp64_fma" and budget_s and elapsed() + 1.1 * w64 > budget_s:
            out["fp64_fma"] = {
                "skipped": f"time budget: {elapsed():.0f} s used, the run takes about "
                           f"{w64:.0f} s, budget {budget_s:.0f} s",
                "measured_in": "profiles/accuracy_forms.md (Orszag-Tang 4096^2 table)"}
            with_floor = False
            continue
        u, st, wall, sel = whole_run(mhdglm, form, n)
        errs = compare.rel_l2_state(u64, u, fields)
        row = {"steps": st.steps, "code": st.code, "done": st.done, "wall_s": round(wall, 2),
               "cell_updates_per_s": round(n * n * st.steps / wall, 1),
               "speedup_vs_fp64": round(w64 / wall * st.steps / s64.steps, 3),
               "relL2_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
               "rollbacks": st.rollbacks}
        if sel:
            row["selector"] = sel
        out[form] = row
    sel = out["selector"]
    out["speedup"] = sel["speedup_vs_fp64"]
    out["gate_met"] = bool(sel["code"] == 0 and
                           all(sel["relL2_vs_fp64"][k] <= GATE_L2 for k in ("rho", "p", "|B|")))
    if with_floor and "relL2_vs_fp64" in out.get("fp64_fma", {}):
        floor = out["fp64_fma"]["relL2_vs_fp64"]
        out["selector_over_floor"] = {k: round(sel["relL2_vs_fp64"][k] / floor[k], 3)
                                      for k in ("rho", "p", "|B|") if floor[k] > 0}
        out["floor_below_gate"] = all(floor[k] <= GATE_L2 for k in ("rho", "p", "|B|"))
    return out


def selector_breakdown(mhdglm, n, t_end=0.5):
    """Whole Orszag-Tang runs at n^2 to t_end in the mixed form, and in the
    mixed form with the selector as shipped, with audits off, with nothing
    flagged (threshold 2), and with neither: what a step costs over the
    mixed step for the selector's own launches, the audits and the FP64
    tiles."""
    from mhdglm import selector

    model = selector.load(selector.DEFAULT_MODEL)
    variants = [("mixed", None), ("selector", {}), ("selector_no_audits", {"audit_frac": 0.0}),
                ("selector_nothing_flagged", {"threshold": 2.0}),
                ("selector_launches_only", {"threshold": 2.0, "audit_frac": 0.0})]
    runs = {}
    for name, kw in variants:
        s = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n, precision=mhdglm.PREC_MIXED))
        s.init_problem("ot")
        if kw is not None:
            s.configure_selector(model, **kw)
        t0 = time.perf_counter()
        st = s.advance(t_end)
        wall = time.perf_counter() - t0
        row = {"steps": st.steps, "code": st.code, "wall_s": round(wall, 3),
               "us_per_step": round(wall / max(st.steps, 1) * 1e6, 2)}
        if kw is not None:
            sel = s.selector_stats()
            row["fp64_tile_fraction"] = round(sel["fp64_tile_steps"] / max(sel["steps"] * s.selector_tiles(), 1), 5)
            row["audits"], row["misses"] = sel["audits"], sel["misses"]
        s.close()
        runs[name] = row
        log(f"bench: selector breakdown {n}^2 {name}: {row}")
    us = {k: v["us_per_step"] for k, v in runs.items()}
    return {"n": n, "t_end": t_end, "runs": runs, "us_per_step_over_mixed": {
        "selector_launches": round(us["selector_launches_only"] - us["mixed"], 2),
        "audits": round(us["selector_nothing_flagged"] - us["selector_launches_only"], 2),
        "fp64_tiles": round(us["selector_no_audits"] - us["selector_launches_only"], 2),
        "all": round(us["selector"] - us["mixed"], 2)}}


def selector_gain(mhdglm, n, t0=0.15, problem="ot", ks=(1, 4, 16, 64), shares=(0.05, 0.2)):
    """What the network's routing buys before a chaotic flow amplifies
    every rounding difference alike: from one FP64 state of `problem` at
    n^2 and t0, k steps in FP64 (the reference), in the mixed form, and in
    the mixed form routed by the shipped network and by a random network
    at the same FP64 tile share (audits off); relL2 against FP64 of rho, p
    and |B| after each k.  Shares: the network's own threshold, then the
    thresholds that flag `shares` of the tiles."""
    import numpy as np

    from mhdglm import compare, selector

    s = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n))
    s.init_problem(problem)
    s.advance(t0)
    u0 = s.get_state(0)
    s.close()
    c = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n, precision=mhdglm.PREC_MIXED))
    c.set_state(u0)
    feats = c.tile_features()
    c.close()
    shipped = selector.load(selector.DEFAULT_MODEL)
    rnd = {"params": selector.init_params(7), "mean": np.zeros(12, np.float32),
           "std": np.ones(12, np.float32), "threshold": np.float32(0.5),
           "budget": np.float32(1e30), "audit_budget": np.float32(1e30)}

    def prob(model):
        x = np.zeros_like(feats)
        x[:, :12] = (feats[:, :12] - model["mean"]) / model["std"]
        return mhdglm.mlp_forward(model["params"], x)

    ps, pr = prob(shipped), prob(rnd)
    own = float(np.mean(ps > float(shipped["threshold"])))
    routes = [("network", float(shipped["threshold"]), own)]
    routes += [("network", float(np.quantile(ps, 1.0 - f)), f) for f in shares]

    def run(k, model=None, thr=None):
        prec = mhdglm.PREC_FP64 if model == "fp64" else mhdglm.PREC_MIXED
        c = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n, precision=prec))
        c.set_state(u0)
        if isinstance(model, dict):
            c.configure_selector(model, threshold=thr, audit_frac=0.0)
        st = c.advance(10.0, k)
        u = c.get_state(0)
        share = None
        if isinstance(model, dict):
            sel = c.selector_stats()
            share = round(sel["fp64_tile_steps"] / max(sel["steps"] * c.selector_tiles(), 1), 4)
        c.close()
        return u, st, share

    out = {"problem": problem, "n": n, "t0": t0, "network_own_share": round(own, 4),
           "rows": []}
    for k in ks:
        ref, _, _ = run(k, "fp64")
        um, _, _ = run(k)
        em = compare.rel_l2_state(ref, um, ("rho", "p", "|B|"))
        row = {"k": k, "mixed": {f: float(f"{v:.3e}") for f, v in em.items()}}
        for name, thr, share in routes:
            thr_r = float(np.quantile(pr, 1.0 - share)) if share > 0 else 2.0
            un, _, sn = run(k, shipped, thr)
            ur, _, sr = run(k, rnd, thr_r)
            en = compare.rel_l2_state(ref, un, ("rho", "p", "|B|"))
            er = compare.rel_l2_state(ref, ur, ("rho", "p", "|B|"))
            row[f"share_{share:.3f}"] = {
                "network": {"fp64_share": sn, **{f: float(f"{v:.3e}") for f, v in en.items()}},
                "random": {"fp64_share": sr, **{f: float(f"{v:.3e}") for f, v in er.items()}},
                "rho_error_removed": {"network": round(1.0 - en["rho"] / em["rho"], 3),
                                      "random": round(1.0 - er["rho"] / em["rho"], 3)}}
        out["rows"].append(row)
        log(f"bench: selector gain {problem} {n}^2 k={k}: {row}")
    return out


def slab_share(mhdglm, n, rows, forms=("fp64", "mixed"), steps=40):
    """One GPU's share of a row-slab split of the n^2 grid, at its real
    shape: the stage kernels on an n x (rows + 4) grid holding rows
    0 .. rows + 3 of the Orszag-Tang state after 200 steps (`rows` owned
    rows and a 2-row halo on each side, as a slab would hold them after its
    halo exchange), against the full grid.  Rates count owned cells only.
    The kernels are the single-grid ones (periodic in y), so the halo rows'
    own updates are discarded work, as they would be on a slab."""
    import numpy as np

    full = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=n))
    full.init_problem("ot")
    full.advance(0.5, 200)
    u = full.get_state(0)
    full.close()
    out = {"n": n, "rows": rows, "state": "Orszag-Tang after 200 FP64 steps", "forms": {}}
    for form in forms:
        r = {}
        for ny, owned in ((n, n * n), (rows + 4, n * rows)):
            s = mhdglm.Solver(mhdglm.SolverConfig(nx=n, ny=ny, precision=mhdglm.PRECISIONS[form]))
            s.set_state(np.ascontiguousarray(u[:, :ny, :]))
            s.advance(0.5, 5)
            ms = s.time_kernels(0.5, steps)
            s.close()
            key = "full" if ny == n else "slab"
            r[key] = {"shape": [n, ny], "stage_ms": [round(x, 4) for x in ms],
                      "step_ms": round(sum(ms), 4),
                      "owned_cell_updates_per_s": round(owned / (sum(ms) * 1e-3))}
        r["slab_vs_full"] = round(r["slab"]["owned_cell_updates_per_s"]
                                  / r["full"]["owned_cell_updates_per_s"], 3)
        out["forms"][form] = r
        log(f"bench: slab share {n}x{rows} {form}: {r}")
    return out


def run_ensemble(mhdglm, torch, n, form, t_end):
    from mhdglm.workloads import ENSEMBLE_SEED, ensemble_workload

    sizes, params = ensemble_workload(n)
    log(f"bench: ensemble of {n} OT instances ({form}), sizes {sorted(set(sizes))}, to t={t_end}")
    ens = mhdglm.Ensemble(mhdglm.SolverConfig(nx=2, ny=2, precision=mhdglm.PRECISIONS[form]),
                          [(s, s) for s in sizes])
    for i, p in enumerate(params):
        ens.init_problem(i, "otv", p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sts = ens.advance(t_end)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ens.close()
    codes = [s.code for s in sts]
    updates = sum(s * s * st.steps for s, st in zip(sizes, sts))
    return {
        "instances": n, "form": form, "t_end": t_end, "seed": ENSEMBLE_SEED,
        "sizes": {str(k): sizes.count(k) for k in sorted(set(sizes))},
        "wall_s": round(wall, 3), "cell_updates": updates,
        "value": round(updates / wall, 1), "unit": "cell-updates/s",
        "all_done": all(st.done == 1 for st in sts), "codes": sorted(set(codes)),
        "instances_done": sum(1 for st in sts if st.done == 1),
        "rollbacks": [{"instance": i, "size": sizes[i], "windows": st.rollbacks,
                       "fp64_steps": st.exact_steps, "restarted": st.restarts}
                      for i, st in enumerate(sts) if st.rollbacks or st.restarts],
        "instances_halted": [{"instance": i, "size": sizes[i], "code": st.code, "t": round(st.t, 4)}
                             for i, st in enumerate(sts) if st.code != 0],
        "steps_min_max": [min(st.steps for st in sts), max(st.steps for st in sts)],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--precision",
                    choices=("fp64", "fp32", "mixed", "mixed_kahan", "fp64_fma", "ref_mixed"),
                    default="fp64")
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--fuse", type=int, default=0,
                    help="SSPRK2 stages per launch: 1, 2 (16x15 tiles), 0 = the autotuned choice")
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--cpu-t-end", type=float, default=0.006)
    ap.add_argument("--cpu-parallel", type=int, default=-1,
                    help="independent single-thread CPU runs at once for the aggregate figure "
                         "(-1: min(visible CPUs, 16); 0 or 1: none)")
    ap.add_argument("--large-n", type=int, default=8192,
                    help="also time FP64 and mixed steps at this size (0 to skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64_fma,mixed,selector,mixed_kahan,fp32,ref_mixed",
                    help="other forms timed beside --precision ('' for none)")
    ap.add_argument("--gate-n", type=int, default=4096,
                    help="grid of the headline's whole runs to t=0.5, FP64 vs selector (0 to skip)")
    ap.add_argument("--no-gate-floor", action="store_true",
                    help="skip the FP64-FMA noise-floor run of the gate")
    ap.add_argument("--accuracy-n", type=int, default=0,
                    help="grid of the whole-run accuracy check of every form (0 to skip)")
    ap.add_argument("--kh-n", type=int, default=0,
                    help="Kelvin-Helmholtz (khb) whole runs, FP64 vs mixed (0 to skip)")
    ap.add_argument("--kh-t-end", type=float, default=5.0)
    ap.add_argument("--kh-forms", default="fp64_fma,mixed",
                    help="forms of the Kelvin-Helmholtz whole runs, against FP64")
    ap.add_argument("--selector-breakdown", type=int, default=0,
                    help="whole runs at this size splitting the selector's cost (0 to skip)")
    ap.add_argument("--selector-n", type=int, default=0,
                    help="Orszag-Tang whole runs to t=0.5, FP64 vs selector (0 to skip)")
    ap.add_argument("--slab-rows", type=int, default=512,
                    help="rows of one GPU's share of a row-slab split of the --n grid: its stage "
                         "kernels at that shape (plus 2-row halos) against the full grid (0 to skip)")
    ap.add_argument("--selector-gain", type=int, default=512,
                    help="grid of the short-horizon routing comparison: network against random "
                         "routing at equal FP64 share, Orszag-Tang and Kelvin-Helmholtz (0 to skip)")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="instances of the batched-ensemble workload (0 to skip)")
    ap.add_argument("--ensemble-form", default="mixed")
    ap.add_argument("--ensemble-t-end", type=float, default=0.5)
    ap.add_argument("--time-budget", type=float, default=430.0,
                    help="seconds this process may run: the gate runs last and leaves out its "
                         "FP64-FMA floor run when that would end past the budget (0: no limit)")
    ap.add_argument("--whole-run", default="",
                    help="only run one whole Orszag-Tang --n^2 run to t=0.5 in this form and print "
                         "its wall time, steps and rate (for profiling the form's kernels)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 8, "launches": 3}), flush=True)
        return

    if args.whole_run:
        import mhdglm
        u, st, wall, sel = whole_run(mhdglm, args.whole_run, args.n)
        print(json.dumps({"whole_run": args.whole_run, "n": args.n, "steps": st.steps,
                          "code": st.code, "wall_s": round(wall, 3),
                          "cell_updates_per_s": round(args.n * args.n * st.steps / wall, 1),
                          "selector": sel}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this solver runs on one GPU (a single grid per device); "
              "multi-GPU runs are not supported -- use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import mhdglm

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(2)
    n = args.n
    res = time_form(mhdglm, torch, args.precision, n, args.steps, args.warmup, args.tile,
                    args.fuse)
    value, wall, ms = res["value"], res["wall"], res["kernel_ms"]
    esize = 4 if args.precision == "fp32" else 8
    fused = args.fuse == 2
    stages_per_launch = args.fuse or (
        mhdglm.load().mhd_autotuned_fuse(mhdglm.PRECISIONS[args.precision], n, n)
        if args.tile < 0 else 1)
    # The step against the HBM roofline: 144 B per FP64 cell-update (72 B in
    # FP32 state) over the step's kernel time (HIP events on the solver
    # stream around the stage kernels and the finish kernel).
    step_ms = ms[0] + ms[1] + ms[2]
    alg_bytes = STEP_BYTES_PER_CELL * esize * n * n
    achieved = alg_bytes / (step_ms * 1e-3) / 1e9
    tname = "step_fused" if fused else "stage"
    tfile = os.path.join(ROOT, "profiles", f"traffic_{tname}_ot{n}_{args.precision}.json")
    traffic = None
    valu = None
    kernels = {}
    prof = None
    if os.path.exists(tfile):
        with open(tfile) as f:
            prof = json.load(f)
    if fused:
        kernels["k_step_fused"] = {"ms": round(ms[0], 4), "compulsory_B_per_cell": 18 * esize}
    else:
        for stage in (1, 2):
            per = STAGE_BYTES_PER_CELL[stage] * esize
            kernels[f"k_stage<{stage}>"] = {
                "ms": round(ms[stage - 1], 4), "compulsory_B_per_cell": per,
                "achieved_GBs": round(per * n * n / (ms[stage - 1] * 1e-3) / 1e9, 1)}
    kernels["k_finish"] = {"ms": round(ms[2], 4)}
    if prof:
        b = [prof.get(f"stage{k}_bytes_per_launch") for k in ((1,) if fused else (1, 2))]
        if all(x is not None for x in b):
            traffic = round(sum(b))
        v = prof.get("stage1_valu")
        if v:
            # wave64 VALU instructions per CU per cycle (rocprofv3 counters),
            # against the FP64 issue limit of 1 (transcendentals counted 4x).
            valu = {"insts_per_launch": round(v["insts"]),
                    "issue_per_cu_cycle": round(v["issue_per_cu_cycle"], 3),
                    "issue_incl_trans": round(v["issue_incl_trans_quarter_rate"], 3),
                    "fp64_issue_limit": 1.0}
            if "flops_fp64_waves" in v and not fused:
                # FP64 FLOP of the two stage launches (SQ_INSTS_VALU_FLOPS_FP64,
                # per wave, x64 lanes) over their live HIP-event time, against
                # the 78.6 TFLOP/s FP64 vector peak.
                v2 = prof.get("stage2_valu", {})
                flops = 64.0 * (v["flops_fp64_waves"] + v2.get("flops_fp64_waves", 0.0))
                tf = flops / ((ms[0] + ms[1]) * 1e-3) / 1e12
                valu["fp64_flops_per_step"] = flops
                valu["fp64_tflops"] = round(tf, 2)
                valu["fp64_peak_tflops"] = FP64_PEAK_TFLOPS
                valu["fp64_frac"] = round(tf / FP64_PEAK_TFLOPS, 4)
                valu["share_of_valu_fp64_arith"] = round(
                    (v["add_f64"] + v["mul_f64"] + v["fma_f64_counter"] + v["trans_f64_counter"])
                    / v["insts"], 3)
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "unit_of_work": f"one SSPRK2 step of {n}x{n} cells, {STEP_BYTES_PER_CELL * esize} B "
                        "per cell-update (read U^n, write U^{n+1})",
        "step_kernel_ms": round(step_ms, 4),
        "frac_of_value": round(value * STEP_BYTES_PER_CELL * esize / 1e9 / HBM_PEAK_GBS, 4),
        "traffic_over_algorithmic": round(traffic / alg_bytes, 3) if traffic else None,
        "traffic_source": os.path.relpath(tfile, ROOT) if traffic else None,
        "kernels": kernels, "valu": valu,
        "limiter": "FP64 VALU work and its latency at 2 waves per SIMD, not HBM bandwidth",
    }
    forms = {args.precision: round(value, 1)}
    for form in [f for f in args.forms.split(",") if f and f != args.precision]:
        r = time_form(mhdglm, torch, form, n, args.steps, args.warmup, args.tile, args.fuse)
        forms[form] = round(r["value"], 1)
        if form == "selector":
            forms["selector_fp64_tile_fraction"] = round(r["fp64_tile_fraction"], 5)
            forms["selector_audits"] = [r["audits"], r["audit_misses"]]
    if "mixed" in forms and args.precision in ("fp64", "mixed"):
        forms["mixed_vs_fp64_speedup"] = round(forms["mixed"] / forms.get("fp64", value), 3)
    if "mixed" in forms and "fp64_fma" in forms:
        forms["mixed_vs_fp64_fma_speedup"] = round(forms["mixed"] / forms["fp64_fma"], 3)
    if "ref_mixed" in forms and args.precision == "fp64":
        forms["ref_mixed_vs_fp64_speedup"] = round(forms["ref_mixed"] / value, 3)
    accuracy = None
    if args.accuracy_n:
        accuracy = accuracy_check(mhdglm, args.accuracy_n,
                                  [f for f in ("fp64_fma", "mixed", "selector", "mixed_kahan",
                                               "fp32", "ref_mixed")
                                   if f in forms or f == "fp64_fma"])
    kh = None
    if args.kh_n:
        kh = accuracy_check(mhdglm, args.kh_n, [f for f in args.kh_forms.split(",") if f],
                            problem="khb", t_end=args.kh_t_end)
    selector_run = None
    if args.selector_n:
        selector_run = accuracy_check(mhdglm, args.selector_n, ["fp64_fma", "mixed", "selector"])
    breakdown = selector_breakdown(mhdglm, args.selector_breakdown) if args.selector_breakdown else None
    slab = slab_share(mhdglm, n, args.slab_rows) if args.slab_rows else None
    gain = None
    if args.selector_gain:
        gain = [selector_gain(mhdglm, args.selector_gain),
                selector_gain(mhdglm, args.selector_gain, t0=2.0, problem="khb")]
    ensemble = None
    if args.ensemble:
        ensemble = run_ensemble(mhdglm, torch, args.ensemble, args.ensemble_form,
                                args.ensemble_t_end)
    large = None
    if args.large_n:
        large = {"n": args.large_n, "steps": 20, "warmup": 5}
        for form in ("fp64", "mixed", "selector"):
            large[form] = round(time_form(mhdglm, torch, form, args.large_n, 20, 5)["value"], 1)
        large["mixed_vs_fp64_speedup"] = round(large["mixed"] / large["fp64"], 3)
        large["selector_vs_fp64_speedup"] = round(large["selector"] / large["fp64"], 3)
    npar = args.cpu_parallel if args.cpu_parallel >= 0 else min(os.cpu_count() or 1, 16)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n, args.cpu_t_end, npar)
    # The headline's whole runs last: the longest section, and the only one
    # that shortens itself to the time budget.
    gate = None
    if args.gate_n:
        gate = gate_check(mhdglm, args.gate_n, with_floor=not args.no_gate_floor,
                          budget_s=args.time_budget or None)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "fp64_fma": "f64",
                  "mixed": "f32 flux / f64 state+update", "mixed_kahan": "f32 flux / f32x2 state",
                  "ref_mixed": "f32 recovery+recon+rhs / f64 flux+state"}[args.precision],
        "gate": gate,
        "forms": forms,
        "accuracy": accuracy,
        "large": large,
        "kh": kh,
        "selector_run": selector_run,
        "selector_breakdown": breakdown,
        "selector_gain": gain,
        "slab_share": slab,
        "ensemble": ensemble,
        "data": "synthetic: Orszag-Tang initial condition evaluated on the host",
        "config": {
            "workload": f"Orszag-Tang {n}x{n}, MUSCL+HLLD+GLM SSPRK2, {args.precision}",
            "nx": n, "ny": n, "problem": "ot", "precision": args.precision,
            "cfl": 0.25, "alpha": 0.18, "hlld_tol": 1e-4,
            "step": "one SSPRK2 step of the whole grid",
            "stages_per_launch": stages_per_launch,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
