"""GPU tests of the learned precision selector (MIXED form).

  * MLP forward on MFMA (BF16 operands, FP32 accumulation) against the CPU
    oracle's BF16 network (oracle/selector_oracle.py), ragged batch sizes;
  * MLP gradients (device backward on the FP32 matrix cores) against the
    oracle's float64 gradients of the same BF16 forward; device training
    lowers the loss on a separable problem and is bitwise repeatable;
  * tile features and FP32-vs-FP64 tile errors against the oracle;
  * routing: every tile FP64 is bitwise the FP64 form, no tile FP64 is
    bitwise the mixed form, a partial selection takes each tile's stage-1
    update from the kernel the flags name, audits leave the state alone and
    a zero budget makes audited tiles sticky;
  * gain: the trained network (Orszag-Tang and Kelvin-Helmholtz training
    runs), on held-out Orszag-Tang variants, catches more of the tiles whose
    FP32 flux error is largest than the best single-feature rule (chosen on
    the training tiles) and several times more than a random choice (the
    audits' share) of the same number of FP64 tiles; from one FP64
    Orszag-Tang state, the shipped network's routing removes several times
    more of the mixed form's error over the next steps than random routing
    of as many tiles;
  * the drop-in executable with MHD_SELECTOR runs the same selected form as
    the library, bitwise.
"""
import os
import re
import sys

import numpy as np
import pytest

import selector_oracle as so

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def flat(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float32).ravel()
                           for k in ("W1", "b1", "W2", "b2", "w3", "b3")])


def unflat(g):
    H, NIN = so.H, so.NIN
    o, out = 0, {}
    for k, n in (("W1", NIN * H), ("b1", H), ("W2", H * H), ("b2", H), ("w3", H), ("b3", 1)):
        out[k] = np.asarray(g[o:o + n], dtype=np.float64)
        o += n
    out["W1"] = out["W1"].reshape(NIN, H)
    out["W2"] = out["W2"].reshape(H, H)
    return out


def padded(x):
    xp = np.zeros((x.shape[0], so.NIN), dtype=np.float32)
    xp[:, :so.NFEAT] = x
    return xp


@pytest.mark.parametrize("n", [1, 17, 200, 1000])
def test_mlp_forward_matches_oracle(mhd, n):
    say(f"mlp forward n={n}")
    rng = np.random.default_rng(n)
    p = so.init_params(7)
    p["b1"] = (0.1 * rng.standard_normal(so.H)).astype(np.float32)
    p["b2"] = (0.1 * rng.standard_normal(so.H)).astype(np.float32)
    p["b3"] = np.array([0.2], dtype=np.float32)
    x = rng.standard_normal((n, so.NFEAT)).astype(np.float32)
    got = mhd.mlp_forward(flat(p), padded(x))
    ref = so.mlp_forward(x, p, bf16_ops=True)["p"]
    f64 = so.mlp_forward(x, {k: v.astype(np.float64) for k, v in p.items()}, bf16_ops=False)["p"]
    d = np.abs(got - ref).max()
    say(f"  max |prob - oracle| {d:.2e}; vs unrounded network {np.abs(got - f64).max():.2e}")
    assert d <= 1e-4


def test_mlp_gradients_match_oracle(mhd):
    say("mlp gradients n=300")
    rng = np.random.default_rng(11)
    p = so.init_params(5)
    p["b1"] = (0.1 * rng.standard_normal(so.H)).astype(np.float32)
    x = rng.standard_normal((300, so.NFEAT)).astype(np.float32)
    y = (rng.random(300) < 0.3).astype(np.float32)
    w = np.where(y > 0, 2.5, 1.0).astype(np.float32)
    g, loss = mhd.mlp_gradients(flat(p), padded(x), y, w)
    gref, lref = so.mlp_grads(x, y.astype(np.float64), w.astype(np.float64), p, bf16_ops=True)
    assert loss == pytest.approx(lref, rel=1e-4)
    gd = unflat(g)
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        ref = np.asarray(gref[k], dtype=np.float64).reshape(gd[k].shape)
        scale = np.abs(ref).max()
        err = np.abs(gd[k] - ref).max()
        say(f"  {k}: max |g| {scale:.3e}, max err {err:.3e}")
        assert scale > 0
        assert err <= 2e-3 * scale + 1e-7, k
    assert np.all(gd["W1"][so.NFEAT:] == 0.0)  # padding rows see zero input


def test_mlp_training_reduces_loss(mhd):
    from mhdglm import selector

    rng = np.random.default_rng(3)
    n = 6000
    x = rng.standard_normal((n, so.NFEAT)).astype(np.float32)
    y = ((x[:, 0] + 0.5 * x[:, 3] - 0.3 * x[:, 7]) > 0.7).astype(np.float32)
    w = np.ones(n, dtype=np.float32)
    p, losses = mhd.mlp_train(selector.init_params(0), padded(x), y, w, epochs=25, batch=256,
                              lr=3e-3, seed=1)
    acc = np.mean((mhd.mlp_forward(p, padded(x)) > 0.5) == (y > 0))
    say(f"training: loss {losses[0]:.4f} -> {losses[-1]:.4f}, accuracy {acc:.4f}")
    assert losses[-1] < 0.3 * losses[0]
    assert acc > 0.97
    assert np.all(p[so.NFEAT * so.H:so.NIN * so.H] == 0.0)


def test_mlp_training_is_deterministic(mhd):
    """The gradient rows of a minibatch's blocks are summed in block order
    (no atomics): the same data and seed train the same weights, bit for
    bit, and the gradients of one batch are the same on every call."""
    from mhdglm import selector

    rng = np.random.default_rng(8)
    n = 3000
    x = padded(rng.standard_normal((n, so.NFEAT)).astype(np.float32))
    y = (rng.random(n) < 0.2).astype(np.float32)
    w = np.where(y > 0, 4.0, 1.0).astype(np.float32)
    runs = [mhd.mlp_train(selector.init_params(2), x, y, w, epochs=4, batch=512, lr=2e-3, seed=5)
            for _ in range(2)]
    assert np.array_equal(runs[0][0].view(np.uint32), runs[1][0].view(np.uint32))
    assert np.array_equal(runs[0][1], runs[1][1])
    g1, l1 = mhd.mlp_gradients(runs[0][0], x, y, w)
    g2, l2 = mhd.mlp_gradients(runs[0][0], x, y, w)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32)) and l1 == l2


def mixed_solver(mhd, nx, ny, problem="ot", steps=0, **kw):
    kw.setdefault("tile", 3)  # 16x15, the tile grid the tests count
    s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=mhd.PREC_MIXED, **kw))
    s.init_problem(problem)
    if steps:
        s.advance(10.0, steps)
    return s


@pytest.mark.parametrize("nx,ny,tile", [(96, 80, 3), (90, 77, 3), (90, 77, 1), (12, 10, 3)])
def test_tile_features_match_oracle(mhd, nx, ny, tile):
    """Partial tiles at the high edges (80 and 77 rows, 90 columns) included:
    their right and top neighbours wrap around the grid; 16x15 and 32x8 tiles
    (the largest box of a tile and its neighbour row and column), and a grid
    smaller than one tile."""
    tx, ty = {1: (32, 8), 3: (16, 15)}[tile]
    say(f"tile features OT {nx}x{ny}, {tx}x{ty} tiles, after 40 steps")
    s = mixed_solver(mhd, nx, ny, steps=40, tile=tile)
    f = s.tile_features()
    ch = s.status().ch_last
    u = s.get_state(0)
    s.close()
    ref = so.tile_features(u, tx, ty, ch)
    assert f.shape == (-(-nx // tx) * -(-ny // ty), so.NIN)
    assert np.all(f[:, so.NFEAT:] == 0.0)
    d = np.abs(f[:, :so.NFEAT] - ref) / (1.0 + np.abs(ref))
    say(f"  max scaled feature difference {d.max():.2e}")
    assert d.max() <= 1e-5


def test_pair_tile_errors_match_oracle(mhd):
    """FP32-flux vs FP64-flux stage-1 difference per tile, against the oracle
    applied to the stage-1 states of a FP64 and a mixed context."""
    say("pair tile errors OT 96x80")
    s = mixed_solver(mhd, 96, 80)
    err = s.pair_tile_errors()
    s.advance(10.0, 1)
    u1m = s.get_state(1)
    s.close()
    s64 = mhd.Solver(mhd.SolverConfig(nx=96, ny=80, precision=mhd.PREC_FP64))
    s64.init_problem("ot")
    s64.advance(10.0, 1)
    u164 = s64.get_state(1)
    s64.close()
    ref = so.tile_errors(u1m, u164, 16, 15)
    say(f"  errors {err.min():.2e} .. {err.max():.2e}")
    assert err.max() > 0
    assert np.allclose(err, ref, rtol=1e-6, atol=0)


def run_steps(mhd, precision, nx, ny, steps, model=None, problem="ot", which=0, **sel):
    s = mhd.Solver(mhd.SolverConfig(nx=nx, ny=ny, precision=precision, tile=3))
    s.init_problem(problem)
    if model is not None:
        s.configure_selector(model, **sel)
    st = s.advance(10.0, steps)
    assert st.code == 0
    u = s.get_state(which)
    stats = s.selector_stats() if model is not None else None
    s.close()
    return u, stats


def random_model(seed=0):
    from mhdglm import selector

    return {"params": selector.init_params(seed), "mean": np.zeros(so.NFEAT, np.float32),
            "std": np.ones(so.NFEAT, np.float32), "threshold": 0.5, "budget": 1e30}


@pytest.mark.parametrize("steps", [1, 40])
def test_selector_all_fp64_is_fp64(mhd, steps):
    say(f"selector threshold -1 (all FP64), {steps} steps")
    u64, _ = run_steps(mhd, mhd.PREC_FP64, 96, 80, steps)
    us, st = run_steps(mhd, mhd.PREC_MIXED, 96, 80, steps, random_model(), threshold=-1.0,
                       audit_frac=0.0)
    assert np.array_equal(u64.view(np.uint64), us.view(np.uint64))
    assert st["fp64_tile_steps"] == 36 * steps and st["steps"] == steps and st["audits"] == 0


@pytest.mark.parametrize("audit_frac", [0.0, 0.3])
def test_selector_none_flagged_is_mixed(mhd, audit_frac):
    say(f"selector threshold 2 (no FP64 tile), audit fraction {audit_frac}")
    um, _ = run_steps(mhd, mhd.PREC_MIXED, 96, 80, 40)
    us, st = run_steps(mhd, mhd.PREC_MIXED, 96, 80, 40, random_model(), threshold=2.0,
                       audit_frac=audit_frac, audit_budget=1e30)
    assert np.array_equal(um.view(np.uint64), us.view(np.uint64))
    assert st["fp64_tile_steps"] == 0 and st["misses"] == 0
    if audit_frac > 0:
        say(f"  audits {st['audits']} of {36 * 40} tile-steps")
        assert 0.15 * 36 * 40 < st["audits"] < 0.45 * 36 * 40
    else:
        assert st["audits"] == 0


def test_selector_counts_only_live_steps(mhd):
    """A run that ends inside a replayed batch of steps: the batch's tail
    steps do nothing, and the selector samples, audits and counts nothing in
    them (audits there would compare stale buffers); advancing a finished
    run again changes nothing."""
    say("selector to t_end inside a batch")
    s = mhd.Solver(mhd.SolverConfig(nx=96, ny=80, precision=mhd.PREC_MIXED, tile=3))
    s.init_problem("ot")
    steps = s.advance(0.05).steps
    s.close()
    batch = 32 if steps % 32 else 31
    s = mhd.Solver(mhd.SolverConfig(nx=96, ny=80, precision=mhd.PREC_MIXED, batch_steps=batch,
                                    tile=3))
    s.init_problem("ot")
    s.configure_selector(random_model(), threshold=2.0, audit_frac=0.3, audit_budget=1e30)
    st = s.advance(0.05)
    a = s.selector_stats()
    st2 = s.advance(0.05)
    b = s.selector_stats()
    s.close()
    say(f"  steps {st.steps}, stats {a}")
    assert st.done == 1 and st.steps == steps and steps % batch != 0
    assert a["steps"] == st.steps and a["misses"] == 0
    assert 0.2 * 36 * st.steps < a["audits"] < 0.4 * 36 * st.steps
    assert st2.steps == st.steps and b == a


@pytest.mark.parametrize("nx,ny", [(96, 80), (2048, 2048)])
def test_selector_partial_routing(mhd, nx, ny):
    """Stage 1 of one step: each tile's update comes from the FP64 kernel if
    flagged, else from the mixed kernel (flags predicted on the host from
    the same features and network).  At 2048^2 the 17,536 tiles take the
    list builder two rounds."""
    model = random_model(4)
    model["mean"] = np.zeros(so.NFEAT, np.float32)
    s = mixed_solver(mhd, nx, ny)
    feats = s.tile_features()
    s.close()
    prob = mhd.mlp_forward(model["params"], feats)
    thr = float(np.median(prob))
    flags = prob > thr
    say(f"partial routing: {flags.sum()} of {flags.size} tiles FP64 (threshold {thr:.6f})")
    assert 0 < flags.sum() < flags.size
    u64, _ = run_steps(mhd, mhd.PREC_FP64, nx, ny, 1, which=1)
    um, _ = run_steps(mhd, mhd.PREC_MIXED, nx, ny, 1, which=1)
    us, st = run_steps(mhd, mhd.PREC_MIXED, nx, ny, 1, model, which=1, threshold=thr,
                       audit_frac=0.0)
    assert st["fp64_tile_steps"] == flags.sum()
    assert not np.array_equal(u64, um)
    tiles_x = -(-nx // 16)
    assert flags.size == tiles_x * -(-ny // 15)
    for t in range(flags.size):
        j0, i0 = (t // tiles_x) * 15, (t % tiles_x) * 16
        sl = (slice(None), slice(j0, j0 + 15), slice(i0, i0 + 16))
        want = u64[sl] if flags[t] else um[sl]
        assert np.array_equal(us[sl].view(np.uint64), want.view(np.uint64)), t


def test_selector_audit_misses_make_tiles_sticky(mhd):
    say("audits with zero budget")
    u64, _ = run_steps(mhd, mhd.PREC_FP64, 96, 80, 30)
    us, st = run_steps(mhd, mhd.PREC_MIXED, 96, 80, 30, random_model(), threshold=2.0,
                       audit_frac=0.2, audit_budget=0.0)
    say(f"  stats {st}")
    assert st["audits"] > 0 and st["misses"] > 0
    assert st["fp64_tile_steps"] > 0
    # once every tile has missed, the run is the FP64 one from then on; here
    # only part of the tiles are sticky, so the state lies between the two.
    assert not np.array_equal(us, u64)


FEATURE_NAMES = ("min beta", "max beta", "d rho", "d p", "d B", "Mach", "psi/B", "div B",
                 "B_n min", "v/c_h", "c_f/c_h", "ke/me")


def top_recall(score, unsafe, k):
    return unsafe[np.argsort(-score, kind="stable")[:k]].mean()


def test_selector_gain_over_single_features_and_random(mhd):
    """Trained on the device on Orszag-Tang and both Kelvin-Helmholtz
    problems (256^2, each run's labels its top 5% of FP32-vs-FP64 stage
    errors), evaluated on two held-out Orszag-Tang variants at 192^2: among
    as many tiles as are unsafe, the network catches more unsafe tiles than
    the best single-feature rule (the feature and sign that rank the training
    tiles best) and several times more than a random choice.  Audits without
    the network sample the mixed tiles at random, so at equal FP64 share
    they catch the random share."""
    from mhdglm import selector

    Xs_, Es_, Ys_ = [], [], []
    for problem, t_end in (("ot", 0.5), ("kh", 5.0), ("khb", 5.0)):
        X, E = selector.collect(problem, n=256, t_end=t_end, snapshots=64, log=say)
        Xs_.append(X)
        Es_.append(E)
        Ys_.append(selector.run_labels(E))
    X, E, Y = np.concatenate(Xs_), np.concatenate(Es_), np.concatenate(Ys_)
    model = selector.fit(X, E, budget=float(np.quantile(Es_[0], 0.95)), labels=Y, epochs=30,
                         log=say)
    # The single-feature rule, chosen on the training tiles.
    train_unsafe, kt = Y > 0, int(Y.sum())
    rules = [(f, sign) for f in range(so.NFEAT) for sign in (1.0, -1.0)]
    f_best, s_best = max(rules, key=lambda r: top_recall(r[1] * X[:, r[0]], train_unsafe, kt))
    rule = ("+" if s_best > 0 else "-") + FEATURE_NAMES[f_best]
    for params in ((0.85, 0.5, 1.2, 0.3, 2.0), (1.1, 2.5, 0.2, 4.0, 1.0)):
        Xv, Ev = selector.collect("otv", params=params, n=192, t_end=0.5, snapshots=48, log=say)
        Xs = np.zeros_like(Xv)
        Xs[:, :so.NFEAT] = (Xv[:, :so.NFEAT] - model["mean"]) / model["std"]
        prob = mhd.mlp_forward(model["params"], Xs)
        unsafe = selector.run_labels(Ev) > 0
        k = int(unsafe.sum())
        recall = top_recall(prob, unsafe, k)
        single = top_recall(s_best * Xv[:, f_best], unsafe, k)
        say(f"held-out {params}: {k} unsafe of {len(Ev)}; network {recall:.3f}, single-feature "
            f"rule {rule} {single:.3f}, random {k / len(Ev):.3f}")
        assert k > 200
        assert recall > single
        assert recall >= 3.0 * k / len(Ev)


def test_shipped_selector_lowers_short_horizon_error(mhd):
    """What the routing buys before chaos: from one FP64 Orszag-Tang state
    (512^2, t=0.15), k steps in FP64, in the mixed form, and in the mixed
    form routed by the shipped network and by a random network at the same
    FP64 tile share (audits off).  Over 1 and 16 steps the network's routing
    removes at least 10% of the mixed form's density error against FP64 at
    its own threshold and at a 20% share, and at least twice what random
    routing of as many tiles removes."""
    from mhdglm import compare, selector

    n = 512
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n))
    s.init_problem("ot")
    assert s.advance(0.15).code == 0
    u0 = s.get_state(0)
    s.close()
    c = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=mhd.PREC_MIXED))
    c.set_state(u0)
    feats = c.tile_features()
    c.close()
    shipped = selector.load(selector.DEFAULT_MODEL)
    rnd = random_model(7)

    def prob(model):
        x = np.zeros_like(feats)
        x[:, :so.NFEAT] = (feats[:, :so.NFEAT] - model["mean"]) / model["std"]
        return mhd.mlp_forward(model["params"], x)

    ps, pr = prob(shipped), prob(rnd)

    def run(k, precision, model=None, thr=None):
        r = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=precision))
        r.set_state(u0)
        if model is not None:
            r.configure_selector(model, threshold=thr, audit_frac=0.0)
        assert r.advance(10.0, k).code == 0
        u = r.get_state(0)
        share = None
        if model is not None:
            st = r.selector_stats()
            share = st["fp64_tile_steps"] / (st["steps"] * r.selector_tiles())
        r.close()
        return u, share

    own = float(np.mean(ps > float(shipped["threshold"])))
    assert 0.02 < own < 0.5
    for k in (1, 16):
        ref, _ = run(k, mhd.PREC_FP64)
        e_mixed = compare.rel_l2_state(ref, run(k, mhd.PREC_MIXED)[0], ("rho",))["rho"]
        for thr, frac in ((float(shipped["threshold"]), own), (float(np.quantile(ps, 0.8)), 0.2)):
            un, sn = run(k, mhd.PREC_MIXED, shipped, thr)
            ur, sr = run(k, mhd.PREC_MIXED, rnd, float(np.quantile(pr, 1.0 - frac)))
            e_net = compare.rel_l2_state(ref, un, ("rho",))["rho"]
            e_rnd = compare.rel_l2_state(ref, ur, ("rho",))["rho"]
            say(f"k={k} share {frac:.3f}: FP64 shares network {sn:.3f} random {sr:.3f}; relL2 rho "
                f"mixed {e_mixed:.3e} network {e_net:.3e} random {e_rnd:.3e}")
            assert abs(sn - sr) < 0.25 * frac
            assert e_net < 0.9 * e_mixed
            assert e_mixed - e_net > 2.0 * (e_mixed - e_rnd)


def test_concurrent_schedule_is_the_serial_one(mhd, monkeypatch):
    """MHD_SELECTOR_CONCURRENT=1 runs the flagged and audit tiles' FP64 list
    launch on a second stream beside the mixed stage kernel (fork and join
    inside the captured step).  From a mid-flow Orszag-Tang state (tiles
    flagged and audited), 40 steps in graph batches of 32 (then 8 steps
    enqueued one by one) and of 1, which the selector rounds up to its
    refresh period of 8: the state, the status and the selector's counters
    are the same in every case, bit for bit."""
    from mhdglm import selector

    n = 256
    s = mhd.Solver(mhd.SolverConfig(nx=n, ny=n))
    s.init_problem("ot")
    assert s.advance(0.2).code == 0
    u0 = s.get_state(0)
    s.close()
    model = selector.load(selector.DEFAULT_MODEL)
    out = {}
    for conc in ("0", "1"):
        monkeypatch.setenv("MHD_SELECTOR_CONCURRENT", conc)
        for batch in (0, 1):
            r = mhd.Solver(mhd.SolverConfig(nx=n, ny=n, precision=mhd.PREC_MIXED,
                                            batch_steps=batch))
            r.set_state(u0)
            r.configure_selector(model, threshold=float(model["threshold"]) * 0.5)
            st = r.advance(10.0, 40)
            out[(conc, batch)] = (r.get_state(0), st.as_dict(), r.selector_stats())
            r.close()
    u, st, stats = out[("0", 0)]
    say(f"serial: {stats}")
    assert st["code"] == 0 and stats["fp64_tile_steps"] > 0 and stats["audits"] > 0
    for key, (u2, st2, stats2) in out.items():
        assert np.array_equal(u.view(np.uint64), u2.view(np.uint64)), key
        assert st2 == st and stats2 == stats, key


def test_cli_selector_run_is_the_library_run(mhd, tmp_path):
    """The drop-in executable with MHD_PRECISION=mixed and MHD_SELECTOR set to
    the shipped model runs the chosen form: its .dat density column and step
    count are those of the same run through the Python binding, bitwise
    (%.17e round-trips doubles), and it reports the selector's counters."""
    from mhdglm import selector

    n, t_end = 128, 0.1
    env = dict(os.environ, MHD_SELECTOR=selector.DEFAULT_MODEL)
    r = mhd.run_cli([n, n, t_end, 0.25, "muscl", "hlld", 0.18], str(tmp_path), precision="mixed",
                    env=env)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Selector: FP64 tile-steps = (\d+) .*audits = (\d+), audit misses = (\d+)",
                  r.stdout)
    assert m, r.stdout
    steps = int(re.search(r"Done\. Steps: (\d+)", r.stdout).group(1))
    (f,) = list(tmp_path.iterdir())
    assert f.name == f"orszag_tang_glm_muscl_hlld_mixed_N{n}x{n}.dat"
    rho = np.loadtxt(f, comments="#", usecols=2).reshape(n, n)
    s = mixed_solver(mhd, n, n, tile=-1)  # the CLI's tile: the tuned one
    s.configure_selector(selector.load(selector.DEFAULT_MODEL))
    st = s.advance(t_end)
    u = s.get_state(0)
    stats = s.selector_stats()
    s.close()
    assert st.steps == steps
    assert (stats["fp64_tile_steps"], stats["audits"], stats["misses"]) == tuple(
        int(x) for x in m.groups())
    assert np.array_equal(rho.view(np.uint64), u[0].view(np.uint64))
