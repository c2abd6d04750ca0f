"""Learned precision selector: training data, training, model files.

The solver produces its own training data (mhd_tile_features and
mhd_pair_tile_errors of a MIXED context): for every tile of a state, 12
features and the largest scaled difference between the FP32-flux and the
FP64-flux stage update.  A tile is "unsafe" (label 1) when that difference
exceeds the budget.  The MLP is trained on the device (mhd_mlp_train); the
model (weights, feature standardisation, threshold, budget) is kept as an
.npz file.

    python -m mhdglm.selector --out model.npz [--n 512] [--snapshots 190]
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

from . import PREC_MIXED, Solver, SolverConfig, mlp_forward, mlp_param_count, mlp_train

NFEAT = 12
NIN = 16
H = 64
# Audits pin a tile to FP64 only for a gross miss: an FP32-vs-FP64 stage
# difference this many times the label budget.  Misses at the budget itself
# are round-off variation (every tile exceeds the budget now and then).
AUDIT_FACTOR = 10.0
# Model trained by `python -m mhdglm.selector` (Orszag-Tang, two variants
# and both Kelvin-Helmholtz problems at 512^2), shipped with the package.
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "selector_model.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised FP32 weights in the device layout (padding rows zero)."""
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((NIN, H)) * np.sqrt(2.0 / NFEAT)
    w1[NFEAT:] = 0.0
    parts = [w1.ravel(), np.zeros(H), (rng.standard_normal((H, H)) * np.sqrt(2.0 / H)).ravel(),
             np.zeros(H), rng.standard_normal(H) * np.sqrt(1.0 / H), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == mlp_param_count()
    return p


# Tile shape of the training data (16x15, the stage kernels' default and the
# autotuned shape from 1024^2 up): features are tile aggregates.
TRAIN_TILE = 3


def collect(problem="ot", params=(), n=512, steps=600, every=4, t_end=None, snapshots=24,
            log=print, tile=TRAIN_TILE):
    """Features and pair errors of every tile of a MIXED run of `problem` at
    n x n from t=0: every `every` steps for `steps` steps, or (t_end given)
    at `snapshots` evenly spaced times in [0, t_end)."""
    s = Solver(SolverConfig(nx=n, ny=n, precision=PREC_MIXED, tile=tile))
    if params:
        from . import problem_state_ex
        s.set_state(problem_state_ex(problem, params, n, n))
    else:
        s.init_problem(problem)
    X, E = [], []
    marks = ([t_end * k / snapshots for k in range(1, snapshots + 1)] if t_end is not None
             else [None] * len(range(0, steps, every)))
    for mark in marks:
        X.append(s.tile_features())
        E.append(s.pair_tile_errors())
        st = s.advance(mark, -1) if mark is not None else s.advance(10.0, every)
        if st.code != 0:
            break
    s.close()
    X, E = np.concatenate(X), np.concatenate(E)
    log(f"selector: collected {len(E)} tiles from {problem} {n}^2")
    return X, E


def fit(X, E, budget=None, unsafe_fraction=0.05, epochs=30, batch=2048, lr=2e-3, seed=0,
        labels=None, log=print):
    """Trains the MLP on (features, pair errors).  budget defaults to the
    (1 - unsafe_fraction) quantile of the errors; `labels` (0/1 per tile)
    replace E > budget when given (budget then only sets the audits').
    Returns the model dict."""
    X = np.asarray(X, dtype=np.float32)
    E = np.asarray(E, dtype=np.float64)
    if budget is None:
        budget = float(np.quantile(E, 1.0 - unsafe_fraction))
    y = (E > budget).astype(np.float32) if labels is None else np.asarray(labels, np.float32)
    mean = X[:, :NFEAT].mean(axis=0)
    std = X[:, :NFEAT].std(axis=0) + 1e-6
    Xs = np.zeros_like(X)
    Xs[:, :NFEAT] = (X[:, :NFEAT] - mean) / std
    pos = max(float(y.sum()), 1.0)
    w = np.where(y > 0, (len(y) - pos) / pos, 1.0).astype(np.float32)
    params, losses = mlp_train(init_params(seed), Xs, y, w, epochs=epochs, batch=batch, lr=lr,
                               seed=seed)
    prob = mlp_forward(params, Xs)
    # Operating point: flag as many tiles as are over budget (the class
    # weights balance the training, so 0.5 would flag several times more).
    threshold = float(np.quantile(prob, 1.0 - pos / len(y)))
    pred = prob > threshold
    tp = float(np.sum(pred & (y > 0)))
    recall = tp / pos
    precision = tp / max(float(pred.sum()), 1.0)
    log(f"selector: {len(y)} tiles, budget {budget:.3e}, unsafe {pos / len(y):.3%}, "
        f"loss {losses[0]:.4f} -> "
        f"{losses[-1]:.4f}, recall {recall:.3f}, precision {precision:.3f}, "
        f"flagged {pred.mean():.3%}")
    return {"params": params, "mean": mean.astype(np.float32), "std": std.astype(np.float32),
            "threshold": np.float32(threshold), "budget": np.float32(budget),
            "audit_budget": np.float32(AUDIT_FACTOR * budget),
            "losses": losses, "recall": recall, "precision": precision,
            "flagged": float(pred.mean())}


def save(model, path):
    np.savez(path, **{k: np.asarray(v) for k, v in model.items()})


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# Orszag-Tang variants (B amplitude, phases of vx, vy, Bx, By) added to the
# training data of the shipped model.
TRAIN_VARIANTS = ((0.9, 0.7, 2.1, 0.4, 1.3), (1.15, 3.0, 0.5, 2.2, 4.1))

# The shipped model's training runs: Orszag-Tang (and the two variants) to
# t = 0.5 and both Kelvin-Helmholtz problems to t = 5, all at 512^2, each
# sampled at evenly spaced times: (problem, parameters, t_end).
TRAIN_RUNS = (("ot", (), 0.5), ("otv", TRAIN_VARIANTS[0], 0.5), ("otv", TRAIN_VARIANTS[1], 0.5),
              ("kh", (), 5.0), ("khb", (), 5.0))


def run_labels(E, unsafe_fraction=0.05):
    """Labels of one run's tiles: the top `unsafe_fraction` of its
    FP32-vs-FP64 stage differences.  The level of those differences is set
    by the flow (Kelvin-Helmholtz tiles all lie above the Orszag-Tang 95th
    percentile, by the FP32 rounding of their larger energies), so one
    global budget would label a whole problem unsafe; within each run the
    labels mark the tiles whose FP32 flux is furthest from FP64."""
    return (E > np.quantile(E, 1.0 - unsafe_fraction)).astype(np.float32)


def training_set_512(n=512, snapshots=190, unsafe_fraction=0.05, log=print):
    """Tiles of TRAIN_RUNS at n^2 (about 1.06e6 tiles at 512^2 with 190
    snapshots each).  Returns (X, E, labels, E_ot): features, errors,
    per-run labels and the errors of the plain Orszag-Tang run (the audits'
    budget reference)."""
    Xs, Es, Ys, ot = [], [], [], None
    for problem, params, t_end in TRAIN_RUNS:
        X, E = collect(problem, params=params, n=n, t_end=t_end, snapshots=snapshots, log=log)
        Xs.append(X)
        Es.append(E)
        Ys.append(run_labels(E, unsafe_fraction))
        if problem == "ot":
            ot = E
    return np.concatenate(Xs), np.concatenate(Es), np.concatenate(Ys), ot


def training_set(sizes=(256, 512, 1024, 2048), early_n=4096, log=print):
    """Tiles of Orszag-Tang runs to t=0.5 at each size (24 snapshots), of the
    two variants at the middle size, and of the first 600 steps at early_n.
    Returns (X, E, E_ot): features, errors, and the errors of the plain
    Orszag-Tang runs (the budget's reference)."""
    parts, ot = [], []
    for n in sizes:
        X, E = collect("ot", n=n, t_end=0.5, snapshots=24, log=log)
        parts.append((X, E))
        ot.append(E)
    mid = sizes[len(sizes) // 2]
    for v in TRAIN_VARIANTS:
        parts.append(collect("otv", params=v, n=mid, t_end=0.5, snapshots=24, log=log))
    if early_n:
        X, E = collect("ot", n=early_n, steps=600, every=100, log=log)
        parts.append((X, E))
        ot.append(E)
    return (np.concatenate([x for x, _ in parts]), np.concatenate([e for _, e in parts]),
            np.concatenate(ot))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=512,
                    help="grid of the training runs (TRAIN_RUNS: OT and KH)")
    ap.add_argument("--snapshots", type=int, default=190)
    ap.add_argument("--ot-sizes", default="",
                    help="instead: Orszag-Tang only, at these sizes (the earlier training set)")
    ap.add_argument("--early-n", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--unsafe-fraction", type=float, default=0.05)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.ot_sizes:
        X, E, Eot = training_set(tuple(int(x) for x in args.ot_sizes.split(",")), args.early_n,
                                 log)
        labels = None
    else:
        X, E, labels, Eot = training_set_512(args.n, args.snapshots, args.unsafe_fraction, log)
    # The audits' budget is a quantile of the Orszag-Tang errors (the
    # benchmark problem).
    budget = float(np.quantile(Eot, 1.0 - args.unsafe_fraction))
    model = fit(X, E, budget=budget, epochs=args.epochs, labels=labels, log=log)
    save(model, args.out)
    log(f"selector: model written to {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
