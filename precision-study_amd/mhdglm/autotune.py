"""Autotuner of the step kernels' tile shape and fusion depth.

    python -m mhdglm.autotune [--forms fp64,mixed] [--sizes 256,1024,4096] [--out FILE]

For every arithmetic form and grid size it times, on the current device,
each tile shape with one launch per SSPRK2 stage and the 16x15 tile with
both stages in one launch (HIP events around the step's kernels, after a
warm-up), and writes the table the library reads when a caller leaves the
tile at -1 (and the fusion depth at 0): lines "form max_cells tile fuse",
sizes ascending, so the best choice for the smallest tuned size at or above
a grid's cell count is used.
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys

from . import PKG_ROOT, PRECISIONS, Solver, SolverConfig

TILES = {0: "16x16", 1: "32x8", 2: "16x8", 3: "16x15", 4: "32x7", 5: "16x7"}
# (tile, stages per launch) candidates: every tile one stage per launch, and
# the two-stage kernel (16x15 only).
CHOICES = [(t, 1) for t in TILES] + [(3, 2)]
DEFAULT_OUT = os.path.join(PKG_ROOT, "autotune.txt")


def label(choice) -> str:
    tile, fuse = choice
    return TILES[tile] + ("" if fuse == 1 else " fused")


def time_choice(form: str, n: int, tile: int, fuse: int = 1, steps: int = 20) -> float:
    """Average ms of one step's kernels (OT n x n), the better of two rounds
    of `steps` steps."""
    s = Solver(SolverConfig(nx=n, ny=n, precision=PRECISIONS[form], tile=tile, fuse=fuse))
    s.init_problem("ot")
    s.advance(0.5, 3)
    best = min(sum(s.time_kernels(0.5, steps)) for _ in range(2))
    s.close()
    return best


def tune(forms, sizes, log=print):
    rows, report = [], []
    for form in forms:
        for n in sorted(sizes):
            times = {ch: time_choice(form, n, *ch) for ch in CHOICES}
            best = min(times, key=times.get)
            rows.append((PRECISIONS[form], n * n, best[0], best[1]))
            report.append((form, n, best, times))
            log(f"autotune {form} {n}^2: " +
                ", ".join(f"{label(c)} {times[c]:.4f} ms" for c in CHOICES) +
                f" -> {label(best)}")
    return rows, report


def write_table(rows, path: str, note: str = ""):
    """rows: (form, cells, tile) or (form, cells, tile, fuse)."""
    with open(path, "w") as f:
        f.write("# step-kernel tile and stages per launch per arithmetic form and grid size,"
                " written by mhdglm.autotune\n")
        f.write(f"# {note}\n# form max_cells tile fuse   (form: 0 fp64, 1 fp32, 2 mixed,"
                " 3 mixed_kahan, 4 fp64_fma, 5 ref_mixed; tile: 0 16x16, 1 32x8, 2 16x8,"
                " 3 16x15, 4 32x7, 5 16x7; fuse: stages per launch)\n")
        by_form = {}
        for row in rows:
            form, cells, tile = row[:3]
            fuse = row[3] if len(row) > 3 else 1
            by_form.setdefault(form, []).append((cells, tile, fuse))
        for form in sorted(by_form):
            entries = sorted(by_form[form])
            for k, (cells, tile, fuse) in enumerate(entries):
                # The largest tuned size also covers anything bigger.
                limit = cells if k + 1 < len(entries) else 2**62
                f.write(f"{form} {limit} {tile} {fuse}\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--forms", default="fp64,fp64_fma,mixed,mixed_kahan,fp32,ref_mixed")
    ap.add_argument("--sizes", default="256,512,1024,2048,4096")
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    forms = [f for f in args.forms.split(",") if f]
    sizes = [int(x) for x in args.sizes.split(",") if x]
    rows, _ = tune(forms, sizes, log=lambda m: print(m, file=sys.stderr, flush=True))
    note = f"tuned {datetime.datetime.now().isoformat(timespec='seconds')} on the current device"
    write_table(rows, args.out, note)
    print(open(args.out).read())
    return 0


if __name__ == "__main__":
    sys.exit(main())
