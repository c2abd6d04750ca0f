"""Autotuner of the stage kernels' tile shape.

    python -m mhdglm.autotune [--forms fp64,mixed] [--sizes 256,1024,4096] [--out FILE]

For every arithmetic form and grid size it times each tile shape on the
current device (HIP events around both stage kernels, after a warm-up) and
writes the table the library reads when a caller leaves the tile at -1:
lines "form max_cells tile", sizes ascending, so the best tile for the
smallest tuned size at or above a grid's cell count is used.
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys

from . import PKG_ROOT, PRECISIONS, Solver, SolverConfig

TILES = {0: "16x16", 1: "32x8", 2: "16x8", 3: "16x15", 4: "32x7", 5: "16x7"}
DEFAULT_OUT = os.path.join(PKG_ROOT, "autotune.txt")


def time_tile(form: str, n: int, tile: int, steps: int = 10) -> float:
    """Average ms of one step's two stage kernels (OT n x n)."""
    s = Solver(SolverConfig(nx=n, ny=n, precision=PRECISIONS[form], tile=tile))
    s.init_problem("ot")
    s.advance(0.5, 3)
    ms = s.time_kernels(0.5, steps)
    s.close()
    return ms[0] + ms[1]


def tune(forms, sizes, log=print):
    rows, report = [], []
    for form in forms:
        for n in sorted(sizes):
            times = {t: time_tile(form, n, t) for t in TILES}
            best = min(times, key=times.get)
            rows.append((PRECISIONS[form], n * n, best))
            report.append((form, n, best, times))
            log(f"autotune {form} {n}^2: " +
                ", ".join(f"{TILES[t]} {times[t]:.4f} ms" for t in TILES) + f" -> {TILES[best]}")
    return rows, report


def write_table(rows, path: str, note: str = ""):
    with open(path, "w") as f:
        f.write("# stage-kernel tile per arithmetic form and grid size, written by mhdglm.autotune\n")
        f.write(f"# {note}\n# form max_cells tile   (form: 0 fp64, 1 fp32, 2 mixed, 3 mixed_kahan,"
                " 4 fp64_fma; tile: 0 16x16, 1 32x8, 2 16x8, 3 16x15, 4 32x7, 5 16x7)\n")
        by_form = {}
        for form, cells, tile in rows:
            by_form.setdefault(form, []).append((cells, tile))
        for form in sorted(by_form):
            entries = sorted(by_form[form])
            for k, (cells, tile) in enumerate(entries):
                # The largest tuned size also covers anything bigger.
                limit = cells if k + 1 < len(entries) else 2**62
                f.write(f"{form} {limit} {tile}\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--forms", default="fp64,fp64_fma,mixed,mixed_kahan,fp32")
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
