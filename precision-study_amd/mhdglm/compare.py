"""Comparator and md5 manifest for the solver's .dat output.

    python -m mhdglm.compare REF.dat TEST.dat [--floor FLOOR.dat]
    python -m mhdglm.compare --manifest DIR          # md5 of every .dat under DIR

The error table keeps the columns and definitions of the reference's
analysis/mhd/compare_precision.py:22-66 for the 13-column 2D layout
(x y rho vx vy vz p Bx By Bz E divB psi): with d = ref - test over N cells,
L1 = mean|d|, L2 = sqrt(mean d^2), Linf = max|d|, relL2 = L2 / RMS(ref).
It adds rows for |B| and psi, which that script does not compare, and the
same divB summary.  With --floor, a second run that differs from REF only
by valid rounding (e.g. FP64 with FMA contraction) gives the noise floor:
two more columns hold relL2(REF, FLOOR) and the ratio of TEST's relL2 to it
(chaotic flows amplify any rounding difference to that level).  The manifest mirrors scripts/make_manifest.sh:4-7
(`md5sum` of every .dat file), plus the md5 of the data rows alone
(`grep -v '^#' | md5sum`), which is the reference's bitwise gate.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

COLUMNS = ["x", "y", "rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "E", "divB", "psi"]
COMPARED = ["rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "E"]


def load_dat(path: str) -> np.ndarray:
    """Data rows of a 13-column .dat file (float64)."""
    a = np.loadtxt(path, comments="#", ndmin=2)
    if a.shape[1] != 13:
        raise ValueError(f"{path}: expected 13 columns, found {a.shape[1]}")
    return a


def fields_from_dat(a: np.ndarray) -> dict:
    """Named columns plus |B|."""
    f = {name: a[:, k] for k, name in enumerate(COLUMNS)}
    f["|B|"] = np.sqrt(f["Bx"] ** 2 + f["By"] ** 2 + f["Bz"] ** 2)
    return f


def fields_from_state(u: np.ndarray, gamma: float = 5.0 / 3.0) -> dict:
    """Named primitive fields of a conserved state (9, ...) in float64."""
    u = np.asarray(u, dtype=np.float64)
    rho = np.maximum(u[0], 1e-12)
    vx, vy, vz = u[1] / rho, u[2] / rho, u[3] / rho
    bx, by, bz = u[5], u[6], u[7]
    p = (gamma - 1.0) * (u[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * (bx * bx + by * by + bz * bz))
    return {"rho": rho, "vx": vx, "vy": vy, "vz": vz, "p": p, "Bx": bx, "By": by, "Bz": bz,
            "E": u[4], "psi": u[8], "|B|": np.sqrt(bx * bx + by * by + bz * bz)}


def norms(ref: np.ndarray, test: np.ndarray) -> dict:
    d = (np.asarray(ref, dtype=np.float64) - np.asarray(test, dtype=np.float64)).ravel()
    rms_ref = float(np.sqrt(np.mean(np.asarray(ref, dtype=np.float64) ** 2)))
    l2 = float(np.sqrt(np.mean(d * d)))
    return {"L1": float(np.mean(np.abs(d))), "L2": l2, "Linf": float(np.max(np.abs(d))),
            "relL2": (l2 / rms_ref) if rms_ref > 0 else 0.0}


def compare_fields(ref: dict, test: dict, names=None) -> dict:
    names = names or (COMPARED + ["|B|", "psi"])
    return {n: norms(ref[n], test[n]) for n in names}


def rel_l2_state(u_ref: np.ndarray, u_test: np.ndarray, names=("rho", "p", "|B|", "psi")) -> dict:
    """relL2 of named fields between two conserved states (9, ny, nx)."""
    fr, ft = fields_from_state(u_ref), fields_from_state(u_test)
    return {n: norms(fr[n], ft[n])["relL2"] for n in names}


def format_table(res: dict, ref_name: str, test_name: str, n: int, divb=None, floor=None,
                 floor_name=None) -> str:
    out = [f"# layout = 2D OT-GLM   N = {n}   cols = 13",
           f"# ref  = {ref_name}", f"# test = {test_name}"]
    head = f"{'var':>4} {'L1':>12} {'L2(RMS)':>12} {'Linf':>12} {'relL2':>10}"
    if floor is not None:
        out.append(f"# floor = {floor_name}")
        head += f" {'floor':>10} {'x floor':>8}"
    out += [head, "-" * len(head)]
    for name, r in res.items():
        line = f"{name:>4} {r['L1']:12.4e} {r['L2']:12.4e} {r['Linf']:12.4e} {r['relL2']:10.2e}"
        if floor is not None:
            fl = floor[name]["relL2"]
            ratio = r["relL2"] / fl if fl > 0 else float("inf") if r["relL2"] > 0 else 1.0
            line += f" {fl:10.2e} {ratio:8.2f}"
        out.append(line)
    if divb is not None:
        (l1a, lia), (l1b, lib) = divb
        out += ["-" * 56, "# divB control (col 12):",
                f"#   ref :  L1={l1a:.3e}  Linf={lia:.3e}", f"#   test:  L1={l1b:.3e}  Linf={lib:.3e}"]
    return "\n".join(out)


def compare_files(ref_path: str, test_path: str, floor_path: str = None) -> str:
    a, b = load_dat(ref_path), load_dat(test_path)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch {a.shape} vs {b.shape}")
    fa = fields_from_dat(a)
    res = compare_fields(fa, fields_from_dat(b))
    floor = None
    if floor_path:
        c = load_dat(floor_path)
        if c.shape != a.shape:
            raise ValueError(f"shape mismatch {a.shape} vs {c.shape}")
        floor = compare_fields(fa, fields_from_dat(c))
    divb = ((float(np.mean(np.abs(a[:, 11]))), float(np.max(np.abs(a[:, 11])))),
            (float(np.mean(np.abs(b[:, 11]))), float(np.max(np.abs(b[:, 11])))))
    return format_table(res, os.path.basename(ref_path), os.path.basename(test_path), a.shape[0],
                        divb, floor, os.path.basename(floor_path) if floor_path else None)


def data_md5(path: str) -> str:
    h = hashlib.md5()
    with open(path, "rb") as f:
        for line in f:
            if not line.startswith(b"#"):
                h.update(line)
    return h.hexdigest()


def file_md5(path: str) -> str:
    h = hashlib.md5()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def manifest(root: str) -> str:
    lines = []
    for d, _, files in sorted(os.walk(root)):
        for name in sorted(files):
            if name.endswith(".dat"):
                p = os.path.join(d, name)
                lines.append(f"{file_md5(p)}  {data_md5(p)}  {os.path.relpath(p, root)}")
    return "\n".join(lines)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) == 2 and argv[0] == "--manifest":
        print(manifest(argv[1]))
        return 0
    floor = None
    if len(argv) == 4 and argv[2] == "--floor":
        floor = argv[3]
        argv = argv[:2]
    if len(argv) != 2:
        print("Usage: python -m mhdglm.compare REF.dat TEST.dat [--floor FLOOR.dat] | "
              "--manifest DIR", file=sys.stderr)
        return 1
    print(compare_files(argv[0], argv[1], floor))
    return 0


if __name__ == "__main__":
    sys.exit(main())
