"""Writes the comparison fixtures of tests/golden/ from the reference's own
logs (text tables only; nothing of the reference is imported or run):

  refmixed_ot512.json -- the reference's mixed GPU build against its FP64 GPU
      build, Orszag-Tang 512^2 to t=0.5 (logs/mhd/raptor_scan/
      ot2d512_cmp_mixed_gpu.txt, step count from ot2d512_orig.log), with the
      A30 loop-wall medians of logs/mhd/perf_baseline/mixed_rr.csv;
  cmp_ot192.json -- Orszag-Tang 192^2 to t=0.5 against the FP64 CPU build:
      the native FP32 build (ot2d_cmp_nativefp32.txt) and the FP64 build with
      recovery, CFL, reconstruction, rhs and GLM regions truncated to FP32
      mantissas (ot2d_cmp_c_recovery+cfl+recon+rhs+glm.txt, the "keepFU"
      row of ot2d_table.txt), and the same build with one region at a time
      truncated (ot2d_cmp_{recovery,cfl,recon,rhs,glm}.txt).

    python tests/golden/make_cmp_fixtures.py [/root/reference]
"""
import csv
import json
import os
import re
import statistics
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_cmp(path):
    """Rows `var L1 L2 Linf relL2` and the divB lines of a comparison table."""
    rows, divb = {}, {}
    for line in open(path, encoding="utf-8"):
        m = re.match(r"\s*(\w+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$", line)
        if m and m.group(1) in ("rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "E"):
            rows[m.group(1)] = {"L1": float(m.group(2)), "L2": float(m.group(3)),
                                "Linf": float(m.group(4)), "relL2": float(m.group(5))}
            rows[m.group(1)]["text"] = [m.group(2), m.group(3), m.group(4), m.group(5)]
        m = re.match(r"#\s+(fp64|fp32):\s+L1=(\S+)\s+Linf=(\S+)", line)
        if m:
            divb[m.group(1)] = {"L1": m.group(2), "Linf": m.group(3)}
    return rows, divb


def steps_of(path):
    m = re.search(r"Done\. Steps: (\d+)", open(path, encoding="utf-8").read())
    return int(m.group(1))


def main(ref="/root/reference"):
    scan = os.path.join(ref, "logs/mhd/raptor_scan")
    rows, _ = parse_cmp(os.path.join(scan, "ot2d512_cmp_mixed_gpu.txt"))
    for r in rows.values():
        del r["text"]
    walls = list(csv.DictReader(open(os.path.join(ref, "logs/mhd/perf_baseline/mixed_rr.csv"))))
    med = {p: round(statistics.median(float(r["loop_wall_s"]) for r in walls if r["prec"] == p), 3)
           for p in ("fp64", "fp32", "mixed")}
    out = {
        "source": "logs/mhd/raptor_scan/ot2d512_cmp_mixed_gpu.txt",
        "what": "reference mixed GPU build (mhd2d_glm_gpu_mixed.cu) vs reference FP64 GPU build, "
                "Orszag-Tang 512x512, MUSCL+HLLD, CFL 0.25, to t=0.5 (5239 steps, "
                "raptor_scan/ot2d512_orig.log)",
        "n": 512, "rows": rows, "t_end": 0.5,
        "steps": steps_of(os.path.join(scan, "ot2d512_orig.log")),
        "a30_loop_wall_s_median": med, "a30_source": "logs/mhd/perf_baseline/mixed_rr.csv",
    }
    with open(os.path.join(HERE, "refmixed_ot512.json"), "w") as f:
        json.dump(out, f, indent=1)

    native, native_divb = parse_cmp(os.path.join(scan, "ot2d_cmp_nativefp32.txt"))
    keep, keep_divb = parse_cmp(os.path.join(scan, "ot2d_cmp_c_recovery+cfl+recon+rhs+glm.txt"))
    out = {
        "n": 192, "t_end": 0.5, "steps_fp64": steps_of(os.path.join(scan, "ot2d_orig.log")),
        "native_fp32": {"source": "logs/mhd/raptor_scan/ot2d_cmp_nativefp32.txt",
                        "what": "FP32 CPU build vs FP64 CPU build (.dat files)",
                        "rows": native, "divB": native_divb},
        "keepFU": {"source": "logs/mhd/raptor_scan/ot2d_cmp_c_recovery+cfl+recon+rhs+glm.txt",
                   "what": "FP64 build with the recovery, cfl, recon, rhs and glm regions "
                           "truncated to FP32 mantissas vs FP64 (the flux and update kept)",
                   "rows": keep, "divB": keep_divb},
        "regions": {},
    }
    for reg in ("recovery", "cfl", "recon", "rhs", "glm"):
        rows, divb = parse_cmp(os.path.join(scan, f"ot2d_cmp_{reg}.txt"))
        out["regions"][reg] = {
            "source": f"logs/mhd/raptor_scan/ot2d_cmp_{reg}.txt",
            "what": f"FP64 build with only the {reg} region truncated to FP32 mantissas vs FP64",
            "rows": rows, "divB": divb}
    with open(os.path.join(HERE, "cmp_ot192.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(*sys.argv[1:])
