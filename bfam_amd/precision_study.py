"""Precision study on the manufactured elastic solution.

The configuration is a periodic unit cube of n^3 hexes at LGL order N (32^3,
N=7 by default: 150,994,944 DOF). Material is homogeneous: rho = 1, cp = sqrt(3),
cs = 1. The exact solution is a product of sines and cosines in x, y, z and t,
driven by the matching forcing. It runs for `steps` LSRK4(5) steps (10^4 by
default) at CFL 0.5.

Forms compared:
  fp64         FP64 state, rates and arithmetic
  fp32_rhs     FP64 state, FP32 rates and arithmetic
  fp32         FP32 state, rates and arithmetic
  fp32_bf16x3  as fp32, with the derivative contractions on the matrix
               cores in split-BF16x3 (two-term bf16 operands, three
               products, FP32 accumulation)

Every `every` steps the script prints one JSON line per form with:
  - the relative L2 error against the exact solution;
  - the relative energy-norm error against the exact solution;
  - the relative L2 difference to the FP64 run.
All of these are FP64 device reductions with compensated sums.

    python -m bfam_amd.precision_study [--n 32] [--order 7] [--steps 10000]
                                       [--out record.jsonl]

--out also writes the JSON lines to a file (the committed record of a run).
"""
import argparse
import json
import math
import sys
import time

import numpy as np

import bfam_amd as bx
import bfam_amd.workloads as W


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--forms", default="fp64,fp32_rhs,fp32,fp32_bf16x3")
    ap.add_argument("--out", default=None, help="also write the JSON lines to this file")
    a = ap.parse_args(argv)
    forms = a.forms.split(",")
    if "fp64" not in forms:
        forms = ["fp64"] + forms
    cfg = W.elastic3d_mms(n=a.n, N=a.order)
    m = cfg["mesh"]
    dof = 9 * m.K * m.Np
    q0 = W.mms_exact(m, cfg["A"], cfg["k"], cfg["omega"], 0.0)
    dev = bx.Device(0)
    subs = {}
    for f in forms:
        s = W.upload(dev, cfg, bx.PRECISIONS[f])
        s.set_fields(q0, 0.0, cfg["dt"])
        subs[f] = s
    del q0
    clock = {f: 0.0 for f in forms}
    done = 0
    out = open(a.out, "w") if a.out else None

    def emit(rec):
        line = json.dumps(rec)
        print(line, flush=True)
        if out:
            out.write(line + "\n")
            out.flush()

    emit({"config": f"elastic3d_mms {a.n}^3 N={a.order}", "dof": dof, "dt": cfg["dt"],
          "steps": a.steps, "forms": forms})
    while done < a.steps:
        k = min(a.every, a.steps - done)
        for f in forms:
            dev.sync()
            t0 = time.perf_counter()
            subs[f].step(k)
            dev.sync()
            clock[f] += time.perf_counter() - t0
        done += k
        for f in forms:
            s = subs[f]
            d2, n2 = s.mms_error(cfg["A"])
            ed, en = s.mms_energy_error(cfg["A"])
            rec = {"step": done, "form": f, "t": s.time()[0],
                   "rel_l2_vs_exact": math.sqrt(d2 / n2),
                   "rel_energy_vs_exact": math.sqrt(ed / en),
                   "ms_per_stage": 1e3 * clock[f] / (5 * done)}
            if f != "fp64":
                x2, y2 = s.l2_diff(subs["fp64"])
                rec["rel_l2_vs_fp64"] = math.sqrt(x2 / y2)
            emit(rec)
    for s in subs.values():
        s.close()
    dev.close()
    if out:
        out.close()


if __name__ == "__main__":
    sys.exit(main())
