"""Granular column collapse, FP64 and FP32-mixed side by side on identical inputs.

    python -m persiansph_amd.granular --dx 0.004 --t-end 0.5 --samples 10 --out granular.json

Column 0.2 x 0.4 (high) x 0.2 m of Drucker-Prager soil (phi 30 deg, c 0, psi 0, E 5 MPa,
nu 0.3, rho 2650, Fail 3) on a rigid floor, cubic spline h = 1.5 dx, artificial stress on.
The FP32-mixed tier sums the pair terms (acceleration, continuity, strain and rotation
rates) in FP32 and keeps the stress update and return mapping in FP64.  At each sample time
both runs stop (output time of Solve) and the runout (largest x of a free particle) is
compared; the report also gives the wall time per step of each tier.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import FP32_MIXED, FP64, Solver, cases


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dx", type=float, default=0.004)
    ap.add_argument("--t-end", type=float, default=0.5)
    ap.add_argument("--samples", type=int, default=10)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    dim = 3
    cfg, p, dt = cases.granular_column(dim=dim, dx=args.dx, scheme=1, precision=FP64)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = FP32_MIXED
    n0 = len(p)
    x_mid = 0.1
    r0 = cases.runout(p)
    log(f"[granular] {n0} particles ({int((p['IsFree'] == 1).sum())} soil), dt0 {dt:.3e}")
    solvers = {"fp64": Solver(cfg, p, dt), "fp32": Solver(cfg32, p, dt)}
    wall = {k: 0.0 for k in solvers}
    steps = {k: 0 for k in solvers}
    rows = []
    for k in range(1, args.samples + 1):
        tk = args.t_end*k/args.samples
        got = {}
        for name, s in solvers.items():
            t0 = time.perf_counter()
            st = s.run(10**9, tout=tk)
            wall[name] += time.perf_counter() - t0
            steps[name] += st.steps
            q = s.particles()
            got[name] = (st, cases.runout(q), len(q))
        (s64, ro64, n64), (s32, ro32, n32) = got["fp64"], got["fp32"]
        row = {"t_fp64": s64.time, "t_fp32": s32.time, "steps_fp64": steps["fp64"], "steps_fp32": steps["fp32"],
               "runout_fp64": ro64, "runout_fp32": ro32,
               "runout_rel_err": abs(ro32 - ro64)/max(ro64 - x_mid, 1e-12),
               "spread_fp64": ro64 - r0, "n_fp64": n64, "n_fp32": n32,
               "ms_per_step_fp64": 1e3*wall["fp64"]/max(steps["fp64"], 1),
               "ms_per_step_fp32": 1e3*wall["fp32"]/max(steps["fp32"], 1)}
        rows.append(row)
        log(json.dumps(row))
    out = {"workload": f"3-D granular column 0.2x0.4x0.2 m, dx={args.dx}, {n0} particles, Leapfrog, to t={args.t_end}",
           "samples": rows,
           "final_runout_rel_err": rows[-1]["runout_rel_err"],
           "speedup_fp32_over_fp64": (wall["fp64"]/max(steps["fp64"], 1))/(wall["fp32"]/max(steps["fp32"], 1))}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "samples"}))


if __name__ == "__main__":
    main()
