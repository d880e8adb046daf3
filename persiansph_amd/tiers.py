"""Side-by-side run of the FP64 and FP32-mixed tiers on identical seeded inputs.

    python -m persiansph_amd.tiers --dx 0.01 --t-end 1.0 --samples 20 --out tiers.json

Both solvers stop at each sample time t_k (the first force pass with Time >= t_k, as an output
time of Solve would).  At each stop the script records:
  - the per-step error: one force pass of each tier on the same state (the FP64 run's), with
    the relative L2 error of the FP32-mixed accelerations and continuity rates;
  - the trajectory difference: density excess, acceleration and position of the two runs
    (these grow with the chaotic divergence of the two trajectories, not with a per-step error);
  - the surge front of each run and its relative difference;
  - the wall time per step of each tier.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import FP32_MIXED, FP64, Solver, analysis, cases


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--samples", type=int, default=20)
    ap.add_argument("--out", default="")
    args = ap.parse_args()

    cfg, p, dt = cases.dam_break_3d(dx=args.dx, scheme=1, precision=FP64)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = FP32_MIXED
    n0 = len(p)
    L0 = 1.2
    solvers = {"fp64": Solver(cfg, p, dt), "fp32": Solver(cfg32, p, dt)}
    del p
    wall = {k: 0.0 for k in solvers}
    steps = {k: 0 for k in solvers}
    rows = []
    for k in range(1, args.samples + 1):
        tk = args.t_end*k/args.samples
        states = {}
        for name, s in solvers.items():
            t0 = time.perf_counter()
            st = s.run(10**9, tout=tk)
            wall[name] += time.perf_counter() - t0
            steps[name] += st.steps
            q = s.particles()
            states[name] = (st, q)
        (s64, q64), (s32, q32) = states["fp64"], states["fp32"]
        common = np.intersect1d(q64["Index"], q32["Index"])
        i64 = np.searchsorted(q64["Index"], common)
        i32 = np.searchsorted(q32["Index"], common)
        fr = q64["IsFree"][i64] == 1
        rho_err = analysis.relative_error(q32["Density"][i32][fr] - 1000.0, q64["Density"][i64][fr] - 1000.0)
        a_err = analysis.relative_error(q32["a"][i32][fr], q64["a"][i64][fr])
        x_err = float(np.abs(q32["x"][i32][fr] - q64["x"][i64][fr]).max())
        # per-step error: both tiers' force pass on the FP64 run's state
        one = {}
        for name, c in (("fp64", cfg), ("fp32", cfg32)):
            s1 = Solver(c, q64, s64.deltat, time=s64.time)
            s1.run(1, tout=0.0)
            one[name] = s1.particles()
            s1.close()
        f1 = one["fp64"]["IsFree"] == 1
        step_a = analysis.relative_error(one["fp32"]["a"][f1], one["fp64"]["a"][f1])
        step_dd = analysis.relative_error(one["fp32"]["dDensity"][f1], one["fp64"]["dDensity"][f1])
        f64 = analysis.front_position(q64, 3)
        f32 = analysis.front_position(q32, 3)
        row = {
            "t_fp64": s64.time, "t_fp32": s32.time, "steps_fp64": steps["fp64"], "steps_fp32": steps["fp32"],
            "n_fp64": int(len(q64)), "n_fp32": int(len(q32)),
            "step_rel_err_accel": step_a, "step_rel_err_ddensity": step_dd,
            "rel_err_density_excess": rho_err, "rel_err_accel": a_err, "max_abs_dx": x_err,
            "front_fp64": f64, "front_fp32": f32,
            "front_rel_err": abs(f32 - f64)/f64, "front_rel_err_travel": abs(f32 - f64)/max(f64 - L0, 1e-9),
            "ms_per_step_fp64": 1e3*wall["fp64"]/max(steps["fp64"], 1),
            "ms_per_step_fp32": 1e3*wall["fp32"]/max(steps["fp32"], 1),
        }
        rows.append(row)
        log(json.dumps(row))
    out = {
        "workload": f"3-D dam break dx={args.dx}, {n0} particles, Leapfrog, to t={args.t_end}",
        "samples": rows,
        "speedup_fp32_over_fp64": (wall["fp64"]/steps["fp64"])/(wall["fp32"]/steps["fp32"]),
        "max_front_rel_err": max(r["front_rel_err"] for r in rows),
        "max_step_rel_err_accel": max(r["step_rel_err_accel"] for r in rows),
        "max_step_rel_err_ddensity": max(r["step_rel_err_ddensity"] for r in rows),
    }
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: v for k, v in out.items() if k != "samples"}))


if __name__ == "__main__":
    main()
