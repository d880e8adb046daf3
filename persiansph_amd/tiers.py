"""The FP64 and FP32-mixed tiers on the 3-D dam break, to t = 1 s, on identical seeded inputs.

    python -m persiansph_amd.tiers --tier fp64 --t-end 1.0 --samples 20 --out fp64.jsonl
    python -m persiansph_amd.tiers --tier fp32 --t-end 1.0 --samples 20 --out fp32.jsonl
    python -m persiansph_amd.tiers --compare fp64.jsonl fp32.jsonl

A tier run stops at each sample time t_k (the first force pass with Time >= t_k, as an output
time of Solve would) and appends one JSON line to --out (flushed, so a run cut short keeps its
samples): the surge front, the plain maximum x of the fluid, the mean density excess, the
steps, list builds and wall time so far.  The FP64 run also measures the per-step error of the
FP32-mixed tier at each sample: one force pass of each tier on the FP64 run's state, with the
relative L2 error of the FP32-mixed accelerations and continuity rates.  --compare joins the
two runs' samples: the front difference relative to the front and to the distance travelled,
and the speed-up of FP32-mixed over FP64 in wall time per simulated second.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import FP32_MIXED, FP64, Solver, analysis, cases

COLUMN_LENGTH = 1.2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_tier(args):
    prec = FP64 if args.tier == "fp64" else FP32_MIXED
    cfg, p, dt = cases.dam_break_3d(dx=args.dx, scheme=1, precision=prec)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = FP32_MIXED
    cfg64 = type(cfg).from_buffer_copy(cfg)
    cfg64.precision = FP64
    n0 = len(p)
    s = Solver(cfg, p, dt)
    del p
    out = open(args.out, "a") if args.out else None
    wall = 0.0
    steps = builds = 0
    for k in range(1, args.samples + 1):
        tk = args.t_end*k/args.samples
        t0 = time.perf_counter()
        st = s.run(10**9, tout=tk)
        wall += time.perf_counter() - t0
        steps += st.steps
        builds += st.list_builds
        q = s.particles()
        fr = q["IsFree"] == 1
        row = {
            "tier": args.tier, "t": st.time, "t_sample": tk, "steps": steps, "list_builds": builds,
            "wall_s": wall, "ms_per_step": 1e3*wall/max(steps, 1), "n": int(len(q)),
            "front": analysis.front_position(q, 3), "max_x": float(q["x"][fr, 0].max()),
            "mean_density_excess": float(np.mean(q["Density"][fr]/1000.0 - 1.0)),
            "deltat": st.deltat,
        }
        if args.tier == "fp64" and not args.no_step_error:
            # per-step error of FP32-mixed: both tiers' force pass on this state
            one = {}
            for name, c in (("fp64", cfg64), ("fp32", cfg32)):
                s1 = Solver(c, q, st.deltat, time=st.time)
                s1.run(1, tout=0.0)
                one[name] = s1.particles()
                s1.close()
            f1 = one["fp64"]["IsFree"] == 1
            row["step_rel_err_accel"] = analysis.relative_error(one["fp32"]["a"][f1], one["fp64"]["a"][f1])
            row["step_rel_err_ddensity"] = analysis.relative_error(one["fp32"]["dDensity"][f1],
                                                                   one["fp64"]["dDensity"][f1])
        log(json.dumps(row))
        if out:
            out.write(json.dumps(row) + "\n")
            out.flush()
    s.close()
    summary = {"tier": args.tier, "workload": f"3-D dam break dx={args.dx}, {n0} particles, Leapfrog",
               "t_end": args.t_end, "steps": steps, "list_builds": builds, "wall_s": wall,
               "ms_per_step": 1e3*wall/max(steps, 1), "particle_updates_per_s": n0*steps/max(wall, 1e-9)}
    print(json.dumps(summary))


def compare(path64, path32):
    def load(p):
        with open(p) as f:
            return {round(r["t_sample"], 9): r for r in map(json.loads, f) if r}
    a, b = load(path64), load(path32)
    rows = []
    for t in sorted(set(a) & set(b)):
        r64, r32 = a[t], b[t]
        f64, f32 = r64["front"], r32["front"]
        rows.append({
            "t_sample": t, "front_fp64": f64, "front_fp32": f32,
            "front_rel_err": abs(f32 - f64)/f64,
            "front_rel_err_of_travel": abs(f32 - f64)/max(f64 - COLUMN_LENGTH, 1e-9),
            "steps_fp64": r64["steps"], "steps_fp32": r32["steps"],
            "wall_fp64": r64["wall_s"], "wall_fp32": r32["wall_s"],
            "step_rel_err_accel": r64.get("step_rel_err_accel"),
            "step_rel_err_ddensity": r64.get("step_rel_err_ddensity"),
        })
    last = rows[-1] if rows else None
    out = {"samples": rows}
    if last:
        out["speedup_fp32_over_fp64_wall"] = last["wall_fp64"]/last["wall_fp32"]
        out["max_front_rel_err"] = max(r["front_rel_err"] for r in rows)
    print(json.dumps(out, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tier", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--t-end", type=float, default=1.0)
    ap.add_argument("--samples", type=int, default=20)
    ap.add_argument("--no-step-error", action="store_true")
    ap.add_argument("--out", default="")
    ap.add_argument("--compare", nargs=2, metavar=("FP64_JSONL", "FP32_JSONL"))
    args = ap.parse_args()
    if args.compare:
        compare(*args.compare)
    else:
        run_tier(args)


if __name__ == "__main__":
    main()
