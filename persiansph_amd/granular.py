"""Granular column collapse, FP64 and FP32-mixed side by side on identical inputs.

    python -m persiansph_amd.granular --dx 0.004 --t-end 0.5 --samples 10 --out granular.json

    python -m persiansph_amd.granular --dx 0.004 --measure 0.05 500

--measure T STEPS advances each tier to t = T, then times STEPS steps with HIP events on the
solver's stream around every stage (sort, wall, force, integrate) and prints one JSON line per
tier: ms/step, particle-updates/s, the stage means, the algorithmic bytes per step and per
force pass (psph_traffic_model, soil tensors included) and the force pass's achieved GB/s.

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


HBM_PEAK_GBS = 8000.0


def measure(cfg, p, dt, t_start, steps, label):
    """Stage times and the force pass's roofline figures of `steps` steps from t = t_start."""
    s = Solver(cfg, p, dt)
    n, nfree = len(p), int((p["IsFree"] == 1).sum())
    if t_start > 0.0:
        s.run(10**9, tf=t_start)
    s.run(5)
    s.set_timing(True)
    t0 = time.perf_counter()
    st = s.run(steps)
    el = time.perf_counter() - t0
    stats = s.kernel_stats()
    s.set_timing(False)
    total_b, force_b = s.traffic_model()
    fms = stats.get("force", (0.0, 0))[0]
    s.close()
    res = {"tier": label, "t": st.time, "steps": st.steps, "ms_per_step": 1e3*el/st.steps,
           "particle_updates_per_s": n*st.steps/el, "free_particle_updates_per_s": nfree*st.steps/el,
           "pairs_per_step": st.pairs, "stage_means_ms": {k: round(v[0], 5) for k, v in stats.items() if v[1]},
           "algorithmic_bytes_per_step": total_b, "force_bytes_per_launch": force_b}
    if fms > 0:
        ach = force_b/(fms*1e-3)/1e9
        res["force_roofline"] = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(ach/HBM_PEAK_GBS, 5), "kernel_ms": round(fms, 5)}
    log(json.dumps(res))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dx", type=float, default=0.004)
    ap.add_argument("--t-end", type=float, default=0.5)
    ap.add_argument("--samples", type=int, default=10)
    ap.add_argument("--out", default="")
    ap.add_argument("--measure", nargs=2, metavar=("T", "STEPS"))
    args = ap.parse_args(argv)
    dim = 3
    cfg, p, dt = cases.granular_column(dim=dim, dx=args.dx, scheme=1, precision=FP64)
    if args.measure:
        cfg32 = type(cfg).from_buffer_copy(cfg)
        cfg32.precision = FP32_MIXED
        for label, c in (("fp64", cfg), ("fp32", cfg32)):
            print(json.dumps(measure(c, p, dt, float(args.measure[0]), int(args.measure[1]), label)), flush=True)
        return
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
