"""Step rates of the boundary and material paths on one GPU, FP64.

    python -m persiansph_amd.bc_bench --steps 500 --out bc.json

Workloads (persiansph_amd.cases, scaled up from the parity cases):
  * periodic: the 1-Poiseuille.cpp channel (periodic x, quintic spline, Takeda viscosity,
    modified Verlet, body force) lengthened to nx x 48 particles;
  * inflow: the in/out-flow channel (InFlowBCFresh/InFlowBCLeave, cubic spline, modified
    Verlet) at dx = 1 mm, L x 0.12 m; the recycling halts the step loop whenever an inflow
    column passes InFlowLoc1 or a column leaves at the right, and runs on the host mirror;
  * soil_water: a water column on a soil bed (CalcForce11 + CalcForce2233 + CalcForce13 with
    Darcy seepage, SWIType 0) in a 0.6 x 0.4 m tank at dx.
Each runs `--warmup` untimed steps, then `--steps` timed steps of Solver.run (which returns
after the device has finished); the line per workload gives ms/step and particle-updates/s.
"""
import argparse
import json
import sys
import time

from . import Solver, cases


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _workloads(args):
    cfg, p, dt, body = cases.poiseuille_2d(nx=args.periodic_nx)
    yield "periodic", f"Poiseuille channel periodic in x, {args.periodic_nx} x 48 particles, dx 2.5e-5 m", cfg, p, dt, body
    cfg, p, dt = cases.channel_inflow_2d(dx=0.001, L=args.inflow_length)
    yield "inflow", f"in/out-flow channel {args.inflow_length} x 0.12 m, dx 1 mm, v_in 1 m/s", cfg, p, dt, None
    cfg, p, dt = cases.soil_water_2d(dx=args.soil_dx, swi_type=0)
    yield "soil_water", f"water column on a soil bed, tank 0.6 x 0.4 m, dx {args.soil_dx} m, SWIType 0", cfg, p, dt, None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--periodic-nx", type=int, default=10000)
    ap.add_argument("--inflow-length", type=float, default=1.0)
    ap.add_argument("--soil-dx", type=float, default=0.001)
    ap.add_argument("--only", default="")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    rows = []
    for name, desc, cfg, p, dt, body in _workloads(args):
        if args.only and name not in args.only.split(","):
            continue
        log(f"[bc_bench] {name}: {len(p)} particles, {int((p['IsFree'] == 1).sum())} free")
        s = Solver(cfg, p, dt)
        if body is not None:
            s.set_body_force(body)
        s.run(args.warmup)
        n0 = len(p)
        t0 = time.perf_counter()
        st = s.run(args.steps)
        el = time.perf_counter() - t0
        row = {"workload": name, "description": desc, "particles": int(st.n), "steps": int(st.steps),
               "ms_per_step": 1e3*el/max(st.steps, 1),
               "particle_updates_per_s": st.n*st.steps/el if el > 0 else 0.0,
               "pairs_per_step": int(st.pairs), "deleted": int(st.deleted), "sim_time": st.time}
        log(f"[bc_bench] {name}: {row['ms_per_step']:.3f} ms/step, n {n0} -> {st.n}")
        rows.append(row)
        s.close()
    out = json.dumps({"precision": "fp64", "rows": rows})
    print(out, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(out + "\n")


if __name__ == "__main__":
    main()
