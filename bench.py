"""Benchmark: particle-updates/s of the Solve hot path on one MI355X.

Workload (default): 3-D weakly-compressible dam break, tank 3.2 x 1.0 x 1.0 m, water column
1.2 m long, 0.6 m high, 1.0 m wide, dx = 0.01 m, three wall layers, 1,079,268 particles
(720,000 fluid), cubic spline h = 1.3 dx, Tait EOS c0 = 62.64 m/s, Morris viscosity 1e-3,
artificial viscosity alpha 0.05, Leapfrog, FP64.  Inputs are synthetic (generated lattice with
seeded jitter), resident in HBM before the timed region.

Prints one JSON line.  `value` is whole-job particle-updates/s (particles x steps / wall time
of exactly --steps steps, synchronised on both sides).  The state is first advanced, untimed, to
the simulated time --t-start (default 0.1 s, the surge well under way), so that the timed steps
see the developed flow and rebuild the neighbour lists at their running rate; `list_builds`
counts the timed steps that sorted the particles and rebuilt the lists.  `roofline` is for the dominant kernel
(the pair-force gather): its algorithmic bytes per launch over its mean launch duration from HIP
events recorded on the solver's stream during the timed region.  `cpu_baseline` times the CPU
oracle (an OpenMP restatement of the reference algorithm) on the host cores on a short sample
of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s: FP64 vs FP32-mixed vs surrogate, 1 GPU, ≤1% front error"
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(precision, dx):
    """HBM bytes per force pass from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of
    this workload (copy-calibrated; written by python -m persiansph_amd.profiling traffic)."""
    if dx != 0.01:
        return None
    path = os.path.join(PROFILES, f"traffic_force_pass_{precision}_dambreak3d_1M.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_pass")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cfg, p, dt, steps, threads):
    """Times the oracle (kind "port") on `steps` steps after one untimed step."""
    from oracle.pyoracle import Oracle
    t0 = time.perf_counter()
    o = Oracle(cfg, p, dt, nthreads=threads)
    log(f"[cpu] oracle setup {time.perf_counter()-t0:.1f}s, {threads} threads")
    o.run(1)
    t0 = time.perf_counter()
    st = o.run(steps)
    el = time.perf_counter() - t0
    n = len(p)
    o.close()
    return {
        "value": n*st.steps/el,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{st.steps} Leapfrog steps of the same {n}-particle workload after 1 warm-up step, "
                  f"{el:.1f} s wall ({el/st.steps*1e3:.0f} ms/step), FP64 OpenMP oracle with the "
                  f"reference's algorithm (AoS particles, cell linked list, half-stencil pair lists, "
                  f"per-particle locks)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--t-start", type=float, default=0.1,
                    help="simulated time (s) to advance, untimed, before the warm-up: the timed steps "
                         "then run on the developed flow, where the neighbour lists are rebuilt at "
                         "their running rate")
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events in the timed region")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a plain device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe*1024*1024//4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4*n, "copy_probe_bytes_written": 4*n, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    from persiansph_amd import FP32_MIXED, FP64, Solver, cases

    prec = FP64 if args.precision == "fp64" else FP32_MIXED
    t0 = time.perf_counter()
    cfg, p, dt = cases.dam_break_3d(dx=args.dx, scheme=1, precision=prec)
    n = len(p)
    nfree = int((p["IsFree"] == 1).sum())
    log(f"[bench] generated {n} particles ({nfree} free) in {time.perf_counter()-t0:.1f}s, dt={dt:.4e}")
    t0 = time.perf_counter()
    s = Solver(cfg, p, dt)
    log(f"[bench] upload {time.perf_counter()-t0:.1f}s grid={s.grid_dict()['cell_no']}")
    adv_steps = adv_builds = 0
    if args.t_start > 0.0:
        t0 = time.perf_counter()
        while True:
            st = s.run(4000, tf=args.t_start)
            adv_steps += st.steps
            adv_builds += st.list_builds
            log(f"[bench] advancing: t={st.time:.4e} after {adv_steps} steps ({adv_builds} list builds), "
                f"{time.perf_counter()-t0:.1f}s")
            if st.stop_reason == 1 or st.steps == 0:
                break
    st = s.run(args.warmup)
    log(f"[bench] warmup {st.steps} steps, t={st.time:.4e}")
    t_timed = st.time
    if not args.no_timing:
        s.set_timing(True)
    t0 = time.perf_counter()
    st = s.run(args.steps)     # returns after the stream is synchronised
    el = time.perf_counter() - t0
    assert st.steps == args.steps, st.steps
    builds = st.list_builds
    ms = el/args.steps*1e3
    ups = n*args.steps/el
    log(f"[bench] {args.steps} steps in {el:.3f}s: {ms:.3f} ms/step, {ups:.4e} particle-updates/s, "
        f"pairs/step={st.pairs}, list builds={builds}")
    stats = s.kernel_stats()
    total_b, force_b = s.traffic_model()
    roof = None
    if not args.no_timing and stats.get("force", (0, 0))[1] > 0:
        fms = stats["force"][0]
        ach = force_b/(fms*1e-3)/1e9
        tb = measured_traffic(args.precision, args.dx)
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach/HBM_PEAK_GBS, 5), "traffic": int(tb) if tb else None,
                "kernel": "force pass: k_nlist (list builds) + k_force11_nl (FP64) / k_force11_f32_nl "
                          "(FP32) + overflow + fold", "kernel_ms": round(fms, 5),
                "algorithmic_bytes_per_launch": int(force_b)}
        log("[bench] kernel means (ms): " + ", ".join(f"{k}={v[0]:.4f}" for k, v in stats.items() if v[1]))
    pair_rate = st.pairs*args.steps/el
    # one visualisation output of the current state: the fields WriteXDMF writes, packed to FP32
    # on the GPU and copied to pinned host memory, against the full FP64 state download
    s.output_begin()                 # the first output allocates its device and pinned buffers
    s.output_end(copy=False)
    t0 = time.perf_counter()
    s.output_begin()
    s.output_end(copy=False)
    packed_ms = (time.perf_counter() - t0)*1e3
    t0 = time.perf_counter()
    s.particles()
    download_ms = (time.perf_counter() - t0)*1e3
    log(f"[bench] output: packed FP32 {packed_ms:.2f} ms, full download {download_ms:.1f} ms")
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(cfg, p, dt, args.cpu_steps, threads)
        log(f"[cpu] {cpu['value']:.4e} particle-updates/s")
    s.close()
    out = {
        "metric": METRIC,
        "value": ups,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32-pairs/f64-state",
        "data": "synthetic",
        "config": {
            "workload": f"3-D dam break, tank 3.2x1.0x1.0 m, column 1.2x0.6x1.0 m, dx={args.dx} m, "
                        f"3 wall layers, {n} particles ({nfree} fluid), cubic spline h=1.3dx, Tait c0=62.64 m/s, "
                        f"Morris mu=1e-3, alpha=0.05, Leapfrog, {args.precision.upper()}",
            "particles": n, "free": nfree, "dt0": dt, "precision": args.precision,
            "t_start": args.t_start, "t_timed_from": t_timed,
        },
        "list_builds": builds,
        "list_builds_before": adv_builds,
        "steps_before": adv_steps,
        "pair_interactions_per_s": pair_rate,
        "pairs_per_step": st.pairs,
        "algorithmic_bytes_per_step": total_b,
        "roofline": roof,
        "cpu_baseline": cpu,
        "output": {"packed_f32_ms": round(packed_ms, 3), "full_download_ms": round(download_ms, 2),
                   "packed_bytes": 104*n},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
