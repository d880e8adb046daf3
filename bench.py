"""Benchmark: particle-updates/s of the Solve hot path on one MI355X, FP64 vs FP32-mixed vs the
learned surrogate.

Workload: 3-D weakly-compressible dam break, tank 3.2 x 1.0 x 1.0 m, water column 1.2 m long,
0.6 m high, 1.0 m wide, dx = 0.01 m, three wall layers, 1,079,268 particles (720,000 fluid),
cubic spline h = 1.3 dx, Tait EOS c0 = 62.64 m/s, Morris viscosity 1e-3, artificial viscosity
alpha 0.05, Leapfrog.  Inputs are synthetic (generated lattice with seeded jitter), resident in
HBM before any timed region.

Prints one JSON line.  `value` is the FP64 tier's whole-job particle-updates/s (particles x steps
/ wall time of exactly --steps steps, synchronised on both sides).  Each tier is first advanced,
untimed, to the simulated time --t-start (default 0.1 s, the surge well under way), so that the
timed steps see the developed flow.  Beside it the line carries:
  * `ms_per_step_amortised`: a further --amort-steps steps timed as one window, so that the
    neighbour-list rebuilds enter at their running rate (the --steps window may hold none);
  * `free_particle_updates_per_s`: the same rate over the fluid particles only;
  * `fp32_mixed`: the FP32-mixed tier on the same inputs in the same process, with its own
    roofline (FP32 byte model) and its speed-up over FP64;
  * `front`: the surge front of both tiers at --t-start and their relative difference;
  * `surrogate`: one call of the 10-layer message-passing network on the FP64 state (graph build,
    forward, state update), as particle-updates/s counting the k = 10 steps it replaces;
  * `roofline`: the dominant kernel (the FP64 pair-force gather), its algorithmic bytes per
    launch over its mean launch duration from HIP events on the solver's stream in the timed
    region; `traffic` is the copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per launch;
  * `cpu_baseline`: the CPU oracle (an OpenMP restatement of the reference algorithm) on the
    host cores, timed over steps after a warm-up on the same workload.
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
BF16_PEAK_TFLOPS = 2500.0
PROFILES = os.path.join(ROOT, "profiles")
COLUMN_LENGTH = 1.2


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


def force_pass_name(precision):
    """The kernels of the timed force pass of this workload."""
    if not os.environ.get("PSPH_NO_TILES"):
        return ("force pass: k_nlist + k_tile_table + k_tile_union (list builds) + " +
                ("k_force11_tile (FP64" if precision == "fp64" else "k_force32_tile (FP32-mixed") +
                ", LDS-staged tiles) + overflow pass")
    return ("force pass: k_nlist (list builds) + " +
            ("k_force11_nl (FP64)" if precision == "fp64" else "k_force11_f32_nl (FP32-mixed)") + " + overflow pass")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(cfg, p, dt, warm, steps, threads):
    """Times the oracle (kind "port") over `steps` steps after `warm` untimed steps."""
    from oracle.pyoracle import Oracle
    t0 = time.perf_counter()
    o = Oracle(cfg, p, dt, nthreads=threads)
    log(f"[cpu] oracle setup {time.perf_counter()-t0:.1f}s, {threads} threads")
    o.run(warm)
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
        "cpu_model": cpu_model(),
        "sample": f"steps {warm + 1}..{warm + st.steps} (after {warm} untimed) of the same {n}-particle "
                  f"workload from t = 0: {el:.1f} s wall ({el/st.steps*1e3:.0f} ms/step) on {threads} "
                  f"OpenMP threads of a {cpu_model()}; FP64 oracle with the reference's algorithm (AoS "
                  f"particles, cell linked list, half-stencil pair lists, per-particle locks)",
    }


def advance(s, t_start, label):
    steps = builds = 0
    if t_start <= 0.0:
        return steps, builds, 0.0
    t0 = time.perf_counter()
    while True:
        st = s.run(4000, tf=t_start)
        steps += st.steps
        builds += st.list_builds
        log(f"[{label}] advancing: t={st.time:.4e} after {steps} steps ({builds} list builds), "
            f"{time.perf_counter()-t0:.1f}s")
        if st.stop_reason == 1 or st.steps == 0:
            return steps, builds, st.time


def time_tier(s, args, n, nfree, label, precision):
    """Warm-up, the timed window of --steps steps (HIP events around the force pass, or every
    kernel with --all-kernel-events), then the amortisation window of --amort-steps steps (no
    events)."""
    st = s.run(args.warmup)
    log(f"[{label}] warmup {st.steps} steps, t={st.time:.4e}")
    t_timed = st.time
    if not args.no_timing:
        s.set_timing(True, force_only=not args.all_kernel_events)
    t0 = time.perf_counter()
    st = s.run(args.steps)     # returns after the stream is synchronised
    el = time.perf_counter() - t0
    assert st.steps == args.steps, st.steps
    stats = s.kernel_stats()
    s.set_timing(False)
    ms = el/args.steps*1e3
    res = {"value": n*args.steps/el, "ms_per_step": ms, "free_particle_updates_per_s": nfree*args.steps/el,
           "list_builds": st.list_builds, "pairs_per_step": st.pairs,
           "pair_interactions_per_s": st.pairs*args.steps/el, "t_timed_from": t_timed}
    log(f"[{label}] {args.steps} steps in {el:.3f}s: {ms:.3f} ms/step, {res['value']:.4e} particle-updates/s, "
        f"pairs/step={st.pairs}, list builds={st.list_builds}")
    total_b, force_b = s.traffic_model()
    res["algorithmic_bytes_per_step"] = total_b
    res["roofline"] = None
    if not args.no_timing and stats.get("force", (0, 0))[1] > 0:
        fms = stats["force"][0]
        ach = force_b/(fms*1e-3)/1e9
        tb = measured_traffic(precision, args.dx)
        res["roofline"] = {
            "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach/HBM_PEAK_GBS, 5), "traffic": int(tb) if tb else None,
            "kernel": force_pass_name(precision), "kernel_ms": round(fms, 5),
            "algorithmic_bytes_per_launch": int(force_b)}
        res["kernel_means_ms"] = {k: round(v[0], 5) for k, v in stats.items() if v[1]}
        log(f"[{label}] kernel means (ms): " + ", ".join(f"{k}={v[0]:.4f}" for k, v in stats.items() if v[1]))
    if args.amort_steps > 0:
        t0 = time.perf_counter()
        sa = s.run(args.amort_steps)
        ea = time.perf_counter() - t0
        res["ms_per_step_amortised"] = ea/sa.steps*1e3
        res["amortised_window"] = {"steps": sa.steps, "list_builds": sa.list_builds,
                                   "value": n*sa.steps/ea}
        log(f"[{label}] amortisation window: {sa.steps} steps, {sa.list_builds} list builds, "
            f"{res['ms_per_step_amortised']:.3f} ms/step")
    return res


def surrogate_leg(q, dx, layers, k, dt):
    """One call of the surrogate network (same architecture as the configs trained in 2-D, here
    in 3-D) on the state q: radius graph, 10 message-passing layers, state update.  Its cost
    does not depend on the weight values (here a seeded initialisation)."""
    import math
    import ctypes
    from persiansph_amd import analysis, gnn
    G = 9.81
    T = k*dt
    net = gnn.Network(3, layers, 2.0*1.3*dx, rho0=1000.0, vel_scale=math.sqrt(2.0*G*2.0), rho_scale=0.01,
                      out_dx=0.5*G*T*T, out_dv=G*T, out_drho=1.0, drift_t=T,
                      deg_scale=float(analysis.bulk_neighbours(3)), train=False, seed=1)
    x, v, rho, fr = q["x"], q["v"], q["Density"], q["IsFree"]
    net.load_state(x, v, rho, fr)
    net.advance()                                           # warm-up call
    ne = ctypes.c_int64()
    gnn._check(gnn._lib().psph_gnn_graph(net._h, ctypes.byref(ne), None, None, 0))
    t0 = time.perf_counter()
    net.load_state(x, v, rho, fr)
    net.advance()
    el = time.perf_counter() - t0
    net.close()
    n, H = len(x), gnn.HIDDEN
    e = ne.value
    flops = 2.0*(e*((32 + H)*H + layers*((3*H + H)*H)) + n*((32 + H)*H + layers*((2*H + H)*H) + 2*H*H))
    return {"value": n*k/el, "unit": "particle-updates/s", "s_per_call": el, "k": k, "layers": layers,
            "edges": e, "tflops": flops/el/1e12, "bf16_peak_frac": flops/el/1e12/BF16_PEAK_TFLOPS,
            "note": "one call counts the k solver steps it replaces; weights seeded (the cost does not "
                    "depend on them); accuracy is judged per window by the plain solver (revert engine)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--t-start", type=float, default=0.1,
                    help="simulated time (s) to advance, untimed, before the warm-up: the timed steps "
                         "then run on the developed flow, where the neighbour lists are rebuilt at "
                         "their running rate")
    ap.add_argument("--amort-steps", type=int, default=1000,
                    help="steps of the amortisation window after the timed steps (list rebuilds at "
                         "their running rate)")
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-mixed tier")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate call")
    ap.add_argument("--surrogate-layers", type=int, default=10)
    ap.add_argument("--cpu-warm", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel events in the timed region")
    ap.add_argument("--all-kernel-events", action="store_true",
                    help="HIP events around every kernel of the timed steps (default: the force pass only)")
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

    from persiansph_amd import FP32_MIXED, FP64, Solver, analysis, cases

    t0 = time.perf_counter()
    cfg, p, dt = cases.dam_break_3d(dx=args.dx, scheme=1, precision=FP64)
    n = len(p)
    nfree = int((p["IsFree"] == 1).sum())
    log(f"[bench] generated {n} particles ({nfree} free) in {time.perf_counter()-t0:.1f}s, dt={dt:.4e}")

    # ---- FP64 tier (the headline value) ----
    t0 = time.perf_counter()
    s64 = Solver(cfg, p, dt)
    log(f"[fp64] upload {time.perf_counter()-t0:.1f}s grid={s64.grid_dict()['cell_no']}")
    adv_steps, adv_builds, t_adv = advance(s64, args.t_start, "fp64")
    q64 = s64.particles()
    front64 = analysis.front_position(q64, 3)
    r64 = time_tier(s64, args, n, nfree, "fp64", "fp64")

    # one visualisation output of the current state: the fields WriteXDMF writes, packed to FP32
    # on the GPU and copied to pinned host memory, against the full FP64 state download
    s64.output_begin()               # the first output allocates its device and pinned buffers
    s64.output_end(copy=False)
    t0 = time.perf_counter()
    s64.output_begin()
    s64.output_end(copy=False)
    packed_ms = (time.perf_counter() - t0)*1e3
    t0 = time.perf_counter()
    s64.particles()
    download_ms = (time.perf_counter() - t0)*1e3
    log(f"[bench] output: packed FP32 {packed_ms:.2f} ms, full download {download_ms:.1f} ms")
    s64.close()

    # ---- FP32-mixed tier on the same inputs ----
    r32 = None
    front = {"t": t_adv, "fp64": front64}
    if not args.no_fp32:
        cfg32 = type(cfg).from_buffer_copy(cfg)
        cfg32.precision = FP32_MIXED
        s32 = Solver(cfg32, p, dt)
        adv32, builds32, t32 = advance(s32, args.t_start, "fp32")
        front32 = analysis.front_position(s32.particles(), 3)
        r32 = time_tier(s32, args, n, nfree, "fp32", "fp32")
        s32.close()
        r32["steps_before"] = adv32
        r32["list_builds_before"] = builds32
        r32["speedup_vs_fp64"] = r64["ms_per_step"]/r32["ms_per_step"]
        if "ms_per_step_amortised" in r32 and "ms_per_step_amortised" in r64:
            r32["speedup_vs_fp64_amortised"] = r64["ms_per_step_amortised"]/r32["ms_per_step_amortised"]
        front.update({"fp32_mixed": front32, "t_fp32_mixed": t32,
                      "rel_err": abs(front32 - front64)/front64,
                      "rel_err_of_travel": abs(front32 - front64)/max(front64 - COLUMN_LENGTH, 1e-9)})
        log(f"[bench] front at t={t_adv:.4f}: fp64 {front64:.5f} m, fp32-mixed {front32:.5f} m "
            f"(rel. err {front['rel_err']:.2e})")

    # ---- surrogate: one network call on the FP64 state ----
    sur = None
    if not args.no_surrogate:
        sur = surrogate_leg(q64, args.dx, args.surrogate_layers, 10, dt)
        log(f"[surrogate] {sur['s_per_call']:.3f} s/call, {sur['value']:.4e} particle-updates/s, "
            f"{sur['edges']} edges, {sur['tflops']:.1f} TFLOP/s")
    del q64

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or host_threads()
        cpu = cpu_baseline(cfg, p, dt, args.cpu_warm, args.cpu_steps, threads)
        log(f"[cpu] {cpu['value']:.4e} particle-updates/s")
    out = {
        "metric": METRIC,
        "value": r64["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3-D dam break, tank 3.2x1.0x1.0 m, column 1.2x0.6x1.0 m, dx={args.dx} m, "
                        f"3 wall layers, {n} particles ({nfree} fluid), cubic spline h=1.3dx, Tait c0=62.64 m/s, "
                        f"Morris mu=1e-3, alpha=0.05, Leapfrog; FP64 headline, FP32-mixed and surrogate beside it",
            "particles": n, "free": nfree, "dt0": dt, "t_start": args.t_start,
            "t_timed_from": r64["t_timed_from"],
        },
        "ms_per_step_amortised": r64.get("ms_per_step_amortised"),
        "amortised_window": r64.get("amortised_window"),
        "free_particle_updates_per_s": r64["free_particle_updates_per_s"],
        "list_builds": r64["list_builds"],
        "list_builds_before": adv_builds,
        "steps_before": adv_steps,
        "pair_interactions_per_s": r64["pair_interactions_per_s"],
        "pairs_per_step": r64["pairs_per_step"],
        "algorithmic_bytes_per_step": r64["algorithmic_bytes_per_step"],
        "kernel_means_ms": r64.get("kernel_means_ms"),
        "roofline": r64["roofline"],
        "fp32_mixed": r32,
        "front": front,
        "surrogate": sur,
        "cpu_baseline": cpu,
        "output": {"packed_f32_ms": round(packed_ms, 3), "full_download_ms": round(download_ms, 2),
                   "packed_bytes": 104*n},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
