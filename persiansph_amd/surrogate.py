"""Surrogate-accelerated stepping: a message-passing network (persiansph_amd.gnn, HIP kernels)
that advances a particle state by k solver steps per call, trained on trajectories of the
plain solver, with the plain solver as its fallback and judge.

Pieces:
  dam_break_variant     seeded 2-D dam-break inputs with a given column aspect ratio
  trajectory            plain-solver frames every T = k dt0 of simulated time
  scales_from           input / output normalisation from trajectories
  train                 Adam training on (frame, frame + k steps) pairs
  predictors            `net_predictor` and the two predictors without a network: ballistic
                        (x += k dt v + g (k dt)^2 / 2, v += g k dt) and drift (x += k dt v)
  prediction_error      k-step error of a predictor on held-out frames
  HybridEngine          windows of M dt0 advanced by the network (M/k calls of T = k dt0),
                        each judged by the plain solver re-run over the whole window from its
                        start: the state the network reached must agree with the solver's
                        (surge front within `front_tol` of the solver's front position, and
                        position error at most `gain_tol` times that of drift stepping over the
                        window) and keep the mean density drift and the gain of kinetic +
                        gravitational energy below thresholds; otherwise the window is reverted
                        to the solver's state.
"""
import math
import time

import numpy as np

from . import analysis, cases
from .gnn import Network
from .solver import Solver, SolverError

G = 9.81
C0 = 10.0*math.sqrt(2.0*G*2.0)


def dam_break_variant(aspect, dx=0.01, area=2.0, tank=(4.0, 3.0)):
    """2-D dam break in the configs[0] tank with a water column of the given area (m^2) and
    height / width aspect ratio; c0 = 10 sqrt(2 g 2.0) for every variant."""
    W = math.sqrt(area/aspect)
    H = math.sqrt(area*aspect)
    return cases.dam_break(2, dx, tank, (W, H), cs=C0, scheme=1)


class Frame:
    __slots__ = ("x", "v", "rho", "free", "t")

    def __init__(self, x, v, rho, free, t):
        self.x, self.v, self.rho, self.free, self.t = x, v, rho, free, t

    @classmethod
    def from_particles(cls, q, t):
        return cls(np.array(q["x"]), np.array(q["v"]), np.array(q["Density"]),
                   np.array(q["IsFree"], dtype=np.int32), t)


def trajectory(cfg, p, dt, t_end, T):
    """Frames of the plain solver at times T, 2T, ... <= t_end (each taken at the first force
    pass with Time >= the sample time, as an output time of Solve; the solver's adaptive dt
    makes the step count per frame vary)."""
    s = Solver(cfg, p, dt)
    frames = [Frame.from_particles(s.particles(), 0.0)]
    n_frames = int(round(t_end/T))
    for m in range(1, n_frames + 1):
        st = s.run(10**9, tout=m*T)
        q = s.particles()
        if len(q) != len(frames[0].x):
            break                          # particles left the domain: end the trajectory
        frames.append(Frame.from_particles(q, st.time))
    s.close()
    return frames


def scales_from(trajs, k, dt, radius, rho0=1000.0, g=G, height=2.0):
    """Normalisation.  Inputs: velocity over the free-fall speed sqrt(2 g height), density
    deviation over 1 %.  Outputs over T = k dt: position change net of the drift T v over
    g T^2 / 2, velocity change over g T, density change over its RMS in the data (free
    particles).  Physical scales keep fast and resting particles on one footing."""
    T = k*dt
    drs = [(b.rho - a.rho)[a.free == 1] for fr in trajs for a, b in zip(fr[:-1], fr[1:])]
    rms = float(np.sqrt(np.mean(np.concatenate(drs)**2))) if drs else 0.0
    return dict(radius=radius, rho0=rho0, vel_scale=math.sqrt(2.0*g*height), rho_scale=0.01,
                out_dx=0.5*g*T*T, out_dv=g*T, out_drho=max(rms, 1e-6*rho0), drift_t=T,
                deg_scale=float(analysis.bulk_neighbours(2)))


def make_network(dim, layers, scales, train=True, seed=1):
    s = scales
    return Network(dim, layers, s["radius"], rho0=s["rho0"], vel_scale=s["vel_scale"],
                   rho_scale=s["rho_scale"], out_dx=s["out_dx"], out_dv=s["out_dv"],
                   out_drho=s["out_drho"], drift_t=s.get("drift_t", 0.0), deg_scale=s.get("deg_scale", 1.0),
                   train=train, seed=seed)


def train(net, trajs, steps, lr=1e-4, seed=0, log=None, log_every=50):
    """`steps` Adam steps, each on one (frame, next frame) pair drawn uniformly from trajs."""
    rng = np.random.default_rng(seed)
    pairs = [(i, j) for i, fr in enumerate(trajs) for j in range(len(fr) - 1)]
    losses = []
    for it in range(steps):
        i, j = pairs[rng.integers(len(pairs))]
        a, b = trajs[i][j], trajs[i][j + 1]
        net.load_state(a.x, a.v, a.rho, a.free)
        net.set_target_state(b.x, b.v, b.rho)
        losses.append(net.train_step(lr))
        if log and (it % log_every == 0 or it == steps - 1):
            log(f"[train] step {it} loss {losses[-1]:.4e}")
    return losses


# ---- predictors: frame -> (x, v, rho) after k steps -------------------------------------
def ballistic_predictor(k, dt, g=G):
    T = k*dt

    def f(a):
        x, v = a.x.copy(), a.v.copy()
        fr = a.free == 1
        x[fr, 0] += T*v[fr, 0]
        x[fr, 1] += T*v[fr, 1] - 0.5*g*T*T
        v[fr, 1] -= g*T
        return x, v, a.rho.copy()
    return f


def drift_predictor(k, dt):
    T = k*dt

    def f(a):
        x = a.x.copy()
        fr = a.free == 1
        x[fr] += T*a.v[fr]
        return x, a.v.copy(), a.rho.copy()
    return f


def net_predictor(net):
    def f(a):
        net.load_state(a.x, a.v, a.rho, a.free)
        return net.advance()
    return f


def prediction_error(trajs, predict, scales):
    """Mean over frame pairs of the normalised RMS error of the free particles' predicted
    change (position, velocity, density), each component scaled like the network output."""
    errs = []
    for fr in trajs:
        for a, b in zip(fr[:-1], fr[1:]):
            x, v, r = predict(a)
            f = a.free == 1
            ex = (x - b.x)[f, :2]/scales["out_dx"]     # the drift is exact, so this is net of it
            ev = (v - b.v)[f, :2]/scales["out_dv"]
            er = (r - b.rho)[f]/scales["out_drho"]
            errs.append(math.sqrt((np.sum(ex*ex) + np.sum(ev*ev) + np.sum(er*er))/(5.0*f.sum())))
    return float(np.mean(errs))


def energy(fr, mass):
    """Kinetic plus gravitational energy of the free particles (per unit depth in 2-D)."""
    f = fr.free == 1
    ke = 0.5*np.sum(mass[f]*np.sum(fr.v[f]*fr.v[f], axis=1))
    pe = G*np.sum(mass[f]*fr.x[f, 1])
    return float(ke + pe)


class HybridEngine:
    """Advances a run in windows of M steps; see the module docstring for the judge."""

    def __init__(self, cfg, p, dt, predictor, k=10, M=50, front_tol=0.01, rho_tol=0.02,
                 energy_tol=0.01, gain_tol=0.8, dim=2):
        self.cfg, self.p0, self.dt = cfg, p, dt
        self.predict, self.k, self.M = predictor, k, M
        self.front_tol, self.rho_tol, self.energy_tol = front_tol, rho_tol, energy_tol
        self.gain_tol = gain_tol
        self.dim = dim
        self.T = k*dt                      # simulated time of one network call
        self.mass = np.array(p["Mass"])
        self.h = float(p["h"][0])
        self.stats = dict(windows=0, accepted=0, reverted=0, net_calls=0, solver_steps=0,
                          t_net=0.0, t_solver=0.0)
        self.history = []

    def _solver(self, fr):
        q = self.p0.copy()
        q["x"], q["v"], q["Density"] = fr.x, fr.v, fr.rho
        q["va"], q["vb"], q["Densitya"], q["Densityb"] = fr.v, fr.v, fr.rho, fr.rho
        return Solver(self.cfg, q, self.dt, time=fr.t)

    def _plain(self, fr, duration):
        """The plain solver from fr for `duration` seconds of simulated time."""
        t0 = time.perf_counter()
        s = self._solver(fr)
        st = s.run(10**9, tf=fr.t + duration)
        out = Frame.from_particles(s.particles(), st.time)
        s.close()
        self.stats["solver_steps"] += st.steps
        self.stats["t_solver"] += time.perf_counter() - t0
        return out

    def _net(self, fr):
        """One network call, or None when the prediction fails or is not finite."""
        t0 = time.perf_counter()
        try:
            x, v, r = self.predict(fr)
        except SolverError:
            x = None
        self.stats["net_calls"] += 1
        self.stats["t_net"] += time.perf_counter() - t0
        if x is None or not (np.all(np.isfinite(x)) and np.all(np.isfinite(v)) and np.all(np.isfinite(r))):
            return None
        return Frame(x, v, r, fr.free, fr.t + self.T)

    def _front(self, fr):
        q = np.zeros(len(fr.x), dtype=[("x", "f8", 3), ("IsFree", "i4"), ("h", "f8")])
        q["x"], q["IsFree"], q["h"] = fr.x, fr.free, self.h
        return analysis.front_position(q, self.dim)

    def run(self, windows, start=None):
        """Advances `windows` windows of M steps from `start` (default: the initial state).
        Each window's record (front error against the solver, relative to the front and to the
        front's travel in the window; density drift; energy gain; position errors of the network
        and of drift stepping; accepted or not) is appended to `history`."""
        fr = start or Frame.from_particles(self.p0, 0.0)
        for _ in range(windows):
            self.stats["windows"] += 1
            e0 = energy(fr, self.mass)
            f0 = self._front(fr)
            s = fr
            for _ in range(self.M//self.k):
                s = self._net(s)
                if s is None:
                    break
            judge = self._plain(fr, self.M*self.dt)         # the solver over the same window
            rec = dict(t0=fr.t, accepted=False)
            if s is not None:
                fj, fn = self._front(judge), self._front(s)
                f = s.free == 1
                drift = float(np.mean(np.abs(s.rho[f]/1000.0 - 1.0)))
                gain = (energy(s, self.mass) - e0)/max(abs(e0), 1e-30)
                f0m = fr.free == 1
                T = judge.t - fr.t
                e_net = float(np.sqrt(np.mean(np.sum((s.x - judge.x)[f0m]**2, axis=1))))
                e_drift = float(np.sqrt(np.mean(np.sum((fr.x + T*fr.v - judge.x)[f0m]**2, axis=1))))
                rec.update(front_net=fn, front_solver=fj, front_err=abs(fn - fj)/max(abs(fj), 1e-30),
                           front_err_of_travel=abs(fn - fj)/max(fj - f0, 2.0*self.h), density_drift=drift,
                           energy_gain=gain, pos_err_net=e_net, pos_err_drift=e_drift)
                rec["accepted"] = (rec["front_err"] <= self.front_tol and drift <= self.rho_tol
                                   and gain <= self.energy_tol and e_net <= self.gain_tol*e_drift)
            self.history.append(rec)
            if rec["accepted"]:
                self.stats["accepted"] += 1
                fr = s
            else:
                self.stats["reverted"] += 1
                fr = judge
        return fr


# ---- command line -----------------------------------------------------------------------
def _log(*a):
    import sys
    print(*a, file=sys.stderr, flush=True)


def _variants(n, seed):
    return np.random.default_rng(seed).uniform(0.5, 2.0, size=n)


def trajectories(aspects, dx, area, steps, k, workers=8, log=None):
    """Plain-solver trajectories of the dam-break variants with the given aspect ratios, in
    variant order.  A ~20 k-particle run leaves most of the GPU idle, so `workers` runs proceed
    at once, each a solver on its own stream driven from its own thread (the library's calls
    release the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()

    def one(item):
        i, asp = item
        cfg, p, dt = dam_break_variant(float(asp), dx=dx, area=area)
        fr = trajectory(cfg, p, dt, steps*dt, k*dt)
        if log:
            log(f"[data] variant {i} aspect {asp:.3f}: {len(p)} particles, {len(fr)} frames "
                f"({time.perf_counter() - t0:.1f}s)")
        return fr
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(one, enumerate(aspects)))


def cmd_train(a):
    import json
    t0 = time.perf_counter()
    trajs = trajectories(_variants(a.variants, a.seed), a.dx, a.area, a.steps, a.k, a.workers, _log)
    t_data = time.perf_counter() - t0
    radius = 2.0*1.3*a.dx
    dt = dam_break_variant(1.0, dx=a.dx, area=a.area)[2]
    sc = scales_from(trajs, a.k, dt, radius)
    net = make_network(2, a.layers, sc, seed=a.seed)
    t1 = time.perf_counter()
    losses = train(net, trajs, a.iters, lr=a.lr, seed=a.seed, log=_log, log_every=max(1, a.iters//20))
    t_train = time.perf_counter() - t1
    np.save(a.out + ".params.npy", net.get_params())
    meta = dict(scales=sc, layers=a.layers, k=a.k, dx=a.dx, area=a.area, variants=a.variants, steps=a.steps,
                iters=a.iters, lr=a.lr, seed=a.seed, loss_first=float(np.mean(losses[:20])),
                loss_last=float(np.mean(losses[-20:])), train_s=t_train, ms_per_iter=1e3*t_train/max(a.iters, 1),
                data_s=t_data, frames=int(sum(len(t) for t in trajs)), workers=a.workers)
    with open(a.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps(meta))


def _load(path, train_mode=False):
    import json
    with open(path + ".json") as f:
        meta = json.load(f)
    net = make_network(2, meta["layers"], meta["scales"], train=train_mode)
    net.set_params(np.load(path + ".params.npy"))
    return net, meta


def cmd_eval(a):
    import json
    net, meta = _load(a.model)
    k, dx = meta["k"], meta["dx"]
    res = []
    for asp in _variants(a.variants, a.seed + 1000):
        cfg, p, dt = dam_break_variant(float(asp), dx=dx, area=meta["area"])
        held = [trajectory(cfg, p, dt, a.steps*dt, k*dt)]
        row = dict(aspect=float(asp), particles=len(p),
                   err_network=prediction_error(held, net_predictor(net), meta["scales"]),
                   err_ballistic=prediction_error(held, ballistic_predictor(k, dt), meta["scales"]),
                   err_drift=prediction_error(held, drift_predictor(k, dt), meta["scales"]))
        span = a.windows*a.M*dt              # simulated time of the comparison
        for name, pred in (("network", net_predictor(net)), ("ballistic", ballistic_predictor(k, dt))):
            eng = HybridEngine(cfg, p, dt, pred, k=k, M=a.M)
            t0 = time.perf_counter()
            eng.run(a.windows)
            el = time.perf_counter() - t0
            st = dict(eng.stats)
            st["wall_s"] = el
            st["simulated_s_per_wall_s"] = span/el
            row["engine_" + name] = st
        # plain solver over the same simulated time (adaptive dt: its own step count)
        s = Solver(cfg, p, dt)
        t0 = time.perf_counter()
        st = s.run(10**9, tf=span)
        el = time.perf_counter() - t0
        s.close()
        row["plain"] = dict(steps=st.steps, wall_s=el, simulated_s_per_wall_s=span/el,
                            particle_updates_per_s=len(p)*st.steps/el)
        _log(json.dumps(row))
        res.append(row)
    out = dict(model=a.model, variants=res,
               mean_err_network=float(np.mean([r["err_network"] for r in res])),
               mean_err_ballistic=float(np.mean([r["err_ballistic"] for r in res])),
               mean_err_drift=float(np.mean([r["err_drift"] for r in res])))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k2: v for k2, v in out.items() if k2 != "variants"}))


def cmd_bench3d(a):
    """Network calls on the 3-D dam break (same architecture): graph build + forward + state
    update per call, against the plain FP64 and FP32-mixed solver steps of the same state."""
    import json
    from .abi import FP32_MIXED
    cfg, p, dt = cases.dam_break_3d(dx=a.dx)
    n = len(p)
    radius = 2.0*1.3*a.dx
    k = a.k
    T = k*dt
    sc = dict(radius=radius, rho0=1000.0, vel_scale=math.sqrt(2.0*G*2.0), rho_scale=0.01, out_dx=0.5*G*T*T,
              out_dv=G*T, out_drho=1.0, drift_t=T, deg_scale=float(analysis.bulk_neighbours(3)))
    net = make_network(3, a.layers, sc, train=False, seed=a.seed)
    fr = Frame.from_particles(p, 0.0)
    t0 = time.perf_counter()
    net.load_state(fr.x, fr.v, fr.rho, fr.free)
    _log(f"[bench3d] {n} particles, graph + inputs {time.perf_counter() - t0:.2f}s")
    ne = net.graph()[1].size if a.edges else None
    net.advance()                                         # warm-up call
    t0 = time.perf_counter()
    for _ in range(a.calls):
        net.load_state(fr.x, fr.v, fr.rho, fr.free)
        net.advance()
    el = (time.perf_counter() - t0)/a.calls
    flops = None
    if ne:
        H = 128
        flops = 2.0*(ne*((32 + H)*H + a.layers*((3*H + H)*H)) + n*((32 + H)*H + a.layers*((2*H + H)*H) + 2*H*H))
    res = dict(particles=n, edges=ne, layers=a.layers, k=k, s_per_call=el,
               surrogate_particle_updates_per_s=n*k/el,
               tflops=(flops/el/1e12) if flops else None)
    for name, prec in (("fp64", 0), ("fp32", FP32_MIXED)):
        c2 = type(cfg).from_buffer_copy(cfg)
        c2.precision = prec
        s = Solver(c2, p, dt)
        s.run(5)
        t0 = time.perf_counter()
        s.run(50)
        res[f"plain_{name}_particle_updates_per_s"] = n*50/(time.perf_counter() - t0)
        s.close()
    net.close()
    print(json.dumps(res))


def cmd_engine3d(a):
    """The revert engine end to end on the 3-D dam break: `windows` windows of M dt0, each
    advanced by M/k calls of the 10-layer network (3-D inputs; the network is not trained in
    3-D, so its windows are expected to revert) and judged by the plain FP64 solver over the
    whole window.  Reports the engine's wall time, accept/revert counts and each window's
    front error, against the plain solver alone over the same simulated time."""
    import json
    cfg, p, dt = cases.dam_break_3d(dx=a.dx)
    n = len(p)
    T = a.k*dt
    sc = dict(radius=2.0*1.3*a.dx, rho0=1000.0, vel_scale=math.sqrt(2.0*G*2.0), rho_scale=0.01,
              out_dx=0.5*G*T*T, out_dv=G*T, out_drho=1.0, drift_t=T,
              deg_scale=float(analysis.bulk_neighbours(3)))
    net = make_network(3, a.layers, sc, train=False, seed=a.seed)
    eng = HybridEngine(cfg, p, dt, net_predictor(net), k=a.k, M=a.M, dim=3)
    t0 = time.perf_counter()
    out = eng.run(a.windows)
    el = time.perf_counter() - t0
    span = out.t
    s = Solver(cfg, p, dt)
    t0 = time.perf_counter()
    st = s.run(10**9, tf=span)
    el_plain = time.perf_counter() - t0
    s.close()
    net.close()
    res = dict(particles=n, layers=a.layers, k=a.k, M=a.M, windows=a.windows, simulated_s=span,
               engine=dict(eng.stats, wall_s=el, history=eng.history),
               plain=dict(steps=st.steps, wall_s=el_plain),
               engine_speedup_over_plain=el_plain/el)
    print(json.dumps(res, default=float))


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m persiansph_amd.surrogate")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--variants", type=int, default=256)
    t.add_argument("--dx", type=float, default=0.01)
    t.add_argument("--area", type=float, default=2.0)
    t.add_argument("--steps", type=int, default=1500)
    t.add_argument("--k", type=int, default=10)
    t.add_argument("--layers", type=int, default=10)
    t.add_argument("--iters", type=int, default=20000)
    t.add_argument("--lr", type=float, default=1e-4)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--workers", type=int, default=8, help="trajectories generated at once")
    t.add_argument("--out", required=True)
    e = sub.add_parser("eval")
    e.add_argument("--model", required=True)
    e.add_argument("--variants", type=int, default=4)
    e.add_argument("--steps", type=int, default=400)
    e.add_argument("--windows", type=int, default=8)
    e.add_argument("--M", type=int, default=50)
    e.add_argument("--seed", type=int, default=1)
    e.add_argument("--out", default="")
    g3 = sub.add_parser("engine3d")
    g3.add_argument("--dx", type=float, default=0.01)
    g3.add_argument("--layers", type=int, default=10)
    g3.add_argument("--k", type=int, default=10)
    g3.add_argument("--M", type=int, default=50)
    g3.add_argument("--windows", type=int, default=4)
    g3.add_argument("--seed", type=int, default=5)
    b = sub.add_parser("bench3d")
    b.add_argument("--dx", type=float, default=0.01)
    b.add_argument("--layers", type=int, default=10)
    b.add_argument("--k", type=int, default=10)
    b.add_argument("--calls", type=int, default=3)
    b.add_argument("--edges", action="store_true", help="also count edges (reads the graph back)")
    b.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval, "bench3d": cmd_bench3d, "engine3d": cmd_engine3d}[a.cmd](a)


if __name__ == "__main__":
    main()
