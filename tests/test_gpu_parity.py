"""GPU parity: the HIP path against the CPU oracle on identical seeded inputs.

Bars:
  * pair sets: identical (same oriented (P1, P2) tuples);
  * one force pass: |gpu - oracle| <= 1e-12 * (|oracle| + sum_j |contribution_j|) per component,
    where the second term is the oracle's sum of absolute pair contributions (the GPU sums the
    same contributions in another order);
  * multi-step trajectories: positions, velocities and densities within the stated tolerance.
"""
import numpy as np
import pytest

from persiansph_amd import Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-12


def _pairs_set(a):
    return set(map(tuple, np.asarray(a).tolist()))


def _check_metric(name, g, o, scale, rel=REL):
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    err = np.abs(g - o)
    tol = rel*(np.abs(o) + scale)
    bad = err > tol
    if bad.any():
        k = np.argwhere(bad)[0]
        raise AssertionError(f"{name}: {bad.sum()} entries beyond {rel}; first {tuple(k)} gpu={g[tuple(k)]!r} "
                             f"oracle={o[tuple(k)]!r} scale={np.broadcast_to(scale, o.shape)[tuple(k)]!r}")


def _small_3d(scheme=1):
    return cases.dam_break(3, 0.04, (0.8, 0.6, 0.4), (0.4, 0.3, 0.4), scheme=scheme,
                           jitter=0.01, seed=7, cs=10.0*np.sqrt(2*9.81*0.3))


CASES = {
    "2d_leapfrog": lambda: cases.dam_break_2d(dx=0.02, scheme=1),
    "2d_verlet": lambda: cases.dam_break_2d(dx=0.04, scheme=0),
    "3d_leapfrog": lambda: _small_3d(1),
    "3d_verlet": lambda: _small_3d(0),
}


@pytest.mark.parametrize("case", ["2d_leapfrog", "3d_leapfrog"])
def test_pair_set_identical(case):
    cfg, p, dt = CASES[case]()
    print(f"[pairs] {case} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    gg, og = g.grid(), o.grid()
    assert list(gg.cell_no) == list(og.cell_no)
    assert list(gg.cell_size) == list(og.cell_size)
    gp, op = g.pairs(), o.pairs()
    assert len(gp) == len(op)
    assert _pairs_set(gp) == _pairs_set(op)


@pytest.mark.parametrize("case", list(CASES))
def test_first_force_pass(case):
    cfg, p, dt = CASES[case]()
    print(f"[force] {case} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt, track_scales=True)
    sg = g.run(10, tout=0.0)
    so = o.run(10, tout=0.0)
    assert sg.stop_reason == so.stop_reason == 2
    assert sg.mid_step == so.mid_step == 1
    assert sg.pairs == so.pairs
    # min over particles of sqrt_h_a*sqrt(h/|a|): |a| carries the reordered-sum rounding
    assert abs(sg.deltatmin - so.deltatmin) <= 1e-12*so.deltatmin
    G, O = g.particles(), o.particles()
    sc = o.scales(len(O))
    fr = O["IsFree"] == 1
    fx = ~fr
    _check_metric("a", G["a"][fr], O["a"][fr], sc[fr, 0:3])
    _check_metric("dDensity", G["dDensity"][fr], O["dDensity"][fr], sc[fr, 3])
    _check_metric("wall Pressure", G["Pressure"][fx], O["Pressure"][fx], sc[fx, 4])
    _check_metric("wall NSv", G["NSv"][fx], O["NSv"][fx], sc[fx, 5:8])
    np.testing.assert_array_equal(G["SumKernel"][fx] > 0, O["SumKernel"][fx] > 0)
    # the walls' own pair sums (Interaction.cpp:166-167): Gravity plus their pairs' share
    _check_metric("wall a", G["a"][fx], O["a"][fx], sc[fx, 0:3])
    _check_metric("wall dDensity", G["dDensity"][fx], O["dDensity"][fx], sc[fx, 3])
    assert np.abs(O["a"][fx] - np.array(cfg.gravity)).max() > 1.0
    # the packed output carries the same walls' accelerations
    g.output_begin()
    out = g.output_end()
    np.testing.assert_array_equal(out["acceleration"], G["a"].astype(np.float32))


@pytest.mark.parametrize("case", list(CASES))
def test_trajectory(case):
    cfg, p, dt = CASES[case]()
    steps = 60
    print(f"[traj] {case} n={len(p)} steps={steps}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    sg = g.run(steps)
    so = o.run(steps)
    assert sg.steps == so.steps == steps
    assert abs(sg.time - so.time) <= 1e-12*so.time
    G, O = g.particles(), o.particles()
    assert len(G) == len(O)
    # a few ulps per step of reordered sums; the flow is smooth at these times
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(G["v"], O["v"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(G["Density"], O["Density"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(G["Pressure"], O["Pressure"], rtol=0, atol=1e-8*np.abs(O["Pressure"]).max())


def test_output_stop_and_resume():
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=0)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    tout = 5.5*dt
    for _ in range(3):
        sg = g.run(1000, tout=tout)
        so = o.run(1000, tout=tout)
        assert sg.stop_reason == so.stop_reason == 2
        assert sg.steps == so.steps
        assert abs(sg.time - so.time) <= 1e-12*so.time
        tout += 3.0*dt
    sg = g.run(4, tf=so.time + 2.5*dt)
    so = o.run(4, tf=so.time + 2.5*dt)
    assert sg.steps == so.steps and sg.stop_reason == so.stop_reason
    G, O = g.particles(), o.particles()
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-12)


def test_particle_leave_deletion():
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=1)
    fl = np.flatnonzero(p["IsFree"] == 1)
    # two isolated fluid particles launched upward out of the box, one in the middle of the
    # particle array (deleting it re-indexes every later particle) and one at its end
    extra = p[fl[:2]].copy()
    extra["x"] = [(2.0, 2.9, 0.0), (3.0, 2.85, 0.0)]
    extra["v"] = [(0.0, 200.0, 0.0), (0.0, 150.0, 0.0)]
    extra["Density"] = 1000.0
    extra["Densityb"] = 1000.0
    p = np.concatenate([p[:100], extra[:1], p[100:], extra[1:]])
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    sg = g.run(40)
    so = o.run(40)
    assert so.deleted == 2
    assert sg.deleted == so.deleted and sg.n == so.n == len(p) - 2
    assert sg.steps == so.steps == 40
    G, O = g.particles(), o.particles()
    np.testing.assert_array_equal(G["ID"], O["ID"])
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("case", ["2d_leapfrog", "3d_leapfrog"])
def test_general_kernels_match_fast_path(case, monkeypatch):
    """The specialised fluid gather (cubic spline, Morris) and the general pair kernel give the
    same force pass to the 1e-12 bar, and both meet it against the oracle."""
    cfg, p, dt = CASES[case]()
    o = Oracle(cfg, p, dt, track_scales=True)
    o.run(1, tout=0.0)
    O_ = o.particles()
    sc = o.scales(len(O_))
    fr = O_["IsFree"] == 1
    res = []
    for general in (False, True):
        if general:
            monkeypatch.setenv("PSPH_GENERAL_KERNELS", "1")
        else:
            monkeypatch.delenv("PSPH_GENERAL_KERNELS", raising=False)
        g = Solver(cfg, p, dt)
        g.run(1, tout=0.0)
        G = g.particles()
        _check_metric("a", G["a"][fr], O_["a"][fr], sc[fr, 0:3])
        _check_metric("dDensity", G["dDensity"][fr], O_["dDensity"][fr], sc[fr, 3])
        res.append(G)
        g.close()
    _check_metric("fast vs general a", res[0]["a"][fr], res[1]["a"][fr], sc[fr, 0:3], rel=2e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_mixed_against_fp64(dim):
    """The FP32-mixed tier (cell-relative FP32 pair arithmetic, Kahan continuity sum, FP64
    state) against the FP64 oracle: one force pass within 1e-5 of the contribution scale,
    and 60-step positions within 1e-7 m."""
    from persiansph_amd import FP32_MIXED
    if dim == 2:
        cfg, p, dt = cases.dam_break_2d(dx=0.02, scheme=1)
    else:
        cfg, p, dt = _small_3d(1)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = FP32_MIXED
    o = Oracle(cfg, p, dt, track_scales=True)
    g = Solver(cfg32, p, dt)
    sg, so = g.run(1, tout=0.0), o.run(1, tout=0.0)
    assert abs(sg.pairs - so.pairs) <= max(2, so.pairs*1e-4)
    G, O_ = g.particles(), o.particles()
    sc = o.scales(len(O_))
    fr = O_["IsFree"] == 1
    _check_metric("fp32 a", G["a"][fr], O_["a"][fr], sc[fr, 0:3], rel=1e-5)
    _check_metric("fp32 dDensity", G["dDensity"][fr], O_["dDensity"][fr], sc[fr, 3], rel=1e-5)
    sg, so = g.run(60), o.run(60)
    G, O_ = g.particles(), o.particles()
    np.testing.assert_allclose(G["x"], O_["x"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(G["Density"], O_["Density"], rtol=1e-5)


def _run_env(monkeypatch, env, cfg, p, dt, steps):
    for k in ("PSPH_NL_CAP", "PSPH_WL_CAP", "PSPH_SKIN", "PSPH_REBUILD_EVERY_STEP"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    g = Solver(cfg, p, dt)
    st = g.run(steps)
    q = g.particles()
    g.close()
    return st, q


@pytest.mark.parametrize("lists", ["per_step", "skin"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_list_overflow_is_exact(monkeypatch, dim, precision, lists):
    """Targets with more neighbours than the lists hold are finished by the overflow pass,
    which continues the same sums in list order: with capacity 8 (most targets overflow on the
    first batch, then the run grows the lists) the states are bitwise those of the default
    capacity.  `per_step`: lists of the support radius built every step; `skin`: lists with the
    skin, walked over the 5^DIM cells, rebuilt every step so that both runs sum in the same
    order (wall sub-lists at capacity 4 as well).  The skin lists are gathered per tile, eight
    lanes per target: there an overflowing target's first entries are summed in eight
    interleaved partial sums before the overflow pass adds the rest in list order, so the two
    capacities agree to the rounding of the regrouped sums (FP64, or FP32 for the FP32-mixed
    tier) instead of bit for bit."""
    from persiansph_amd import FP32_MIXED
    if dim == 2:
        cfg, p, dt = cases.dam_break_2d(dx=0.02, scheme=1)
    else:
        cfg, p, dt = _small_3d(1)
    if precision == "fp32":
        cfg = type(cfg).from_buffer_copy(cfg)
        cfg.precision = FP32_MIXED
    print(f"[nlist] dim={dim} {precision} n={len(p)}", flush=True)
    steps = 40
    env = {"PSPH_SKIN": "0"} if lists == "per_step" else {"PSPH_REBUILD_EVERY_STEP": "1"}
    small = {"PSPH_NL_CAP": "8"} if lists == "per_step" else {"PSPH_NL_CAP": "8", "PSPH_WL_CAP": "4"}
    base = _run_env(monkeypatch, env, cfg, p, dt, steps)
    over = _run_env(monkeypatch, {**env, **small}, cfg, p, dt, steps)
    assert base[0].steps == over[0].steps == steps
    assert base[0].pairs == over[0].pairs
    if lists == "skin":
        b, o = base[1], over[1]
        x, v, rho, a = (1e-12, 1e-10, 1e-12, 1e-10) if precision == "fp64" else (1e-8, 1e-5, 1e-6, 1e-5)
        np.testing.assert_allclose(o["x"], b["x"], rtol=0, atol=x)
        np.testing.assert_allclose(o["v"], b["v"], rtol=0, atol=v)
        np.testing.assert_allclose(o["Density"], b["Density"], rtol=rho, atol=0)
        np.testing.assert_allclose(o["a"], b["a"], rtol=0, atol=a*np.abs(b["a"]).max())
        return
    for f in ("x", "v", "Density", "a"):
        np.testing.assert_array_equal(base[1][f], over[1][f], err_msg=f)
