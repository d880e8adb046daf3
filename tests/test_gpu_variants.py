"""GPU parity of the general fluid kernels across the reference's physics switches.

Each variant of a small 2-D dam break (dx = 0.04, ~1,500 particles) is checked against the
oracle on the same inputs:
  * the first force pass: |gpu - oracle| <= 1e-12 * (|oracle| + sum_j |contribution_j|);
  * a 30-step trajectory: positions within 1e-11 m, densities within 1e-11 relative.
The variants cover Kernel_Set (cubic spline, Wendland C2, quintic spline), Viscosity_Eq_Set
(Morris, Shao, incompressible, Takeda), Gradient_Approach_Set (squared and multiplied
density), XSPH and Shepard under modified Verlet, artificial stress (TI) and the
non-Newtonian viscosity models (T0 > 0, Bingham and Cross).
"""
import numpy as np
import pytest

from persiansph_amd import Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-12


def _check(name, g, o, scale, rel=REL):
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    bad = np.abs(g - o) > rel*(np.abs(o) + scale)
    if bad.any():
        k = tuple(np.argwhere(bad)[0])
        raise AssertionError(f"{name}: {bad.sum()} entries beyond {rel}; first {k} gpu={g[k]!r} oracle={o[k]!r}")


def _variant(kernel_type=0, visc_eq=0, gradient_type=0, scheme=1, xsph=0.0, ti=0.0, t0=0.0, vism=0,
             mu=0.5):
    """The dam break with viscosity mu, seeded random fluid velocities (|v| <= 0.5 m/s per
    component) and densities perturbed by up to 0.3 % (some pressures negative), so that the
    viscosity laws, XSPH, artificial stress and the shear-rate models all contribute."""
    cfg, p, dt = cases.dam_break_2d(dx=0.04, scheme=scheme, kernel_type=kernel_type)
    fl = p["IsFree"] == 1
    rng = np.random.default_rng(11)
    v = rng.uniform(-0.5, 0.5, size=(int(fl.sum()), 3))
    v[:, 2] = 0.0
    p["v"][fl] = v
    p["Density"][fl] *= 1.0 + rng.uniform(-3e-3, 3e-3, size=int(fl.sum()))
    p["Densityb"][fl] = p["Density"][fl]
    p["Mu"] = mu                       # a viscous fluid, so the viscosity laws weigh in
    p["MuRef"] = mu
    cfg = type(cfg).from_buffer_copy(cfg)
    cfg.visc_eq = visc_eq
    cfg.gradient_type = gradient_type
    cfg.xsph = xsph
    if ti > 0.0:
        p["TI"] = ti
        p["TIn"] = 4.0
        p["TIInitDist"] = 0.04
    if t0 > 0.0:
        fl = p["IsFree"] == 1
        p["T0"][fl] = t0
        p["VisM"][fl] = vism
        p["m"][fl] = 300.0
    return cfg, p, dt


VARIANTS = {
    "wendland": dict(kernel_type=1),
    "quintic": dict(kernel_type=2),
    "shao": dict(visc_eq=1),
    "incompressible": dict(visc_eq=2),
    "takeda": dict(visc_eq=3),
    "multiplied_density": dict(gradient_type=1),
    "verlet_xsph_shepard": dict(scheme=0, xsph=0.5),
    "artificial_stress": dict(ti=0.3),
    "bingham": dict(t0=0.5, vism=0),
    "cross": dict(t0=0.5, vism=1),
}


@pytest.mark.parametrize("name", list(VARIANTS))
def test_variant_force_pass_and_trajectory(name):
    cfg, p, dt = _variant(**VARIANTS[name])
    print(f"[variant] {name} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt, track_scales=True)
    sg, so = g.run(1, tout=0.0), o.run(1, tout=0.0)
    assert sg.pairs == so.pairs
    G, O = g.particles(), o.particles()
    sc = o.scales(len(O))
    fr = O["IsFree"] == 1
    fx = ~fr
    _check("a", G["a"][fr], O["a"][fr], sc[fr, 0:3])
    _check("dDensity", G["dDensity"][fr], O["dDensity"][fr], sc[fr, 3])
    _check("wall Pressure", G["Pressure"][fx], O["Pressure"][fx], sc[fx, 4])
    steps = 30
    sg, so = g.run(steps), o.run(steps)
    assert sg.steps == so.steps
    G, O = g.particles(), o.particles()
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(G["Density"], O["Density"], rtol=1e-11, atol=0)
    g.close()


@pytest.mark.parametrize("name", ["takeda", "artificial_stress"])
def test_general_pair_lists_are_exact(name, monkeypatch):
    """Without list reuse the general kernels take each target's pairs from a list built every
    step in the walk order; the walk itself serves only entries beyond the capacity.  Lists of
    capacity 8 (most targets overflow, the host then grows them) and no lists at all give
    bitwise the same run."""
    cfg, p, dt = _variant(**VARIANTS[name])
    runs = {}
    monkeypatch.setenv("PSPH_NO_GREUSE", "1")
    for mode in ("lists", "cap8", "nolists"):
        monkeypatch.delenv("PSPH_NL_CAP", raising=False)
        monkeypatch.delenv("PSPH_NO_GLIST", raising=False)
        if mode == "cap8":
            monkeypatch.setenv("PSPH_NL_CAP", "8")
        if mode == "nolists":
            monkeypatch.setenv("PSPH_NO_GLIST", "1")
        print(f"[variant] {name} {mode}", flush=True)
        g = Solver(cfg, p, dt)
        st1 = g.run(1, tout=0.0)
        a1 = g.particles()["a"].copy()
        st = g.run(12)
        q = g.particles()
        runs[mode] = (st1.pairs, a1, q["x"].copy(), q["Density"].copy(), st.pairs)
        g.close()
    ref = runs["lists"]
    for mode in ("cap8", "nolists"):
        got = runs[mode]
        assert got[0] == ref[0] and got[4] == ref[4], mode
        for k in (1, 2, 3):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=mode)


_REUSE_MODES = {
    "reuse": {},
    "reuse_cap8": {"PSPH_NL_CAP": "8", "PSPH_WL_CAP": "4"},
    "rebuild_every_step": {"PSPH_REBUILD_EVERY_STEP": "1"},
}


@pytest.mark.parametrize("name", ["quintic", "verlet_xsph_shepard", "artificial_stress"])
def test_general_list_reuse_keeps_the_pair_set(name, monkeypatch):
    """With list reuse the general kernels keep skin lists (candidates within the support plus
    a skin) between builds, for free and wall targets, and keep the entries that are pairs of
    the reference's set at the current positions.  Against per-step lists: the same pair counts
    at every step, the first force pass within 1e-12 of the largest acceleration (the sums run
    in another order), and 40-step trajectories within 1e-11 m and 1e-11 relative density.
    Lists of capacity 8 (overflow at builds: the host grows them) and a rebuild at every step
    give the same."""
    cfg, p, dt = _variant(**VARIANTS[name])
    env = ("PSPH_NO_GREUSE", "PSPH_NL_CAP", "PSPH_WL_CAP", "PSPH_REBUILD_EVERY_STEP")

    def run(mode):
        for k in env:
            monkeypatch.delenv(k, raising=False)
        if mode == "lists":
            monkeypatch.setenv("PSPH_NO_GREUSE", "1")
        for k, v in _REUSE_MODES.get(mode, {}).items():
            monkeypatch.setenv(k, v)
        print(f"[reuse] {name} {mode}", flush=True)
        g = Solver(cfg, p, dt)
        st1 = g.run(1, tout=0.0)
        a1 = g.particles()["a"].copy()
        pairs = []
        for _ in range(40):
            pairs.append(g.run(1).pairs)
        q = g.particles()
        out = (st1.pairs, a1, q["x"].copy(), q["Density"].copy(), pairs)
        g.close()
        return out

    ref = run("lists")
    fr = p["IsFree"] == 1
    amax = np.abs(ref[1][fr]).max()
    for mode in _REUSE_MODES:
        got = run(mode)
        assert got[0] == ref[0] and got[4] == ref[4], mode
        assert np.abs(got[1][fr] - ref[1][fr]).max() <= 1e-12*amax, mode
        np.testing.assert_allclose(got[2], ref[2], rtol=0, atol=1e-11, err_msg=mode)
        np.testing.assert_allclose(got[3], ref[3], rtol=1e-11, atol=0, err_msg=mode)
