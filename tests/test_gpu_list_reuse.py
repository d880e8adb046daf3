"""Neighbour-list reuse on the fast fluid path.

Lists are built with the support radius plus a skin and reused until some free particle has
moved half the skin; every step keeps exactly the reference's pairs from them.  The bars:
  * against the oracle: the trajectory tolerances of the per-step path, with the pair count of
    the last force pass identical, over runs that rebuild several times;
  * against rebuilding every step (no reuse): agreement to rounding of reordered sums;
  * overflowing lists and host-side changes (output stops, pair queries, deletions) force a
    rebuild and keep the results.
"""
import numpy as np
import pytest

from persiansph_amd import FP32_MIXED, Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

ENV = ("PSPH_SKIN", "PSPH_NL_CAP", "PSPH_WL_CAP", "PSPH_REBUILD_EVERY_STEP")


def _small_3d(scheme=1):
    return cases.dam_break(3, 0.04, (0.8, 0.6, 0.4), (0.4, 0.3, 0.4), scheme=scheme,
                           jitter=0.01, seed=7, cs=10.0*np.sqrt(2*9.81*0.3))


CASES = {
    "2d": lambda: cases.dam_break_2d(dx=0.02, scheme=1),
    "3d": lambda: _small_3d(1),
}


def _solver(monkeypatch, env, cfg, p, dt):
    for k in ENV:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    return Solver(cfg, p, dt)


def _run(s, steps, chunk=60):
    """Runs `steps` steps in calls of `chunk`; returns the last call's status with `steps` set
    to the steps of all calls, and the list builds of all calls."""
    builds, done, st = 0, 0, None
    while done < steps:
        st = s.run(min(chunk, steps - done))
        assert st.steps > 0
        builds += st.list_builds
        done += st.steps
    st.steps = done
    return st, builds


def _close(G, O, x=1e-10, v=1e-8, rho=1e-10):
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=x)
    np.testing.assert_allclose(G["v"], O["v"], rtol=0, atol=v)
    np.testing.assert_allclose(G["Density"], O["Density"], rtol=rho, atol=0)


@pytest.mark.parametrize("case", list(CASES))
def test_reused_lists_follow_the_oracle(monkeypatch, case):
    """A small skin (1 % of the support radius) makes the run rebuild every few dozen steps;
    the trajectory stays on the oracle's and the last force pass has the oracle's pair count."""
    cfg, p, dt = CASES[case]()
    steps = 240
    print(f"[reuse-oracle] {case} n={len(p)} steps={steps}", flush=True)
    g = _solver(monkeypatch, {"PSPH_SKIN": "0.01"}, cfg, p, dt)
    o = Oracle(cfg, p, dt)
    sg, builds = _run(g, steps)
    so = o.run(steps)
    print(f"[reuse-oracle] {case} list builds {builds} of {steps} steps", flush=True)
    assert sg.steps == so.steps
    assert 2 < builds < steps
    assert sg.pairs == so.pairs
    assert abs(sg.time - so.time) <= 1e-12*so.time
    _close(g.particles(), o.particles())
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", list(CASES))
def test_reuse_matches_rebuilding_every_step(monkeypatch, case, precision):
    """The default skin (a few rebuilds in this window), a small skin, and lists that
    overflow (every step rebuilds until the host has grown them) agree with a run that sorts
    and builds its lists every step, to rounding of the reordered sums."""
    cfg, p, dt = CASES[case]()
    if precision == "fp32":
        cfg = type(cfg).from_buffer_copy(cfg)
        cfg.precision = FP32_MIXED
    steps = 160
    print(f"[reuse-every] {case} {precision} n={len(p)}", flush=True)
    ref = _solver(monkeypatch, {"PSPH_SKIN": "0"}, cfg, p, dt)
    sr, br = _run(ref, steps)
    R = ref.particles()
    ref.close()
    assert br == steps
    tol = dict(x=1e-11, v=1e-9, rho=1e-11) if precision == "fp64" else dict(x=1e-8, v=1e-5, rho=1e-6)
    for env in ({}, {"PSPH_SKIN": "0.01"}, {"PSPH_NL_CAP": "8", "PSPH_WL_CAP": "4"}):
        g = _solver(monkeypatch, env, cfg, p, dt)
        sg, bg = _run(g, steps, chunk=40)
        print(f"[reuse-every] {case} {precision} env={env}: {bg} list builds", flush=True)
        assert sg.steps == sr.steps and sg.pairs == sr.pairs
        if not env:
            assert bg < steps//10
        _close(g.particles(), R, **tol)
        g.close()


@pytest.mark.parametrize("case", list(CASES))
def test_host_changes_force_a_rebuild(monkeypatch, case):
    """Output stops (the step is finished by the next call), a pair query (which re-sorts the
    particles) and more steps after it, all with lists that are being reused."""
    cfg, p, dt = CASES[case]()
    g = _solver(monkeypatch, {"PSPH_SKIN": "0.01"}, cfg, p, dt)
    o = Oracle(cfg, p, dt)
    tout = 40.5*dt
    for _ in range(3):
        sg, so = g.run(1000, tout=tout), o.run(1000, tout=tout)
        assert sg.stop_reason == so.stop_reason == 2
        assert sg.steps == so.steps
        tout += 37.0*dt
    sg, so = g.run(5), o.run(5)            # finishes the pending step
    assert sg.steps == so.steps and sg.mid_step == so.mid_step == 0
    gp, op = g.pairs(), o.pairs()
    assert len(gp) == len(op)
    assert set(map(tuple, gp.tolist())) == set(map(tuple, op.tolist()))
    sg, so = g.run(80), o.run(80)
    assert sg.steps == so.steps == 80 and sg.pairs == so.pairs
    _close(g.particles(), o.particles())
    g.close()


def test_deletion_with_reused_lists(monkeypatch):
    """Particles leaving the domain are deleted (the survivors re-uploaded) while lists are
    being reused; the next step rebuilds them for the new particle set."""
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=1)
    fl = np.flatnonzero(p["IsFree"] == 1)
    extra = p[fl[:2]].copy()
    extra["x"] = [(2.0, 2.9, 0.0), (3.0, 2.85, 0.0)]
    extra["v"] = [(0.0, 200.0, 0.0), (0.0, 150.0, 0.0)]
    extra["Density"] = 1000.0
    extra["Densityb"] = 1000.0
    p = np.concatenate([p[:100], extra[:1], p[100:], extra[1:]])
    g = _solver(monkeypatch, {"PSPH_SKIN": "0.05"}, cfg, p, dt)
    o = Oracle(cfg, p, dt)
    sg, so = g.run(60), o.run(60)
    assert so.deleted == 2 and sg.deleted == so.deleted and sg.n == so.n
    G, O = g.particles(), o.particles()
    np.testing.assert_array_equal(G["ID"], O["ID"])
    np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-10)
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", list(CASES))
def test_tile_gather_matches_list_gather(monkeypatch, case, precision):
    """The gather per tile (neighbour records staged in LDS, eight lanes per target; FP32-mixed:
    the cell offsets from the staged cells) against the per-target list gather (PSPH_NO_TILES)
    and against tiles that mostly read their neighbours from HBM (PSPH_TILE_ITEMS=64: few tiles
    fit): the same pairs, and states that agree to the rounding of the regrouped sums (FP32
    sums for the FP32-mixed tier)."""
    cfg, p, dt = CASES[case]()
    if precision == "fp32":
        cfg = type(cfg).from_buffer_copy(cfg)
        cfg.precision = FP32_MIXED
    steps = 120
    res = {}
    for name, env in (("tiles", {}), ("lists", {"PSPH_NO_TILES": "1"}), ("fallback", {"PSPH_TILE_ITEMS": "64"})):
        for k in ("PSPH_NO_TILES", "PSPH_TILE_ITEMS"):
            monkeypatch.delenv(k, raising=False)
        print(f"[tiles] {case} {precision} {name}", flush=True)
        g = _solver(monkeypatch, env, cfg, p, dt)
        st, builds = _run(g, steps, chunk=40)
        res[name] = (st, g.particles())
        g.close()
    st0, R = res["lists"]
    tol = {} if precision == "fp64" else dict(x=1e-8, v=1e-5, rho=1e-6)
    for name in ("tiles", "fallback"):
        st, G = res[name]
        assert st.steps == st0.steps and st.pairs == st0.pairs, name
        _close(G, R, **tol)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", list(CASES))
def test_lean_steps_equal_full_steps(monkeypatch, case, precision):
    """Steps that keep their lists are queued without the sort and list-build kernels, which are
    no-ops on such steps; a tail before a step that must rebuild hands it back to the host; and
    a lean step after another step of the same queue takes the views that step's integration
    wrote instead of running k_reorder.  Queuing every step in full (PSPH_NO_LEAN), and lean
    steps that each run k_reorder (PSPH_NO_FUSED_VIEWS), give the same run bit for bit, with
    the same list builds, over calls of several lengths (a small skin: many rebuilds)."""
    cfg, p, dt = CASES[case]()
    if precision == "fp32":
        cfg = type(cfg).from_buffer_copy(cfg)
        cfg.precision = FP32_MIXED
    res = {}
    for name, env in (("lean", {}), ("full", {"PSPH_NO_LEAN": "1"}), ("unfused", {"PSPH_NO_FUSED_VIEWS": "1"})):
        for k in ("PSPH_NO_LEAN", "PSPH_NO_FUSED_VIEWS"):
            monkeypatch.delenv(k, raising=False)
        print(f"[lean] {case} {precision} {name}", flush=True)
        g = _solver(monkeypatch, dict(env, PSPH_SKIN="0.01"), cfg, p, dt)
        builds = 0
        for chunk in (1, 7, 40, 300):
            st, b = _run(g, chunk, chunk=chunk)
            builds += b
        res[name] = (builds, st, g.particles())
        g.close()
    b0, s0, A = res["lean"]
    assert b0 >= 3, b0
    for name in ("full", "unfused"):
        b1, s1, B = res[name]
        assert b0 == b1, (name, b0, b1)
        assert s0.time == s1.time and s0.pairs == s1.pairs, name
        for f in ("x", "v", "Density", "a"):
            np.testing.assert_array_equal(A[f], B[f], err_msg=f"{name} {f}")
