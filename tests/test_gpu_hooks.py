"""Host-synchronised user hooks GeneralBefore / GeneralAfter (Domain.cpp:1767, 1770) on the GPU
path, against the oracle running the same Python hook functions at the same points.

Every step stops twice (psph_set_hooks): before the wall smoothing, where the state is what
StartAcceleration leaves, and after the pair sums.  The test's hook functions change the
accelerations and density rates of the free particles as a function of their state; the GPU
takes the changed values back with psph_hook_accel and finishes the step.
"""
import numpy as np
import pytest

from persiansph_amd import Solver, abi, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def before_hook(q):
    """GeneralBefore: a body force growing with height and a density source on free particles."""
    fr = q["IsFree"] == 1
    q["a"][fr, 0] += 0.5 + 2.0*q["x"][fr, 1]
    q["dDensity"][fr] += 0.25


def after_hook(q):
    """GeneralAfter: a damping of the free particles' motion after the pair sums."""
    fr = q["IsFree"] == 1
    q["a"][fr] -= 0.3*q["v"][fr]


def _hooked_run(s, steps, before=None, after=None, tout=float("inf"), dt_out=None, log=None):
    """Runs `steps` steps through the hook stops (and output stops every dt_out)."""
    done = 0
    while done < steps:
        st = s.run(steps - done, tout=tout)
        done += st.steps
        if st.stop_reason == abi.STOP_HOOK_BEFORE:
            q = s.particles()
            fr = q["IsFree"] == 1
            # what StartAcceleration leaves (Domain.cpp:998-1016)
            assert np.all(q["a"][~fr] == np.array(s.cfg.gravity)) and np.all(q["dDensity"] == 0.0)
            assert np.all(q["Pressure"][~fr] == 0.0) and np.all(q["NSv"][~fr] == 0.0)
            before(q)
            s.hook_accel(q["a"], q["dDensity"])
            if log is not None:
                log.append(("before", st.time))
        elif st.stop_reason in (abi.STOP_HOOK_AFTER, abi.STOP_OUTPUT):
            if after is not None:
                q = s.particles()
                after(q)
                s.hook_accel(q["a"], q["dDensity"])
            if log is not None:
                log.append(("after" if st.stop_reason == abi.STOP_HOOK_AFTER else "output", st.time))
            if st.stop_reason == abi.STOP_OUTPUT:
                tout += dt_out
        else:
            break
    return done


def _oracle_run(o, steps, tout=float("inf"), dt_out=None, log=None):
    done = 0
    while done < steps:
        st = o.run(steps - done, tout=tout)
        done += st.steps
        if st.stop_reason == abi.STOP_OUTPUT:
            if log is not None:
                log.append(("output", st.time))
            tout += dt_out
        else:
            break
    return done


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("which", ["before", "after", "both"])
def test_hooks_against_oracle(which, scheme):
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=scheme)
    b = before_hook if which in ("before", "both") else None
    a = after_hook if which in ("after", "both") else None
    steps = 30
    print(f"[hooks] {which} scheme={scheme} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    g.set_hooks(b is not None, a is not None)
    o = Oracle(cfg, p, dt)
    o.set_hooks(b, a)
    assert _hooked_run(g, steps, b, a) == steps
    assert _oracle_run(o, steps) == steps
    G, O_ = g.particles(), o.particles()
    np.testing.assert_allclose(G["x"], O_["x"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(G["v"], O_["v"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(G["Density"], O_["Density"], rtol=1e-11, atol=0)
    # the hooks changed the run, by far more than the parity bar on positions above (30 steps
    # of 35 us: the after hook's damping moves particles by some 1e-8 m)
    plain = Oracle(cfg, p, dt)
    plain.run(steps)
    assert np.abs(plain.particles()["x"] - O_["x"]).max() > 100*1e-11


def test_hooks_with_output_stops():
    """With outputs every 0.5 dt0, the GeneralAfter hook runs before the output of the same step
    (Domain.cpp:1770-1773): the stop sequence and the final state follow the oracle's."""
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=0)
    steps, dt_out = 20, 0.5*dt
    g = Solver(cfg, p, dt)
    g.set_hooks(True, True)
    o = Oracle(cfg, p, dt)
    o.set_hooks(before_hook, after_hook)
    lg, lo = [], []
    _hooked_run(g, steps, before_hook, after_hook, tout=dt_out, dt_out=dt_out, log=lg)
    _oracle_run(o, steps, tout=dt_out, dt_out=dt_out, log=lo)
    outs_g = [t for k, t in lg if k == "output"]
    outs_o = [t for k, t in lo if k == "output"]
    assert len(outs_g) == len(outs_o) >= 4
    np.testing.assert_allclose(outs_g, outs_o, rtol=1e-12)
    assert sum(1 for k, _ in lg if k == "before") == steps
    G, O_ = g.particles(), o.particles()
    np.testing.assert_allclose(G["x"], O_["x"], rtol=0, atol=1e-11)


def test_user_acc_hook_equals_body_force():
    """1-Poiseuille.cpp's UserAcc as a host GeneralBefore equals the device body-force hook bit
    for bit (Gravity + 0.002 is the same sum either way)."""
    cfg, p, dt, body = cases.poiseuille_2d()
    steps = 100
    s1 = Solver(cfg, p, dt)
    s1.set_body_force(body)
    s1.run(steps)
    s2 = Solver(cfg, p, dt)
    s2.set_hooks(True, False)

    def user_acc(q):
        fr = q["IsFree"] == 1
        q["a"][fr] += np.asarray(body)
    assert _hooked_run(s2, steps, user_acc, None) == steps
    A, B = s1.particles(), s2.particles()
    np.testing.assert_array_equal(A["x"], B["x"])
    np.testing.assert_array_equal(A["v"], B["v"])


def test_hook_accel_needs_a_stop():
    cfg, p, dt = cases.dam_break_2d(dx=0.05)
    s = Solver(cfg, p, dt)
    q = s.particles()
    with pytest.raises(Exception) as e:
        s.hook_accel(q["a"])
    assert "PSPH_ERR_STATE" in str(e.value)


def test_repeated_hooked_runs_are_identical():
    """Values handed back at a stop reach the device in full before the step resumes: identical
    hooked runs give identical states, bit for bit."""
    cfg, p, dt = cases.dam_break_2d(dx=0.05, scheme=0)
    out = []
    for _ in range(3):
        g = Solver(cfg, p, dt)
        g.set_hooks(True, True)
        assert _hooked_run(g, 12, before_hook, after_hook) == 12
        out.append(g.particles())
        g.close()
    for q in out[1:]:
        for f in ("x", "v", "Density"):
            np.testing.assert_array_equal(q[f], out[0][f], err_msg=f)
