"""How far two oracle runs of the small granular column drift apart from a 1-ulp difference in
the initial positions.  This bounds the step count over which the GPU soil trajectory can be
held to the oracle at roundoff tolerance.  With c = 0 the Drucker-Prager apex is the stress-free
state; around step 40 the column unloads to |Sigma| ~ 1e-14, and the plastic corrector's flow
direction G/sqrt(J2) * ShearStress (Particle.cpp:704-729) is then the direction of a deviator
made of roundoff, so the two runs part by O(1e-1) of |Sigma| within 50 steps."""
import numpy as np
import pytest

from persiansph_amd import cases
from oracle.pyoracle import Oracle


def _drift(steps):
    cfg, p, dt = cases.granular_column(dim=2, dx=0.01, size=(0.1, 0.2, 0.0), floor_half=0.2, scheme=1)
    q = p.copy()
    fr = q["IsFree"] == 1
    q["x"][fr] = np.nextafter(q["x"][fr], np.inf)
    a, b = Oracle(cfg, p, dt), Oracle(cfg, q, dt)
    a.run(steps)
    b.run(steps)
    A, B = a.particles(), b.particles()
    a.close()
    b.close()
    return (float(np.abs(A["x"] - B["x"]).max()),
            float(np.abs(A["Sigma"] - B["Sigma"]).max()/np.abs(A["Sigma"]).max()))


def test_roundoff_stays_small_for_20_steps():
    dx, ds = _drift(20)
    assert dx < 1e-15 and ds < 1e-13, (dx, ds)


def test_apex_cut_amplifies_roundoff_by_50_steps():
    dx, ds = _drift(50)
    assert ds > 1e-3, ds
