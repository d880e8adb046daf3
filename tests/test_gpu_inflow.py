"""GPU parity with in/out-flow boundaries (InFlowBCFresh, InFlowBCLeave and the in/out-flow
branch of Domain::Move; Source/Domain.cpp:1299-1325, 1327-1635).

The 2-D channel of persiansph_amd.cases.channel_inflow_2d runs on the GPU path and on the oracle
from the same inputs, in chunks of steps; after each chunk both hold the same particles in the
same order (the recycling overwrites, creates and deletes exactly as Array::Push / DelItems
do): the count and the InOut flags agree exactly, positions within 1e-11 m, velocities within
1e-9 m/s and densities 1e-11 relative.  Cases: inflow and outflow (modified Verlet and
Leapfrog), inflow only (the particles that leave at the right are deleted, the inflow list is
rebuilt), and an outflow velocity scaled by InPart/OutPart (mass conservation) with an outflow
density.
"""
import numpy as np
import pytest

from persiansph_amd import Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "in_out_verlet": dict(in_out_flow=3),
    "in_out_leapfrog": dict(in_out_flow=3, scheme=1),
    "inflow_only": dict(in_out_flow=1),
    "outflow_mass_conservation": dict(in_out_flow=3, out_v=0.8, out_density=1000.5, mass_conservation=True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_inflow_parity(case):
    cfg, p, dt = cases.channel_inflow_2d(**CASES[case])
    print(f"[inflow] {case} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    created = deleted = recycled = 0
    n0 = len(p)
    prev = o.particles()
    for chunk in range(6):
        sg, so = g.run(250), o.run(250)
        assert sg.steps == so.steps == 250
        assert sg.deleted == so.deleted
        deleted += so.deleted
        G, O = g.particles(), o.particles()
        assert len(G) == len(O), (chunk, len(G), len(O))
        np.testing.assert_array_equal(G["InOut"], O["InOut"])
        np.testing.assert_array_equal(G["IsFree"], O["IsFree"])
        np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(G["v"], O["v"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(G["Density"], O["Density"], rtol=1e-11, atol=0)
        created = len(O) - n0 + deleted
        m = min(len(prev), len(O))
        recycled += int((O["x"][:m, 0] < prev["x"][:m, 0] - 0.25).sum())   # moved back upstream
        prev = O
    assert (O["InOut"] == 1).any()
    if cfg.in_out_flow == 1:
        assert deleted > 0                     # outflow without the outflow zone deletes
    assert recycled > 0 or created > 0         # the recycling ran
    g.close()


def test_inflow_growth_past_the_buffers(monkeypatch):
    """With no spare room in the device buffers (PSPH_INOUT_SLACK=0), the first recycling that
    creates more particles than it deletes re-creates the buffers for more particles instead of
    failing; the run stays on the oracle's, whose Particles array simply grows."""
    monkeypatch.setenv("PSPH_INOUT_SLACK", "0")
    cfg, p, dt = cases.channel_inflow_2d(in_out_flow=1)
    print(f"[inflow growth] n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    n0, nmax = len(p), len(p)
    for chunk in range(8):
        sg, so = g.run(250), o.run(250)
        assert sg.steps == so.steps == 250
        G, O = g.particles(), o.particles()
        assert len(G) == len(O)
        nmax = max(nmax, len(O))
        np.testing.assert_array_equal(G["InOut"], O["InOut"])
        np.testing.assert_allclose(G["x"], O["x"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(G["v"], O["v"], rtol=0, atol=1e-9)
    assert nmax > n0                            # the buffers had to grow
    out = g.output_f32()                        # outputs after the growth
    np.testing.assert_array_equal(out["position"], G["x"].astype(np.float32))
    g.close()
