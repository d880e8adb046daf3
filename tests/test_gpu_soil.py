"""GPU parity for the solid/soil path: CalcForce2233 gather, artificial stress (TIR), the fused
Jaumann + Drucker-Prager update (Mat3Leapfrog / Mat3MVerlet), against the CPU oracle."""
import numpy as np
import pytest

from persiansph_amd import Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _check(name, g, o, scale, rel):
    g = np.asarray(g, dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    err = np.abs(g - o)
    tol = rel*(np.abs(o) + scale)
    bad = err > tol
    assert not bad.any(), (name, int(bad.sum()), float((err/np.maximum(tol, 1e-300)).max()))


def _soil(dim, scheme):
    if dim == 2:
        return cases.granular_column(dim=2, dx=0.01, size=(0.1, 0.2, 0.0), floor_half=0.2, scheme=scheme)
    return cases.granular_column(dim=3, dx=0.02, size=(0.1, 0.2, 0.1), floor_half=0.2, scheme=scheme)


@pytest.mark.parametrize("dim,scheme", [(2, 1), (2, 0), (3, 1), (3, 0)])
def test_soil_first_force_pass(dim, scheme):
    cfg, p, dt = _soil(dim, scheme)
    print(f"[soil force] dim={dim} scheme={scheme} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt, track_scales=True)
    sg, so = g.run(1, tout=0.0), o.run(1, tout=0.0)
    assert sg.pairs == so.pairs
    G, O_ = g.particles(), o.particles()
    sc = o.scales(len(O_))
    fr = O_["IsFree"] == 1
    _check("a", G["a"][fr], O_["a"][fr], sc[fr, 0:3], 1e-12)
    _check("dDensity", G["dDensity"][fr], O_["dDensity"][fr], sc[fr, 3], 1e-12)
    sr_idx = [0, 1, 2, 4, 5, 8]
    _check("StrainRate", G["StrainRate"][fr][:, sr_idx], O_["StrainRate"][fr][:, sr_idx], sc[fr, 9:15], 1e-12)
    rr_idx = [1, 2, 5]
    _check("RotationRate", G["RotationRate"][fr][:, rr_idx], O_["RotationRate"][fr][:, rr_idx], sc[fr, 15:18], 1e-12)
    # walls: smoothed stress (Sigma = sum K sigma_j / sum K) and their own pair sums (the
    # reference adds every pair into both particles)
    fx = ~fr
    np.testing.assert_allclose(G["Sigma"][fx], O_["Sigma"][fx], rtol=1e-12, atol=1e-12*np.abs(O_["Sigma"]).max())
    _check("wall a", G["a"][fx], O_["a"][fx], sc[fx, 0:3], 1e-12)
    _check("wall dDensity", G["dDensity"][fx], O_["dDensity"][fx], sc[fx, 3], 1e-12)


# 20 steps: this column unloads to the stress-free Drucker-Prager apex around step 40, where
# the plastic flow direction ShearStress/sqrt(J2) is set by roundoff, so even two oracle runs
# 1 ulp apart part by O(1e-1) of |Sigma| by step 50 (tests/test_oracle_soil_sensitivity.py);
# over 20 steps they agree to ~1e-15.
@pytest.mark.parametrize("dim,scheme", [(2, 1), (2, 0), (3, 1), (3, 0)])
def test_soil_trajectory(dim, scheme):
    cfg, p, dt = _soil(dim, scheme)
    steps = 20
    print(f"[soil traj] dim={dim} scheme={scheme} n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt)
    sg, so = g.run(steps), o.run(steps)
    assert sg.steps == so.steps == steps
    assert abs(sg.time - so.time) <= 1e-12*so.time
    G, O_ = g.particles(), o.particles()
    smax = np.abs(O_["Sigma"]).max()
    np.testing.assert_allclose(G["x"], O_["x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(G["v"], O_["v"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(G["Sigma"], O_["Sigma"], rtol=0, atol=1e-9*smax)
    np.testing.assert_allclose(G["Strain"], O_["Strain"], rtol=0, atol=1e-12)
    assert abs(cases.runout(G) - cases.runout(O_)) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_soil_fp32_mixed_against_fp64(dim):
    """FP32-mixed soil tier (CalcForce2233 pair arithmetic and sums in FP32, Kahan continuity
    sum, FP64 stress update and Drucker-Prager return mapping) against the FP64 oracle: one
    force pass within 1e-5 of the contribution scale, 20-step positions within 1e-9 m and
    stresses within 1e-4 of the largest stress."""
    from persiansph_amd import FP32_MIXED
    cfg, p, dt = _soil(dim, 1)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = FP32_MIXED
    print(f"[soil fp32] dim={dim} n={len(p)}", flush=True)
    g = Solver(cfg32, p, dt)
    o = Oracle(cfg, p, dt, track_scales=True)
    sg, so = g.run(1, tout=0.0), o.run(1, tout=0.0)
    assert sg.pairs == so.pairs
    G, O_ = g.particles(), o.particles()
    sc = o.scales(len(O_))
    fr = O_["IsFree"] == 1
    _check("a", G["a"][fr], O_["a"][fr], sc[fr, 0:3], 1e-5)
    _check("dDensity", G["dDensity"][fr], O_["dDensity"][fr], sc[fr, 3], 1e-5)
    sr_idx = [0, 1, 2, 4, 5, 8]
    _check("StrainRate", G["StrainRate"][fr][:, sr_idx], O_["StrainRate"][fr][:, sr_idx], sc[fr, 9:15], 1e-5)
    g2, o2 = Solver(cfg32, p, dt), Oracle(cfg, p, dt)
    g2.run(20)
    o2.run(20)
    G, O_ = g2.particles(), o2.particles()
    np.testing.assert_allclose(G["x"], O_["x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(G["Sigma"], O_["Sigma"], rtol=0, atol=1e-4*np.abs(O_["Sigma"]).max())
