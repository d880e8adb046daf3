"""Pairs of different materials in the oracle: CalcForce12 (fluid-solid, Interaction.cpp:379-506)
with the FSI smoothing of PrimaryComputeAcceleration (Domain.cpp:1089-1160), CalcForce13
(soil-water, Interaction.cpp:508-653) with Seepage (Functions.cpp:266-298), the saturation of
soil under water (Domain.cpp:1060-1088, 1209-1233) and the permeability of InitialChecks
(Domain.cpp:1683-1716).

Pins (none restates the oracle's code):
  * momentum: a CalcForce13 pair (SWIType 0) moves m_s a_s + m_w a_w by nothing; a CalcForce12
    pair gives its two particles opposite accelerations (both sides use the fluid's mass);
  * Darcy's law: water streaming through a soil lattice at v feels a = -mu v/(k rho_w) in the
    interior (the SPH sum of the kernel ~ 1);
  * buoyancy: with SWIType 2 soil under hydrostatic water feels +(rho_w/rho_s) g from the pore
    pressure in the interior;
  * the FSI pressure of a solid particle with one fluid neighbour is that neighbour's pressure
    plus the hydrostatic term; with two it is divided by the kernel sum twice (the reference
    queues the particle once per pair);
  * saturation: soil with water above it loses RhoF of density and takes the mass V (rho0 - RhoF);
  * the Kozeny-Carman, Ergun and Den Adel permeabilities of InitialChecks.
"""
import math

import numpy as np
import pytest

from persiansph_amd.abi import default_config, new_particles
from oracle.pyoracle import Oracle


def _pp(n, mat, rho, h, mass, free=True):
    p = new_particles(n)
    p["Material"] = mat
    p["Density"] = rho
    p["RefDensity"] = rho
    p["h"] = h
    p["Mass"] = mass
    p["V"] = mass/rho
    p["IsFree"] = 1 if free else 0
    p["Cs"] = 10.0
    p["PresEq"] = 1
    p["MuRef"] = 1e-3
    p["Mu"] = 1e-3
    return p


def _box(cfg):
    """A 1 m domain around the few particles of a test (Domain::DomMin/DomMax)."""
    cfg.dom_min[0] = cfg.dom_min[1] = 0.0
    cfg.dom_max[0] = cfg.dom_max[1] = 1.0
    return cfg


def _soil(p, k=1e-9, st=0):
    p["SeepageType"] = st
    p["k"] = k
    p["G"] = 1e6
    p["K"] = 2e6
    p["phi"] = 0.5
    return p


def test_calcforce13_conserves_momentum():
    h = 0.013
    s = _soil(_pp(1, 3, 2000.0, h, 0.3))
    w = _pp(1, 1, 1000.0, h, 0.1)
    s["x"][0] = (0.5, 0.5, 0.0)
    w["x"][0] = (0.5 + 0.011, 0.5 + 0.004, 0.0)
    w["v"][0] = (0.3, -0.2, 0.0)
    p = np.concatenate([s, w])
    cfg = _box(default_config(dimension=2, kernel_type=0, scheme=1))
    cfg.swi_type = 0
    o = Oracle(cfg, p, 1e-5)
    st = o.run(1, tout=0.0)
    assert st.pairs == 1
    q = o.particles()
    mom = q["Mass"][0]*q["a"][0] + q["Mass"][1]*q["a"][1]
    assert np.all(np.abs(mom) <= 1e-15*np.abs(q["Mass"][0]*q["a"][0]).max())
    assert np.abs(q["a"][1]).max() > 0.0
    # the water is dragged against its motion relative to the soil
    assert np.dot(q["a"][1], w["v"][0]) < 0.0


def test_calcforce12_opposite_accelerations():
    h = 0.013
    f = _pp(1, 1, 1000.0, h, 0.1)
    s = _pp(1, 2, 1200.0, h, 0.2)
    f["x"][0] = (0.5, 0.5, 0.0)
    s["x"][0] = (0.5 + 0.012, 0.5 - 0.003, 0.0)
    f["v"][0] = (0.1, 0.05, 0.0)
    s["v"][0] = (-0.05, 0.02, 0.0)
    s["G"] = 1e6
    for order in (0, 1):
        p = np.concatenate([f, s] if order == 0 else [s, f])
        cfg = _box(default_config(dimension=2, kernel_type=0, scheme=1))
        cfg.fsi = 1
        o = Oracle(cfg, p, 1e-5)
        o.run(1, tout=0.0)
        q = o.particles()
        np.testing.assert_array_equal(q["a"][0], -q["a"][1])
        assert np.abs(q["a"][0]).max() > 0.0


def _overlap(nx, ny, dx, vw=(0.0, 0.0), swi=0, k=1e-8, gravity=0.0, hydro=False):
    """A soil lattice (fixed) and a water lattice (free) on the same sites, offset by half a
    spacing; the water moves at vw."""
    h = 1.3*dx
    xs = np.array([(i*dx, j*dx, 0.0) for j in range(ny) for i in range(nx)])
    rho_s, rho_w = 2000.0, 1000.0
    s = _soil(_pp(len(xs), 3, rho_s, h, rho_s*dx*dx, free=False), k=k)
    s["x"] = xs
    w = _pp(len(xs), 1, rho_w, h, rho_w*dx*dx)
    w["x"] = xs + np.array([0.5*dx, 0.5*dx, 0.0])
    w["v"][:, 0], w["v"][:, 1] = vw
    w["PresEq"] = 0
    w["Cs"] = 100.0               # nearly incompressible: m/rho of the water stays dx^2
    if hydro:
        H = (ny + 0.5)*dx
        w["Pressure"] = 0.0
        w["P0"] = 0.0
        # linear EOS (PresEq 0: P = P0 + Cs^2 (d - d0)): the density that gives rho g (H - y)
        w["Density"] = rho_w + rho_w*gravity*(H - w["x"][:, 1])/(w["Cs"]**2)
    p = np.concatenate([s, w])
    cfg = default_config(dimension=2, kernel_type=0, scheme=1, gravity=(0.0, -gravity, 0.0))
    cfg.swi_type = swi
    return cfg, p, rho_s, rho_w


def test_darcy_drag_in_the_interior():
    dx, k, v = 0.01, 1e-8, 1e-4
    cfg, p, rho_s, rho_w = _overlap(24, 24, dx, vw=(v, 0.0), k=k)
    o = Oracle(cfg, p, 1e-6)
    o.run(1, tout=0.0)
    q = o.particles()
    w = q["Material"] == 1
    x = q["x"]
    inner = w & (x[:, 0] > 6*dx) & (x[:, 0] < 17*dx) & (x[:, 1] > 6*dx) & (x[:, 1] < 17*dx)
    want = -1e-3*v/(k*rho_w)
    np.testing.assert_allclose(q["a"][inner, 0], want, rtol=0.02)
    assert np.abs(q["a"][inner, 1]).max() <= 0.02*abs(want)


def test_pore_pressure_buoyancy_in_the_interior():
    dx, g = 0.01, 9.81
    cfg, p, rho_s, rho_w = _overlap(24, 24, dx, swi=2, gravity=g, hydro=True, k=1e30)
    o = Oracle(cfg, p, 1e-6)
    o.run(1, tout=0.0)
    q = o.particles()
    s = q["Material"] == 3
    # the soil is fixed, so its acceleration is only the sum of its pair terms plus gravity
    x = q["x"]
    inner = s & (x[:, 0] > 6*dx) & (x[:, 0] < 17*dx) & (x[:, 1] > 6*dx) & (x[:, 1] < 17*dx)
    ay = q["a"][inner, 1] + g
    np.testing.assert_allclose(ay, rho_w*g/rho_s, rtol=0.03)


def _fsi_case(nf):
    h = 0.013
    s = _pp(1, 2, 1200.0, h, 0.2)
    s["G"] = 1e6
    s["x"][0] = (0.5, 0.5, 0.0)
    f = _pp(nf, 1, 1000.0, h, 0.1)
    f["x"][0] = (0.5 + 0.009, 0.5 + 0.007, 0.0)
    if nf > 1:
        f["x"][1] = (0.5 - 0.006, 0.5 - 0.010, 0.0)
    f["Density"] = [1001.0, 1003.0][:nf]
    p = np.concatenate([s, f])
    cfg = _box(default_config(dimension=2, kernel_type=0, scheme=1, gravity=(0.0, -9.81, 0.0)))
    cfg.fsi = 1
    return cfg, p


def _xij_ref(g, x, i, j):
    """x_P1 - x_P2 of the pair (i, j) as the reference's search orients it: P1 is in the cell
    with the smaller key, or within one cell the particle with the larger index (LIFO chains)."""
    ki = [int(np.floor((x[i][c] - g.blpf[c])/g.cell_size[c])) for c in range(2)]
    kj = [int(np.floor((x[j][c] - g.blpf[c])/g.cell_size[c])) for c in range(2)]
    key = lambda k: k[0] + g.cell_no[0]*k[1]
    first_i = key(ki) < key(kj) or (key(ki) == key(kj) and i > j)
    return (x[i] - x[j]) if first_i else (x[j] - x[i])


def test_fsi_pressure_of_one_and_two_neighbours():
    from oracle.pyoracle import lib
    G = -9.81
    cfg, p = _fsi_case(1)
    o = Oracle(cfg, p, 1e-5)
    o.run(1, tout=0.0)
    q, g = o.particles(), o.grid()
    # one pair: the kernel cancels, leaving P_f + G.(x_P1 - x_P2) rho_f
    xij = _xij_ref(g, q["x"], 0, 1)
    assert q["FSIPressure"][0] == pytest.approx(q["Pressure"][1] + (G*xij[1])*q["Density"][1], rel=1e-14)
    cfg, p = _fsi_case(2)
    o = Oracle(cfg, p, 1e-5)
    o.run(1, tout=0.0)
    q, g = o.particles(), o.grid()
    K = [lib().psph_o_kernel(2, 0, float(np.linalg.norm(q["x"][0] - q["x"][j]))/0.013, 0.013) for j in (1, 2)]
    t = [q["Pressure"][j]*K[j-1] + (G*_xij_ref(g, q["x"], 0, j)[1])*q["Density"][j]*K[j-1] for j in (1, 2)]
    assert q["FSISumKernel"][0] == pytest.approx(K[0] + K[1], rel=1e-14)
    assert q["FSIPressure"][0] == pytest.approx((t[0] + t[1])/(K[0] + K[1])**2, rel=1e-13)


def test_saturation_under_water():
    h, dx = 0.013, 0.01
    s = _soil(_pp(1, 3, 2000.0, h, 0.2, free=False))
    s["x"][0] = (0.5, 0.5, 0.0)
    s["RhoF"] = 1000.0
    s["V"] = 1e-4
    w = _pp(1, 1, 1000.0, h, 0.1)
    w["x"][0] = (0.5 + 0.002, 0.5 + dx, 0.0)        # above the soil particle
    p = np.concatenate([s, w])
    cfg = _box(default_config(dimension=2, kernel_type=0, scheme=1))
    cfg.swi_type = 0
    o = Oracle(cfg, p, 1e-5)
    o.run(1, tout=0.0)
    q = o.particles()
    assert q["SatCheck"][0] == 1 and q["IsSat"][0] == 1
    assert q["RefDensity"][0] == 1000.0 and q["Density"][0] == 1000.0
    assert q["Mass"][0] == 1e-4*1000.0
    # water below the soil does not saturate it
    w["x"][0] = (0.5 + 0.002, 0.5 - dx, 0.0)
    o = Oracle(cfg, np.concatenate([s, w]), 1e-5)
    o.run(1, tout=0.0)
    assert o.particles()["IsSat"][0] == 0


@pytest.mark.parametrize("st", [1, 2, 3])
def test_permeability_of_initial_checks(st):
    n0, d = 0.4, 1e-3
    s = _soil(_pp(1, 3, 2000.0, 0.013, 0.2), st=st)
    s["n0"] = n0
    s["d"] = d
    s["k"] = 0.0
    cfg = _box(default_config(dimension=2))
    o = Oracle(cfg, s, 1e-5)
    q = o.particles()
    base = 180.0 if st == 1 else 150.0
    assert q["k"][0] == pytest.approx(n0**3*d**2/(base*(1 - n0)**2), rel=1e-15)
    k2 = {1: 0.0, 2: 1.75*(1 - n0)/(n0**3*d), 3: 0.4/(n0**2*d)}[st]
    assert q["k2"][0] == pytest.approx(k2, rel=1e-15)
    assert q["n"][0] == n0


def test_soil_water_window_of_the_gpu_comparison():
    """The GPU comparison of the soil-water runs covers 20 steps: two oracle runs whose free
    particles start 1 ulp apart agree to 1e-15 m over them, and part by far more by step 30,
    when the bed yields (the return mapping's flow direction amplifies roundoff)."""
    from persiansph_amd import cases
    cfg, p, dt = cases.soil_water_2d(swi_type=0, scheme=0)
    q = p.copy()
    fl = np.where(q["IsFree"] == 1)[0]
    q["x"][fl, 0] = np.nextafter(q["x"][fl, 0], np.inf)
    a, b = Oracle(cfg, p, dt), Oracle(cfg, q, dt)
    a.run(20)
    b.run(20)
    assert np.abs(a.particles()["x"] - b.particles()["x"]).max() <= 1e-15
    a.run(10)
    b.run(10)
    assert np.abs(a.particles()["x"] - b.particles()["x"]).max() > 1e-12


def test_erosion_window_of_the_gpu_comparison():
    """The stirred erosion case (suspended stress-free grains, SWIType 1) is compared over 4
    steps: runs 1 ulp apart agree to 1e-15 m over them and part after."""
    from persiansph_amd import cases
    cfg, p, dt = cases.soil_water_2d(swi_type=1, scheme=0, grains=7, stir=0.2)
    q = p.copy()
    fl = np.where(q["IsFree"] == 1)[0]
    q["x"][fl, 0] = np.nextafter(q["x"][fl, 0], np.inf)
    a, b = Oracle(cfg, p, dt), Oracle(cfg, q, dt)
    a.run(4)
    b.run(4)
    assert np.abs(a.particles()["x"] - b.particles()["x"]).max() <= 1e-15
    a.run(6)
    b.run(6)
    assert np.abs(a.particles()["x"] - b.particles()["x"]).max() > 1e-12
