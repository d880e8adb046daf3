"""Input generation for the benchmark and test workloads.

`add_box_length` restates SPH::Domain::AddBoxLength (Source/Domain.cpp:294-495) for
non-random boxes, and the dam-break builders follow the per-particle setup of the
reference drivers 3-DamBreak.cpp:25-105 and 4-DamBreak3D.cpp:25-126 (walls fixed and
no-slip, fluid above the column deleted, hydrostatic initial density).  Where a workload
asks for lattice jitter, it is drawn from a seeded numpy generator and applied to the fluid
only, identically for every solver the inputs are fed to.
"""
import math

import numpy as np

from .abi import default_config, new_particles


def add_box_length(dim, V, Lx, Ly, Lz, r, density, h, ptype=1, rotation=0, fixed=False):
    """Lattice box of particles in the order AddBoxLength pushes them (Domain.cpp:294-495).

    ptype 1 is cubic packing, ptype 0 hexagonal close packing (rotation 0 only in 3-D).
    Returns a structured array with x, Mass, Density, RefDensity, h, IsFree set.
    """
    V = [float(v) for v in V]
    xs = []
    if dim == 3:
        if ptype == 1:
            k = 0
            zp = V[2]
            while zp <= V[2] + Lz - r:
                j = 0
                yp = V[1]
                while yp <= V[1] + Ly - r:
                    i = 0
                    xp = V[0]
                    while xp <= V[0] + Lx - r:
                        xs.append((V[0] + (2.0*i+1)*r, V[1] + (2.0*j+1)*r, V[2] + (2.0*k+1)*r))
                        i += 1
                        xp = V[0] + (2*i+1)*r
                    j += 1
                    yp = V[1] + (2.0*j+1)*r
                k += 1
                zp = V[2] + (2.0*k+1)*r
        else:
            k = 0
            zp = V[2]
            while zp <= V[2] + Lz - r:
                j = 0
                yp = V[1]
                while yp <= V[1] + Ly - r:
                    i = 0
                    xp = V[0]
                    while xp <= V[0] + Lx - r:
                        if (k % 2 != 0) and (j % 2 != 0):
                            x = V[0] + (2*i+(j % 2)+(k % 2)-1)*r
                        else:
                            x = V[0] + (2*i+(j % 2)+(k % 2)+1)*r
                        y = V[1] + (math.sqrt(3.0)*(j+(1.0/3.0)*(k % 2))+1)*r
                        z = V[2] + ((2*math.sqrt(6.0)/3)*k+1)*r
                        xs.append((x, y, z))
                        i += 1
                        if (k % 2 != 0) and (j % 2 != 0):
                            xp = V[0] + (2*i+(j % 2)+(k % 2)-1)*r
                        else:
                            xp = V[0] + (2*i+(j % 2)+(k % 2)+1)*r
                    j += 1
                    yp = V[1] + (math.sqrt(3.0)*(j+(1.0/3.0)*(k % 2))+1)*r
                k += 1
                zp = V[2] + ((2*math.sqrt(6.0)/3)*k+1)*r
        p = new_particles(len(xs))
        p["x"] = np.array(xs, dtype=np.float64)
        # particle mass in 3-D (Domain.cpp:401-418)
        mx = list(V)
        for q in range(len(xs)):
            for c in range(3):
                if p["x"][q, c] > mx[c]:
                    mx[c] = p["x"][q, c]
        mx = [m + r for m in mx]
        t = [mx[c] - V[c] for c in range(3)]
        p["Mass"] = t[0]*t[1]*t[2]*density/len(xs)
    elif dim == 2:
        masses = []
        if ptype == 0 and rotation == 0:
            j = 0
            yp = V[1]
            while yp <= V[1] + Ly - r:
                i = 0
                xp = V[0]
                while xp <= V[0] + Lx - r:
                    xs.append((V[0] + (2*i+(j % 2)+1)*r, V[1] + (math.sqrt(3.0)*j+1)*r, 0.0))
                    masses.append((math.sqrt(3.0)*r*r)*density)
                    i += 1
                    xp = V[0] + (2*i+(j % 2)+1)*r
                j += 1
                yp = V[1] + (math.sqrt(3.0)*j+1)*r
        elif ptype == 0:
            i = 0
            xp = V[0]
            while xp <= V[0] + Lx - r:
                j = 0
                yp = V[1]
                while yp <= V[1] + Ly - r:
                    xs.append((V[0] + (math.sqrt(3.0)*i+1)*r, V[1] + (2*j+(i % 2)+1)*r, 0.0))
                    masses.append((math.sqrt(3.0)*r*r)*density)
                    j += 1
                    yp = V[1] + (2*j+(i % 2)+1)*r
                i += 1
                xp = V[0] + (math.sqrt(3.0)*i+1)*r
        else:
            j = 0
            yp = V[1]
            while yp <= V[1] + Ly - r:
                i = 0
                xp = V[0]
                while xp <= V[0] + Lx - r:
                    xs.append((V[0] + (2*i+1)*r, V[1] + (2*j+1)*r, 0.0))
                    masses.append(2.0*r*2.0*r*density)
                    i += 1
                    xp = V[0] + (2*i+1)*r
                j += 1
                yp = V[1] + (2*j+1)*r
        p = new_particles(len(xs))
        p["x"] = np.array(xs, dtype=np.float64).reshape(-1, 3)
        p["Mass"] = np.array(masses)
    else:
        raise ValueError("dimension must be 2 or 3")
    p["Density"] = density
    p["RefDensity"] = density
    p["h"] = h
    p["IsFree"] = 0 if fixed else 1
    p["ID"] = 1
    # V = Mass/RefDensity in the Particle constructor (Particle.cpp:88): 2-D boxes construct
    # with their mass, 3-D boxes with mass 0 and set it afterwards (Domain.cpp:401-418)
    p["V"] = p["Mass"]/density if dim == 2 else 0.0
    return p


def _fast_cubic_box(dim, V, counts, r, density, h):
    """Vectorised cubic-packing box for large 3-D cases; identical positions to
    add_box_length(ptype=1) (x fastest, then y, then z, as the reference's loops nest)."""
    nx, ny, nz = counts
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    X = V[0] + (2.0*i+1)*r
    Y = V[1] + (2.0*j+1)*r
    Z = V[2] + (2.0*k+1)*r
    n = nx*ny*nz
    p = new_particles(n)
    x = np.empty((n, 3))
    x[:, 0] = np.tile(X, ny*nz)
    x[:, 1] = np.tile(np.repeat(Y, nx), nz)
    x[:, 2] = np.repeat(Z, nx*ny)
    p["x"] = x
    mx = [max(V[c], x[:, c].max()) + r for c in range(3)]
    t = [mx[c] - V[c] for c in range(3)]
    p["Mass"] = t[0]*t[1]*t[2]*density/n
    p["Density"] = density
    p["RefDensity"] = density
    p["h"] = h
    p["ID"] = 1
    return p


def _loop_count(v0, L, r):
    """Number of lattice sites the reference's `while (xp <= V+L-r)` loop visits
    (xp starts at V, then V + (2i+1) r; Domain.cpp:371-384)."""
    n = 0
    xp = v0
    while xp <= v0 + L - r:
        n += 1
        xp = v0 + (2*n+1)*r
    return n


def dam_break(dim, dx, tank, column, h_factor=1.3, rho0=1000.0, g=9.81, mu=1.0e-3,
              alpha=0.05, cs=None, scheme=1, layers=3, jitter=0.0, seed=12345,
              precision=0, kernel_type=0):
    """Weakly-compressible dam break following 3-DamBreak.cpp / 4-DamBreak3D.cpp.

    tank = (TL, TH) in 2-D or (TL, TH, TW) in 3-D (length x, height y, width z);
    column = (L, H) or (L, H, W).  Walls are `layers` lattice layers outside the tank
    on the left, right, bottom (and front/back in 3-D), no lid.  Returns
    (config, particles, dt) with dt = 0.2 h / c0.
    """
    H = column[1]
    h = h_factor*dx
    if cs is None:
        cs = 10.0*math.sqrt(2.0*g*H)
    r = dx/2.0
    lay = float(layers)
    if dim == 2:
        TL, TH = tank
        L = column[0]
        V = (-lay*dx, -lay*dx, 0.0)
        p = add_box_length(2, V, 2*lay*dx + TL + dx/10.0, lay*dx + TH + dx/10.0, 0.0, r, rho0, h, 1, 0, False)
    else:
        TL, TH, TW = tank
        L, H, W = column
        V = (-lay*dx, -lay*dx, -lay*dx)
        Ls = (2*lay*dx + TL + dx/10.0, lay*dx + TH + dx/10.0, 2*lay*dx + TW + dx/10.0)
        counts = tuple(_loop_count(V[c], Ls[c], r) for c in range(3))
        p = _fast_cubic_box(3, V, counts, r, rho0, h)
    x = p["x"]
    wall = (x[:, 0] < 0.0) | (x[:, 1] < 0.0) | (x[:, 0] > TL)
    if dim == 3:
        wall |= (x[:, 2] < 0.0) | (x[:, 2] > TW)
    fluid_keep = ~wall & (x[:, 1] <= H) & (x[:, 0] <= L)
    if dim == 3:
        fluid_keep &= (x[:, 2] <= W)
    keep = wall | fluid_keep          # DelParticles(11), order preserving (array.h:726-738)
    p = p[keep]
    wall = wall[keep]
    p["Cs"] = cs
    p["PresEq"] = 1
    p["Mu"] = mu
    p["MuRef"] = mu
    p["Material"] = 1
    p["Shepard"] = 1
    p["Alpha"] = alpha
    p["IsFree"] = np.where(wall, 0, 1)
    p["NoSlip"] = np.where(wall, 1, 0)
    p["ID"] = np.where(wall, 2, 1)
    fl = ~wall
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter*dx, jitter*dx, size=(int(fl.sum()), 3))
        if dim == 2:
            d[:, 2] = 0.0
        p["x"][fl] = p["x"][fl] + d
    yb = p["x"][fl, 1]
    # scalar libm pow, as the reference drivers' std::pow (numpy's vectorised power may differ
    # in the last bit)
    base = 1.0 + 7.0*g*(H - yb)/(cs*cs)
    rho = rho0*np.array([math.pow(b, 1.0/7.0) for b in base.tolist()])
    p["Density"][fl] = rho
    p["Densityb"][fl] = rho
    cfg = default_config(dimension=dim, kernel_type=kernel_type, visc_eq=0, gradient_type=0,
                         scheme=scheme, precision=precision, gravity=(0.0, -g, 0.0))
    dt = 0.2*h/cs
    return cfg, p, dt


def dam_break_2d(dx=0.02, scheme=1, **kw):
    """2-D dam break: tank 4.0 x 3.0 m, column 1.0 x 2.0 m, 3 wall layers."""
    return dam_break(2, dx, (4.0, 3.0), (1.0, 2.0), scheme=scheme, **kw)


def dam_break_3d(dx=0.01, scheme=1, jitter=0.01, seed=12345, tank=(3.2, 1.0, 1.0),
                 column=(1.2, 0.6, 1.0), cs=10.0*math.sqrt(2.0*9.81*2.0), **kw):
    """3-D dam break: tank 3.2 x 1.0 x 1.0 m (x, y up, z), column 1.2 long, 0.6 high and
    1.0 wide, 3 wall layers, no lid, fluid jitter +-0.01 dx.  The sound speed is the 2-D
    case's c0 = 10 sqrt(2 g 2.0) = 62.64 m/s, so dt = 0.2 h / c0 = 4.15e-5 s at dx = 0.01."""
    return dam_break(3, dx, tank, column, scheme=scheme, jitter=jitter, seed=seed, cs=cs, **kw)


def _lattice(dim, V, counts, r, density, h):
    """Cubic lattice box, x fastest (AddBoxLength order); 2-D boxes have z = 0 and mass 4 r^2 rho."""
    if dim == 3:
        return _fast_cubic_box(3, V, counts, r, density, h)
    nx, ny = counts[0], counts[1]
    p = _fast_cubic_box(3, (V[0], V[1], 0.0), (nx, ny, 1), r, density, h)
    p["x"][:, 2] = 0.0
    p["Mass"] = 2.0*r*2.0*r*density
    return p


def granular_column(dim=3, dx=0.004, size=(0.2, 0.4, 0.2), floor_half=0.5, rho=2650.0,
                    phi_deg=30.0, cohesion=0.0, psi_deg=0.0, E=5.0e6, nu=0.3, h_factor=1.5,
                    cs=None, ti=0.5, tin=2.55, alpha=0.1, beta=0.1, scheme=1, layers=3,
                    g=9.81, precision=0, fail=3, geostatic=True):
    """Granular column collapse with the Drucker-Prager soil model (Material 3, Fail 3).

    The column (size = (Lx, Ly up, Lz)) stands on a rigid floor of `layers` lattice layers
    extending floor_half beyond the column centre in x (and z in 3-D).  Parameters follow
    Old_Tests/Dam_Scouring.cpp:187-202 (Fail 3, TI 0.5, TIn 2.55, TIInitDist = dx) and
    Old_Tests/ReposeAngle.cpp:93-99 (alpha = beta = 0.1); K = E/(3(1-2nu)), G = E/(2(1+nu)).
    The sound speed for artificial viscosity defaults to sqrt(K/rho) and dt = 0.2 h/Cs.
    The initial stress is geostatic: sigma_yy = -rho g (Ly - y), lateral K0 = nu/(1-nu).
    """
    h = h_factor*dx
    r = dx/2.0
    K = E/(3.0*(1.0-2.0*nu))
    G = E/(2.0*(1.0+nu))
    if cs is None:
        cs = math.sqrt(K/rho)
    Lx, Ly, Lz = size
    nx, ny, nz = int(round(Lx/dx)), int(round(Ly/dx)), int(round(Lz/dx))
    soil = _lattice(dim, (0.0, 0.0, 0.0), (nx, ny, nz), r, rho, h)
    nf = int(round(2*floor_half/dx))
    fx0 = Lx/2.0 - nf*dx/2.0
    if dim == 3:
        fz0 = Lz/2.0 - nf*dx/2.0
        fl = _lattice(3, (fx0, -layers*dx, fz0), (nf, layers, nf), r, rho, h)
    else:
        fl = _lattice(2, (fx0, -layers*dx, 0.0), (nf, layers, 1), r, rho, h)
    fl["IsFree"] = 0
    fl["ID"] = 2
    p = np.concatenate([fl, soil])
    p["Material"] = 3
    p["Fail"] = fail
    p["phi"] = math.radians(phi_deg)
    p["psi"] = math.radians(psi_deg)
    p["c"] = cohesion
    p["K"] = K
    p["G"] = G
    p["Cs"] = cs
    p["Alpha"] = alpha
    p["Beta"] = beta
    p["TI"] = ti
    p["TIn"] = tin
    p["TIInitDist"] = dx
    if geostatic:
        free = p["IsFree"] == 1
        y = p["x"][:, 1]
        syy = np.where(free, -rho*g*(Ly - y), 0.0)
        k0 = nu/(1.0-nu)
        sig = np.zeros((len(p), 9))
        sig[:, 4] = syy
        sig[:, 0] = k0*syy
        if dim == 3:
            sig[:, 8] = k0*syy
        p["Sigma"] = sig
    cfg = default_config(dimension=dim, kernel_type=0, visc_eq=0, gradient_type=0,
                         scheme=scheme, precision=precision, gravity=(0.0, -g, 0.0))
    dt = 0.2*h/cs
    return cfg, p, dt


def elastic_block(dim=2, dx=0.01, size=(0.06, 0.1, 0.06), floor_half=0.08, rho=1200.0, E=1.0e6,
                  nu=0.3, v0=-0.5, fail=0, sigmay=5.0e3, ti=0.3, alpha=0.1, beta=0.0, scheme=1,
                  layers=3, g=9.81):
    """An elastic solid block (Material 2) dropped with velocity v0 (y) onto a rigid floor of
    `layers` fixed Material 2 layers: the impact loads the block well beyond its weight.

    Linear EOS (PresEq 0) with Cs = sqrt(K/rho), K = E/(3(1-2nu)), G = E/(2(1+nu)); Fail 1 adds
    the von Mises scale-back at Sigmay (Particle.cpp:293-300, 613-620).  Artificial stress TI
    with TIn = 2.55 and TIInitDist = dx, artificial viscosity alpha/beta, dt = 0.2 h/Cs."""
    h = 1.3*dx
    r = dx/2.0
    K = E/(3.0*(1.0-2.0*nu))
    G = E/(2.0*(1.0+nu))
    cs = math.sqrt(K/rho)
    Lx, Ly, Lz = size
    nx, ny, nz = int(round(Lx/dx)), int(round(Ly/dx)), int(round(Lz/dx))
    block = _lattice(dim, (0.0, 0.0, 0.0), (nx, ny, nz), r, rho, h)
    nf = int(round(2*floor_half/dx))
    fx0 = Lx/2.0 - nf*dx/2.0
    if dim == 3:
        fz0 = Lz/2.0 - nf*dx/2.0
        fl = _lattice(3, (fx0, -layers*dx, fz0), (nf, layers, nf), r, rho, h)
    else:
        fl = _lattice(2, (fx0, -layers*dx, 0.0), (nf, layers, 1), r, rho, h)
    fl["IsFree"] = 0
    fl["ID"] = 2
    block["v"][:, 1] = v0
    p = np.concatenate([fl, block])
    p["Material"] = 2
    p["Fail"] = fail
    p["Sigmay"] = sigmay
    p["PresEq"] = 0
    p["K"] = K
    p["G"] = G
    p["Cs"] = cs
    p["Alpha"] = alpha
    p["Beta"] = beta
    p["TI"] = ti
    p["TIn"] = 2.55
    p["TIInitDist"] = dx
    cfg = default_config(dimension=dim, kernel_type=0, visc_eq=0, gradient_type=0,
                         scheme=scheme, gravity=(0.0, -g, 0.0))
    dt = 0.2*h/cs
    return cfg, p, dt


def runout(p):
    """Runout of a granular column: the largest x reached by a free particle."""
    free = p["IsFree"] == 1
    return float(p["x"][free, 0].max())


def poiseuille_2d(dx=2.5e-5, body_force=0.002, rho=998.21, mu=1.002e-3, cs=0.07, nx=20, ny=48):
    """Plane Poiseuille flow as 1-Poiseuille.cpp:36-80 sets it up: a cubic box of nx x ny
    particles from (0, -ny/2 dx), periodic in x; particles with |y| >= 20 dx are fixed
    no-slip walls (ID 2); quintic spline, Takeda viscosity, squared-density form, modified
    Verlet, linear EOS, h = 1.1 dx, dt = 0.2 h / Cs.  The flow is driven by the body force
    of its GeneralBefore hook (a += (body_force, 0, 0) on free particles), which goes to
    Solver.set_body_force.  Returns (config, particles, dt, body)."""
    h = 1.1*dx
    r = dx/2.0
    p = add_box_length(2, (0.0, -(ny/2)*dx, 0.0), nx*dx + dx/10.0, ny*dx + dx/10.0, 0.0, r, rho, h, 1, 0, False)
    p["Cs"] = cs
    p["PresEq"] = 0
    p["Mu"] = mu
    p["MuRef"] = mu
    p["Material"] = 1
    p["ID"] = 1
    yb = p["x"][:, 1]
    wall = (yb >= 20.0*dx) | (yb <= -20.0*dx)
    p["ID"][wall] = 2
    p["IsFree"][wall] = 0
    p["NoSlip"][wall] = 1
    cfg = default_config(dimension=2, kernel_type=2, visc_eq=3, gradient_type=0, scheme=0)
    cfg.periodic[0] = 1
    cfg.R = r
    dt = 0.2*h/cs
    return cfg, p, dt, (body_force, 0.0, 0.0)


def couette_2d(dx=2.5e-5, v_lid=2.5e-5, rho=998.21, mu=1.002e-3, cs=0.08, nx=20, ny=48):
    """Plane Couette flow as 2-Couette.cpp:33-91 sets it up: a cubic box of nx x ny particles
    from (0, -4 dx), periodic in x; particles with y >= 40 dx form the moving no-slip lid
    (ID 3, v = (v_lid, 0, 0)), those with y < 0 the fixed no-slip floor (ID 2); LES on with
    CSmag = 0 (the shear rate is formed, the viscosity stays MuRef); quintic spline, Takeda
    viscosity, modified Verlet, linear EOS, h = 1.1 dx, dt = 0.2 h / Cs."""
    h = 1.1*dx
    r = dx/2.0
    p = add_box_length(2, (0.0, -4.0*dx, 0.0), nx*dx + dx/10.0, ny*dx + dx/10.0, 0.0, r, rho, h, 1, 0, False)
    p["LES"] = 1
    p["CSmag"] = 0.0
    p["Cs"] = cs
    p["PresEq"] = 0
    p["Mu"] = mu
    p["MuRef"] = mu
    p["Material"] = 1
    p["ID"] = 1
    yb = p["x"][:, 1]
    lid = yb >= 40.0*dx
    floor = yb < 0.0
    p["ID"][lid] = 3
    p["IsFree"][lid] = 0
    p["NoSlip"][lid] = 1
    p["v"][lid, 0] = v_lid
    p["ID"][floor] = 2
    p["IsFree"][floor] = 0
    p["NoSlip"][floor] = 1
    cfg = default_config(dimension=2, kernel_type=2, visc_eq=3, gradient_type=0, scheme=0)
    cfg.periodic[0] = 1
    cfg.R = r
    dt = 0.2*h/cs
    return cfg, p, dt


def _tank_2d(dx, TL, TH, layers=3):
    """Cubic lattice covering a 2-D tank and `layers` wall layers left, right and below."""
    lay = float(layers)
    return add_box_length(2, (-lay*dx, -lay*dx, 0.0), 2*lay*dx + TL + dx/10.0, lay*dx + TH + dx/10.0,
                          0.0, dx/2.0, 1000.0, 1.3*dx, 1, 0, False)


def soil_water_2d(dx=0.01, swi_type=0, k=1e-6, rho_s=2000.0, rho_w=1000.0, rho_f=1000.0, scheme=0,
                  seepage_type=0, fsi=False, grains=0, stir=0.0, seed=3):
    """A water column collapsing onto a soil bed in a 2-D tank (the setting of the reference's
    Old_Tests/Dam_Scouring*.cpp): tank 0.6 x 0.4 m with 3 layers of fixed no-slip water walls
    (Material 1), a soil bed 0.1 m deep on a fixed soil floor (Material 3, Drucker-Prager,
    phi = 30 deg, c = 0, E = 5 MPa, nu = 0.3, TI 0.5), a water column 0.2 x 0.2 m on the bed at
    the left, the bed under geostatic stress.  Soil-water interaction SWIType `swi_type` with
    Darcy permeability k (or the seepage type with n0 = 0.4, d = 1 mm), RhoF = rho_f for the
    saturation of soil under water.  grains > 0 turns every grains-th water particle into a
    free soil grain suspended in the column (few soil neighbours: the surface-erosion branch of
    SWIType 1); stir > 0 gives the water seeded random velocities up to stir m/s.
    Cubic spline, h = 1.3 dx, water Cs = 10 sqrt(2 g 0.3), soil Cs = sqrt(K/rho_s)."""
    g = 9.81
    TL, TH, Hs, Lw, Hw = 0.6, 0.4, 0.1, 0.2, 0.2
    p = _tank_2d(dx, TL, TH)
    x = p["x"]
    wall = (x[:, 0] < 0.0) | (x[:, 0] > TL) | (x[:, 1] < 0.0)
    soil = ~wall & (x[:, 1] <= Hs)
    water = ~wall & ~soil & (x[:, 1] <= Hs + Hw) & (x[:, 0] <= Lw)
    floor = (x[:, 1] < 0.0) & (x[:, 0] >= 0.0) & (x[:, 0] <= TL)
    keep = wall | soil | water
    p, wall, soil, water, floor = p[keep], wall[keep], soil[keep], water[keep], floor[keep]
    if grains > 0:
        wi = np.where(water)[0]
        g_idx = wi[::grains]
        soil[g_idx] = True
        water[g_idx] = False
    E, nu = 5.0e6, 0.3
    K = E/(3.0*(1.0 - 2.0*nu))
    G = E/(2.0*(1.0 + nu))
    cs_w = 10.0*math.sqrt(2.0*g*0.3)
    cs_s = math.sqrt(K/rho_s)
    sm = soil | floor
    p["Material"] = np.where(sm, 3, 1)
    p["IsFree"] = np.where(wall, 0, 1)
    p["NoSlip"] = np.where(wall & ~floor, 1, 0)
    p["ID"] = np.where(wall, 2, np.where(soil, 3, 1))
    p["Density"] = np.where(sm, rho_s, rho_w)
    p["RefDensity"] = p["Density"]
    p["Mass"] = p["Density"]*dx*dx
    p["V"] = p["Mass"]/p["RefDensity"]
    p["Cs"] = np.where(sm, cs_s, cs_w)
    p["PresEq"] = np.where(sm, 0, 1)
    p["Mu"] = np.where(sm, 0.0, 1e-3)
    p["MuRef"] = p["Mu"]
    p["Alpha"] = np.where(sm, 0.1, 0.05)
    p["Beta"] = np.where(sm, 0.1, 0.0)
    p["G"] = np.where(sm, G, 0.0)
    p["K"] = np.where(sm, K, 0.0)
    p["phi"] = np.where(sm, math.radians(30.0), 0.0)
    p["Fail"] = np.where(sm, 3, 0)
    p["TI"] = np.where(soil, 0.5, 0.0)
    p["TIn"] = np.where(soil, 2.55, 4.0)
    p["TIInitDist"] = np.where(soil, dx, 0.0)
    p["SeepageType"] = np.where(sm, seepage_type, 0)
    p["k"] = np.where(sm, k, 0.0)
    p["n0"] = np.where(sm, 0.4, 0.0)
    p["d"] = np.where(sm, 1e-3, 0.0)
    p["RhoF"] = np.where(sm, rho_f, 0.0)
    # geostatic stress of the bed under its own weight and the water column's (away from the
    # stress-free apex of the yield surface, where the flow direction is made of roundoff)
    y, xx = p["x"][:, 1], p["x"][:, 0]
    syy = np.where(soil, -rho_s*g*(Hs - y) - np.where(xx <= Lw, rho_w*g*Hw, 0.0), 0.0)
    sig = np.zeros((len(p), 9))
    sig[:, 4] = syy
    sig[:, 0] = nu/(1.0 - nu)*syy
    p["Sigma"] = sig
    if stir > 0.0:
        rng = np.random.default_rng(seed)
        v = rng.uniform(-stir, stir, size=(int(water.sum()), 3))
        v[:, 2] = 0.0
        p["v"][water] = v
    cfg = default_config(dimension=2, kernel_type=0, scheme=scheme, gravity=(0.0, -g, 0.0))
    cfg.swi_type = swi_type
    cfg.fsi = 1 if fsi else 0
    dt = 0.2*1.3*dx/max(cs_w, cs_s)
    return cfg, p, dt


def fsi_2d(dx=0.01, scheme=0, E=1.0e6, rho_s=1200.0):
    """Fluid-structure interaction in a 2-D tank (FSI on): a water column 0.2 x 0.3 m at the
    left of a 0.6 x 0.4 m tank with fixed no-slip walls (Material 1), and a free elastic block
    (Material 2, E, nu = 0.3, no failure) 0.06 x 0.15 m standing on the floor against the
    column's right face.
    Cubic spline, h = 1.3 dx, water Cs = 10 sqrt(2 g 0.3), solid Cs = sqrt(K/rho_s)."""
    g = 9.81
    TL, TH = 0.6, 0.4
    p = _tank_2d(dx, TL, TH)
    x = p["x"]
    wall = (x[:, 0] < 0.0) | (x[:, 0] > TL) | (x[:, 1] < 0.0)
    water = ~wall & (x[:, 0] <= 0.2) & (x[:, 1] <= 0.3)
    solid = ~wall & (x[:, 0] > 0.2) & (x[:, 0] <= 0.26) & (x[:, 1] <= 0.15)
    keep = wall | water | solid
    p, wall, water, solid = p[keep], wall[keep], water[keep], solid[keep]
    nu = 0.3
    K = E/(3.0*(1.0 - 2.0*nu))
    G = E/(2.0*(1.0 + nu))
    cs_w = 10.0*math.sqrt(2.0*g*0.3)
    cs_s = math.sqrt(K/rho_s)
    p["Material"] = np.where(solid, 2, 1)
    p["IsFree"] = np.where(wall, 0, 1)
    p["NoSlip"] = np.where(wall, 1, 0)
    p["ID"] = np.where(wall, 2, np.where(solid, 4, 1))
    p["Density"] = np.where(solid, rho_s, 1000.0)
    p["RefDensity"] = p["Density"]
    p["Mass"] = p["Density"]*dx*dx
    p["V"] = p["Mass"]/p["RefDensity"]
    p["Cs"] = np.where(solid, cs_s, cs_w)
    p["PresEq"] = np.where(solid, 0, 1)
    p["Mu"] = np.where(solid, 0.0, 1e-3)
    p["MuRef"] = p["Mu"]
    p["Alpha"] = np.where(solid, 0.1, 0.05)
    p["G"] = np.where(solid, G, 0.0)
    p["K"] = np.where(solid, K, 0.0)
    cfg = default_config(dimension=2, kernel_type=0, scheme=scheme, gravity=(0.0, -g, 0.0))
    cfg.fsi = 1
    dt = 0.2*1.3*dx/max(cs_w, cs_s)
    return cfg, p, dt


def channel_inflow_2d(dx=0.01, L=0.5, H=0.12, v_in=1.0, cs=20.0, mu=1e-3, in_out_flow=3, scheme=0,
                      out_v=0.0, out_density=0.0, mass_conservation=False):
    """A 2-D channel with in/out-flow boundaries (Boundary::InOutFlow): fluid filling a channel
    L x H between three layers of fixed no-slip walls above and below, the inflow zone at the
    left imposing v = (v_in, 0, 0) and the reference density 1000, the outflow zone at the right
    (out_v, out_density; 0 keeps the particles' own), particles that leave recycled into the
    inflow.  Cubic spline, h = 1.3 dx, linear EOS with Cs, Morris viscosity mu."""
    lay = 3
    p = add_box_length(2, (0.0, -lay*dx, 0.0), L + dx/10.0, H + 2*lay*dx + dx/10.0, 0.0, dx/2.0, 1000.0,
                       1.3*dx, 1, 0, False)
    y = p["x"][:, 1]
    wall = (y < 0.0) | (y > H)
    p["IsFree"] = np.where(wall, 0, 1)
    p["NoSlip"] = np.where(wall, 1, 0)
    p["ID"] = np.where(wall, 2, 1)
    p["Material"] = 1
    p["Cs"] = cs
    p["PresEq"] = 0
    p["Mu"] = mu
    p["MuRef"] = mu
    p["v"][~wall, 0] = v_in
    cfg = default_config(dimension=2, kernel_type=0, scheme=scheme)
    cfg.in_out_flow = in_out_flow
    cfg.in_v[0] = v_in
    cfg.in_density = 1000.0
    cfg.out_v[0] = out_v
    cfg.out_density = out_density
    cfg.mass_conservation = 1 if mass_conservation else 0
    cfg.initial_dist = dx
    dt = 0.2*1.3*dx/cs
    return cfg, p, dt
