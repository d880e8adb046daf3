"""Python mirror of the C ABI data types in include/persiansph_amd.h.

`PARTICLE_DTYPE` is a numpy structured dtype laid out exactly like `psph_particle`
(FP64 fields first, then int32 flags), so a numpy array of it can be handed to the C ABI
as a plain pointer.  Defaults follow the SPH::Particle constructor
(Source/Particle.cpp:25-117) and the SPH::Domain constructor (Source/Domain.cpp:54-102).
"""
import ctypes

import numpy as np

_D = np.float64
_I = np.int32

_PARTICLE_FIELDS = [
    ("x", _D, 3), ("v", _D, 3), ("va", _D, 3), ("vb", _D, 3), ("a", _D, 3),
    ("NSv", _D, 3), ("VXSPH", _D, 3),
    ("Density", _D, 1), ("Densitya", _D, 1), ("Densityb", _D, 1), ("dDensity", _D, 1),
    ("Pressure", _D, 1), ("Mass", _D, 1),
    ("ZWab", _D, 1), ("SumDen", _D, 1), ("SumKernel", _D, 1),
    ("Mu", _D, 1), ("ShearRate", _D, 1), ("SBar", _D, 1),
    ("StrainRate", _D, 9), ("RotationRate", _D, 9),
    ("ShearStress", _D, 9), ("ShearStressa", _D, 9), ("ShearStressb", _D, 9),
    ("Sigma", _D, 9), ("Sigmaa", _D, 9), ("Sigmab", _D, 9),
    ("Strain", _D, 9), ("Straina", _D, 9), ("Strainb", _D, 9),
    ("TIR", _D, 9),
    ("h", _D, 1), ("Cs", _D, 1), ("P0", _D, 1), ("RefDensity", _D, 1), ("FPMassC", _D, 1),
    ("Alpha", _D, 1), ("Beta", _D, 1), ("MuRef", _D, 1), ("T0", _D, 1), ("m", _D, 1), ("CSmag", _D, 1),
    ("TI", _D, 1), ("TIn", _D, 1), ("TIInitDist", _D, 1),
    ("G", _D, 1), ("K", _D, 1), ("Sigmay", _D, 1), ("c", _D, 1), ("phi", _D, 1), ("psi", _D, 1),
    ("FSIPressure", _D, 1), ("FSISumKernel", _D, 1), ("FSINSv", _D, 3),
    ("n", _D, 1), ("n0", _D, 1), ("k", _D, 1), ("k2", _D, 1), ("d", _D, 1), ("V", _D, 1),
    ("RhoF", _D, 1), ("S", _D, 1),
    ("IsFree", _I, 1), ("NoSlip", _I, 1), ("Shepard", _I, 1), ("LES", _I, 1),
    ("Material", _I, 1), ("PresEq", _I, 1), ("VisM", _I, 1), ("Fail", _I, 1),
    ("ID", _I, 1), ("ct", _I, 1), ("FirstStep", _I, 1), ("InOut", _I, 1),
    ("ShepardCounter", _I, 1), ("ShepardStep", _I, 1),
    ("SeepageType", _I, 1), ("IsSat", _I, 1), ("SatCheck", _I, 1), ("reserved2", _I, 1),
    ("Index", _I, 1), ("reserved", _I, 1),
]

PARTICLE_DTYPE = np.dtype([(n, t, (k,)) if k > 1 else (n, t) for n, t, k in _PARTICLE_FIELDS],
                          align=True)
PARTICLE_FIELDS = [(n, k) for n, _, k in _PARTICLE_FIELDS]
assert PARTICLE_DTYPE.itemsize == 174 * 8 + 20 * 4, PARTICLE_DTYPE.itemsize

FP64 = 0
FP32_MIXED = 1

STOP_STEPS, STOP_TEND, STOP_OUTPUT, STOP_HOOK_BEFORE, STOP_HOOK_AFTER = 0, 1, 2, 3, 4

ERRORS = {
    0: "PSPH_OK", -1: "PSPH_ERR_INVALID", -2: "PSPH_ERR_UNSUPPORTED", -3: "PSPH_ERR_HIP",
    -4: "PSPH_ERR_TIMESTEP", -5: "PSPH_ERR_CFL", -6: "PSPH_ERR_NO_DEVICE",
    -7: "PSPH_ERR_STATE", -8: "PSPH_ERR_CAPACITY",
}


class Config(ctypes.Structure):
    """psph_config: the Domain fields read inside Solve (Source/Domain.h:89-164)."""
    _fields_ = [
        ("dimension", ctypes.c_int32), ("kernel_type", ctypes.c_int32),
        ("visc_eq", ctypes.c_int32), ("gradient_type", ctypes.c_int32),
        ("scheme", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("periodic", ctypes.c_int32 * 3), ("in_out_flow", ctypes.c_int32),
        ("swi_type", ctypes.c_int32), ("fsi", ctypes.c_int32),
        ("cellfac", ctypes.c_double), ("gravity", ctypes.c_double * 3),
        ("xsph", ctypes.c_double), ("sqrt_h_a", ctypes.c_double),
        ("dom_max", ctypes.c_double * 3), ("dom_min", ctypes.c_double * 3),
        ("trpr", ctypes.c_double * 3), ("blpf", ctypes.c_double * 3),
        ("R", ctypes.c_double), ("mu_max", ctypes.c_double), ("cs_max", ctypes.c_double),
        ("in_v", ctypes.c_double * 3), ("out_v", ctypes.c_double * 3),
        ("in_density", ctypes.c_double), ("out_density", ctypes.c_double),
        ("bc_cellfac", ctypes.c_double), ("initial_dist", ctypes.c_double),
        ("mass_conservation", ctypes.c_int32), ("reserved_io", ctypes.c_int32),
    ]


class Status(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("deltat", ctypes.c_double), ("deltatmin", ctypes.c_double),
        ("steps", ctypes.c_int64), ("pairs", ctypes.c_int64), ("deleted", ctypes.c_int64),
        ("n", ctypes.c_size_t), ("stop_reason", ctypes.c_int32), ("mid_step", ctypes.c_int32),
        ("list_builds", ctypes.c_int64),
    ]


class OutputF32(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_size_t),
        ("position", ctypes.c_void_p), ("velocity", ctypes.c_void_p), ("acceleration", ctypes.c_void_p),
        ("pressure", ctypes.c_void_p), ("density", ctypes.c_void_p), ("mass", ctypes.c_void_p),
        ("h", ctypes.c_void_p), ("tag", ctypes.c_void_p), ("sigma", ctypes.c_void_p),
        ("strain", ctypes.c_void_p),
    ]


# datasets of Domain::WriteXDMF (Domain.cpp:1934-1963) held by psph_output_f32: name, type, width
OUTPUT_FIELDS = [("position", "f4", 3), ("velocity", "f4", 3), ("acceleration", "f4", 3),
                 ("pressure", "f4", 1), ("density", "f4", 1), ("mass", "f4", 1), ("h", "f4", 1),
                 ("tag", "i4", 1), ("sigma", "f4", 6), ("strain", "f4", 6)]


class Grid(ctypes.Structure):
    _fields_ = [
        ("blpf", ctypes.c_double * 3), ("trpr", ctypes.c_double * 3),
        ("cell_size", ctypes.c_double * 3),
        ("hmax", ctypes.c_double), ("rhomax", ctypes.c_double),
        ("mu_max", ctypes.c_double), ("cs_max", ctypes.c_double),
        ("cell_no", ctypes.c_int32 * 3), ("reserved", ctypes.c_int32),
    ]


def default_config(dimension=2, kernel_type=0, visc_eq=0, gradient_type=0, scheme=0,
                   precision=FP64, gravity=(0.0, 0.0, 0.0), xsph=0.0, sqrt_h_a=0.0025):
    """A psph_config with the Domain constructor's defaults (Domain.cpp:54-102)."""
    c = Config()
    c.dimension = dimension
    c.kernel_type = kernel_type
    c.visc_eq = visc_eq
    c.gradient_type = gradient_type
    c.scheme = scheme
    c.precision = precision
    c.cellfac = 3.0 if kernel_type == 2 else 2.0          # Kernel_Set, Domain.cpp:117-121
    for k in range(3):
        c.gravity[k] = gravity[k]
        c.dom_max[k] = -100000000000.0
        c.dom_min[k] = 100000000000.0
    c.xsph = xsph
    c.sqrt_h_a = sqrt_h_a
    c.bc_cellfac = 3.9                                     # Boundary::Boundary, Boundary_Condition.cpp:39
    return c


def new_particles(n):
    """n particles with the Particle constructor's defaults (Particle.cpp:25-117).  The
    constructor also sets V = Mass/RefDensity from its arguments; set V with the masses."""
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["TIn"] = 4.0
    p["FPMassC"] = 1.0
    p["IsFree"] = 1
    p["m"] = 300.0
    p["FirstStep"] = 1
    p["ShepardStep"] = 40
    p["CSmag"] = 0.17
    return p


def grid_dict(g):
    return {
        "blpf": list(g.blpf), "trpr": list(g.trpr), "cell_size": list(g.cell_size),
        "hmax": g.hmax, "rhomax": g.rhomax, "mu_max": g.mu_max, "cs_max": g.cs_max,
        "cell_no": list(g.cell_no),
    }
