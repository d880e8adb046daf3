"""ctypes binding of the CPU oracle (oracle/libpsph_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product package never imports this module.
"""
import ctypes
import os

import numpy as np

from persiansph_amd.abi import PARTICLE_DTYPE, Config, Grid, Status

_HERE = os.path.dirname(os.path.abspath(__file__))
# PSPH_ORACLE_LIB points the binding at another build of the oracle (the mutant builds of
# tests/test_oracle_mutants.py); by default it loads the library `make -C oracle` builds
LIB_PATH = os.environ.get("PSPH_ORACLE_LIB") or os.path.join(_HERE, "libpsph_oracle.so")

_lib = None

# void (*)(psph_particle *p, size_t n, void *user)
HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        for name in ("psph_o_kernel", "psph_o_grad_kernel", "psph_o_laplace_kernel", "psph_o_secderiv_kernel"):
            f = getattr(L, name)
            f.restype = d
            f.argtypes = [i, i, d, d]
        L.psph_o_eos.restype = d
        L.psph_o_eos.argtypes = [i, d, d, d, d]
        L.psph_o_sound_speed.restype = d
        L.psph_o_sound_speed.argtypes = [i, d, d, d]
        L.psph_o_density_solid.restype = d
        L.psph_o_density_solid.argtypes = [i, d, d, d, d]
        L.psph_o_viscous_force.restype = None
        L.psph_o_viscous_force.argtypes = [i, p, d, d, d, d, p, i, i, d, d, p, p]
        L.psph_o_eig_sym3.restype = None
        L.psph_o_eig_sym3.argtypes = [p, p, p]
        L.psph_o_tir.restype = None
        L.psph_o_tir.argtypes = [i, p, d, d, p]
        L.psph_o_particle_move.restype = i
        L.psph_o_particle_move.argtypes = [p, d, i, d]
        L.psph_o_create.restype = i
        L.psph_o_create.argtypes = [ctypes.POINTER(Config), p, ctypes.c_size_t, d, d, i, i, ctypes.POINTER(p)]
        L.psph_o_run.restype = i
        L.psph_o_run.argtypes = [p, ctypes.c_int64, d, d, ctypes.POINTER(Status)]
        L.psph_o_get.restype = i
        L.psph_o_get.argtypes = [p, p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.psph_o_grid.restype = i
        L.psph_o_grid.argtypes = [p, ctypes.POINTER(Grid)]
        L.psph_o_pairs.restype = i
        L.psph_o_pairs.argtypes = [p, p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.psph_o_scales.restype = i
        L.psph_o_scales.argtypes = [p, p, ctypes.c_size_t]
        L.psph_o_set_body_force.restype = None
        L.psph_o_set_body_force.argtypes = [p, p]
        L.psph_o_set_hooks.restype = None
        L.psph_o_set_hooks.argtypes = [p, HOOK, HOOK, p]
        L.psph_o_destroy.restype = None
        L.psph_o_destroy.argtypes = [p]
        L.psph_o_last_error.restype = ctypes.c_char_p
        L.psph_o_last_error.argtypes = []
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().psph_o_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """The reference's Solve loop on the CPU (FP64), same semantics as the GPU Solver."""

    def __init__(self, cfg, particles, dt, time=0.0, nthreads=1, track_scales=False):
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        h = ctypes.c_void_p()
        _check(lib().psph_o_create(ctypes.byref(cfg), _ptr(p), len(p), time, dt, nthreads,
                                   1 if track_scales else 0, ctypes.byref(h)))
        self._h = h
        self.cfg = cfg

    def run(self, max_steps, tf=float("inf"), tout=float("inf")):
        st = Status()
        _check(lib().psph_o_run(self._h, max_steps, tf, tout, ctypes.byref(st)))
        return st

    def particles(self):
        n = ctypes.c_size_t()
        lib().psph_o_get(self._h, None, 0, ctypes.byref(n))
        out = np.zeros(n.value, dtype=PARTICLE_DTYPE)
        _check(lib().psph_o_get(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def set_body_force(self, accel):
        a = np.ascontiguousarray(accel, dtype=np.float64)
        lib().psph_o_set_body_force(self._h, _ptr(a))

    def set_hooks(self, before=None, after=None):
        """GeneralBefore / GeneralAfter as Python functions of a structured array of the
        particles (reference order), which they may change in place."""
        def wrap(f):
            if f is None:
                return HOOK()
            def cb(ptr, n, user):
                buf = (ctypes.c_char*(n*PARTICLE_DTYPE.itemsize)).from_address(ptr)
                f(np.frombuffer(buf, dtype=PARTICLE_DTYPE))
            return HOOK(cb)
        self._hooks = (wrap(before), wrap(after))      # kept alive with the oracle
        lib().psph_o_set_hooks(self._h, self._hooks[0], self._hooks[1], None)

    def grid(self):
        g = Grid()
        _check(lib().psph_o_grid(self._h, ctypes.byref(g)))
        return g

    def pairs(self):
        n = ctypes.c_size_t()
        lib().psph_o_pairs(self._h, None, 0, ctypes.byref(n))
        out = np.zeros((n.value, 2), dtype=np.int64)
        _check(lib().psph_o_pairs(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def scales(self, n):
        out = np.zeros((n, 32))
        _check(lib().psph_o_scales(self._h, _ptr(out), n))
        return out

    def close(self):
        if self._h:
            lib().psph_o_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# scalar helpers used by the closed-form tests
def kernel(dim, kt, q, h):
    return lib().psph_o_kernel(dim, kt, q, h)


def grad_kernel(dim, kt, q, h):
    return lib().psph_o_grad_kernel(dim, kt, q, h)


def laplace_kernel(dim, kt, q, h):
    return lib().psph_o_laplace_kernel(dim, kt, q, h)


def secderiv_kernel(dim, kt, q, h):
    return lib().psph_o_secderiv_kernel(dim, kt, q, h)


def eos(eq, cs, p0, rho, rho0):
    return lib().psph_o_eos(eq, cs, p0, rho, rho0)


def sound_speed(eq, cs, rho, rho0):
    return lib().psph_o_sound_speed(eq, cs, rho, rho0)


def density_solid(eq, cs, p0, p, rho0):
    return lib().psph_o_density_solid(eq, cs, p0, p, rho0)


def viscous_force(viseq, mu, di, dj, gk, vab, dim, kt, rij, h, xij, vij):
    vi = np.zeros(3)
    vab = np.ascontiguousarray(vab, dtype=np.float64)
    xij = np.ascontiguousarray(xij, dtype=np.float64)
    vij = np.ascontiguousarray(vij, dtype=np.float64)
    lib().psph_o_viscous_force(viseq, _ptr(vi), mu, di, dj, gk, _ptr(vab), dim, kt, rij, h, _ptr(xij), _ptr(vij))
    return vi


def eig_sym3(m):
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
    vals = np.zeros(3)
    vecs = np.zeros(9)
    lib().psph_o_eig_sym3(_ptr(m), _ptr(vals), _ptr(vecs))
    return vals, vecs.reshape(3, 3)


def tir(dim, sigma, ti, density):
    s = np.ascontiguousarray(sigma, dtype=np.float64).reshape(9)
    out = np.zeros(9)
    lib().psph_o_tir(dim, _ptr(s), ti, density, _ptr(out))
    return out.reshape(3, 3)


def particle_move(p, dt, scheme, i22):
    """Advance one particle record (numpy scalar of PARTICLE_DTYPE, modified in place)."""
    arr = np.ascontiguousarray(np.atleast_1d(p), dtype=PARTICLE_DTYPE)
    _check(lib().psph_o_particle_move(_ptr(arr), dt, scheme, i22))
    return arr[0]
