"""ctypes binding of the C ABI (include/persiansph_amd.h) of the GPU path.

The shared library is built in-tree (persiansph_amd/lib/libpersiansph_amd.so).  There is no
CPU fallback: if the library is missing or no GPU is visible, construction raises.
"""
import ctypes
import os

import numpy as np

from .abi import ERRORS, OUTPUT_FIELDS, PARTICLE_DTYPE, Config, Grid, OutputF32, Status, grid_dict

_HERE = os.path.dirname(os.path.abspath(__file__))
# PSPH_LIB names another build of the same library (same C ABI), e.g. to compare builds
LIB_PATH = os.environ.get("PSPH_LIB") or os.path.join(_HERE, "lib", "libpersiansph_amd.so")

EXPORTS = [
    "psph_create", "psph_upload", "psph_run", "psph_download", "psph_neighbour_pairs",
    "psph_neighbour_pairs_noalloc", "psph_get_grid", "psph_set_timing", "psph_kernel_stats", "psph_stream",
    "psph_traffic_model", "psph_destroy", "psph_last_error", "psph_version",
    "psph_output_begin", "psph_output_end", "psph_set_body_force", "psph_set_hooks", "psph_hook_accel",
]

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"persiansph_amd HIP library not built: {LIB_PATH} missing "
                               "(run __graft_entry__.build() or make -C persiansph_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        p, d, i64, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_size_t
        L.psph_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(p)]
        L.psph_upload.argtypes = [p, p, sz, d, d]
        L.psph_run.argtypes = [p, i64, d, d, ctypes.POINTER(Status)]
        L.psph_download.argtypes = [p, p, sz, ctypes.POINTER(sz)]
        L.psph_neighbour_pairs.argtypes = [p, p, sz, ctypes.POINTER(sz)]
        L.psph_neighbour_pairs_noalloc.argtypes = [p, p, sz, ctypes.POINTER(sz)]
        L.psph_get_grid.argtypes = [p, ctypes.POINTER(Grid)]
        L.psph_set_timing.argtypes = [p, ctypes.c_int]
        L.psph_kernel_stats.argtypes = [p, ctypes.c_char_p, sz, p, p, sz, ctypes.POINTER(sz)]
        L.psph_stream.argtypes = [p, ctypes.POINTER(p)]
        L.psph_traffic_model.argtypes = [p, ctypes.POINTER(d), ctypes.POINTER(d)]
        L.psph_destroy.argtypes = [p]
        L.psph_output_begin.argtypes = [p]
        L.psph_set_body_force.argtypes = [p, p]
        L.psph_set_hooks.argtypes = [p, ctypes.c_int, ctypes.c_int]
        L.psph_hook_accel.argtypes = [p, p, p, sz]
        L.psph_output_end.argtypes = [p, ctypes.POINTER(OutputF32)]
        L.psph_last_error.argtypes = [ctypes.c_char_p, sz]
        for f in EXPORTS:
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


class SolverError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().psph_last_error(buf, 1024)
        raise SolverError(rc, buf.value.decode())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Solver:
    """GPU Solve loop behind the C ABI; same interface as the CPU oracle's wrapper."""

    def __init__(self, cfg, particles, dt, time=0.0):
        h = ctypes.c_void_p()
        _check(lib().psph_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.cfg = cfg
        p = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        try:
            _check(lib().psph_upload(self._h, _ptr(p), len(p), time, dt))
        except Exception:
            self.close()
            raise

    def run(self, max_steps, tf=float("inf"), tout=float("inf")):
        st = Status()
        _check(lib().psph_run(self._h, max_steps, tf, tout, ctypes.byref(st)))
        return st

    def particles(self):
        n = ctypes.c_size_t()
        _check(lib().psph_download(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=PARTICLE_DTYPE)
        _check(lib().psph_download(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out

    def set_body_force(self, accel):
        """Every free particle's acceleration starts each force pass at Gravity + accel (the
        body-force GeneralBefore hook, psph_set_body_force)."""
        a = np.ascontiguousarray(accel, dtype=np.float64)
        _check(lib().psph_set_body_force(self._h, _ptr(a)))

    def set_hooks(self, before, after):
        """Stop every step at the GeneralBefore / GeneralAfter hook points (psph_set_hooks)."""
        _check(lib().psph_set_hooks(self._h, 1 if before else 0, 1 if after else 0))

    def hook_accel(self, a, ddensity=None):
        """At a stop inside a step: the accelerations (and density rates) of every particle in
        download order, as a hook left them (psph_hook_accel)."""
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
        dd = None if ddensity is None else np.ascontiguousarray(ddensity, dtype=np.float64)
        _check(lib().psph_hook_accel(self._h, _ptr(a), None if dd is None else _ptr(dd), len(a)))

    def output_begin(self):
        """Starts the packed FP32 output of the current state (psph_output_begin)."""
        _check(lib().psph_output_begin(self._h))

    def output_end(self, copy=True):
        """The packed output started by output_begin: WriteXDMF's datasets as numpy copies,
        (n, width) float32 (tag: int32) in reference particle order (copy=False: only waits
        for the packed state and returns the particle count)."""
        o = OutputF32()
        _check(lib().psph_output_end(self._h, ctypes.byref(o)))
        if not copy:
            return o.n
        out = {}
        for name, dt, w in OUTPUT_FIELDS:
            buf = (ctypes.c_char*(4*w*o.n)).from_address(getattr(o, name))
            a = np.frombuffer(buf, dtype=dt).copy()
            out[name] = a.reshape(o.n, w) if w > 1 else a
        return out

    def output_f32(self):
        self.output_begin()
        return self.output_end()

    def pairs(self, noalloc=False):
        """The oriented pair set of the current positions (psph_neighbour_pairs, or with
        noalloc=True psph_neighbour_pairs_noalloc, which allocates no device memory)."""
        f = lib().psph_neighbour_pairs_noalloc if noalloc else lib().psph_neighbour_pairs
        n = ctypes.c_size_t()
        _check(f(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 2), dtype=np.int64)
        _check(f(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def grid(self):
        g = Grid()
        _check(lib().psph_get_grid(self._h, ctypes.byref(g)))
        return g

    def set_timing(self, on, force_only=False):
        """Per-kernel HIP events on the solver's stream (force_only: around the force pass only)."""
        _check(lib().psph_set_timing(self._h, (2 if force_only else 1) if on else 0))

    def kernel_stats(self):
        names = ctypes.create_string_buffer(512)
        cap = 32
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        nk = ctypes.c_size_t()
        _check(lib().psph_kernel_stats(self._h, names, 512, _ptr(ms), _ptr(cnt), cap, ctypes.byref(nk)))
        keys = names.value.decode().split(",")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}

    def traffic_model(self):
        t = ctypes.c_double()
        f = ctypes.c_double()
        _check(lib().psph_traffic_model(self._h, ctypes.byref(t), ctypes.byref(f)))
        return t.value, f.value

    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().psph_stream(self._h, ctypes.byref(s)))
        return s.value

    def grid_dict(self):
        return grid_dict(self.grid())

    def close(self):
        if getattr(self, "_h", None):
            lib().psph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
