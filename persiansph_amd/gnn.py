"""ctypes binding of the learned surrogate's C ABI (include/psph_gnn.h).

`Network` wraps one message-passing network on the GPU: parameters, the loaded particle
graph, forward / advance, and training steps run by the library's hand-written HIP kernels.
There is no CPU fallback (construction raises without the library or a GPU).
"""
import ctypes

import numpy as np

from .solver import SolverError, _check, lib

GNN_EXPORTS = [
    "psph_gnn_create", "psph_gnn_destroy", "psph_gnn_num_params", "psph_gnn_get_params",
    "psph_gnn_set_params", "psph_gnn_load_state", "psph_gnn_load_inputs", "psph_gnn_graph",
    "psph_gnn_predict", "psph_gnn_advance", "psph_gnn_set_target_state", "psph_gnn_set_targets",
    "psph_gnn_train_step", "psph_gnn_get_grads",
]

HIDDEN = 128
NODE_IN = 32
EDGE_IN = 32


class GnnConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("layers", ctypes.c_int32), ("hidden", ctypes.c_int32),
        ("train", ctypes.c_int32), ("radius", ctypes.c_double), ("rho0", ctypes.c_double),
        ("vel_scale", ctypes.c_double), ("rho_scale", ctypes.c_double), ("out_dx", ctypes.c_double),
        ("out_dv", ctypes.c_double), ("out_drho", ctypes.c_double), ("drift_t", ctypes.c_double),
        ("deg_scale", ctypes.c_double), ("seed", ctypes.c_uint64),
    ]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        p, d, i64, i32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int32
        L.psph_gnn_create.argtypes = [ctypes.POINTER(GnnConfig), ctypes.POINTER(p)]
        L.psph_gnn_destroy.argtypes = [p]
        L.psph_gnn_num_params.argtypes = [p]
        L.psph_gnn_num_params.restype = i64
        L.psph_gnn_get_params.argtypes = [p, p, i64]
        L.psph_gnn_set_params.argtypes = [p, p, i64]
        L.psph_gnn_load_state.argtypes = [p, p, p, p, p, i64]
        L.psph_gnn_load_inputs.argtypes = [p, p, p, p, p, i64, i64]
        L.psph_gnn_graph.argtypes = [p, ctypes.POINTER(i64), p, p, i64]
        L.psph_gnn_predict.argtypes = [p, p]
        L.psph_gnn_advance.argtypes = [p, p, p, p]
        L.psph_gnn_set_target_state.argtypes = [p, p, p, p]
        L.psph_gnn_set_targets.argtypes = [p, p, p]
        L.psph_gnn_train_step.argtypes = [p, d, i32, ctypes.POINTER(d)]
        L.psph_gnn_get_grads.argtypes = [p, p, i64]
        for f in GNN_EXPORTS:
            if f != "psph_gnn_num_params":
                getattr(L, f).restype = ctypes.c_int
        _bound = True
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def layer_shapes(layers):
    """(k_in, 128) of every Linear layer in parameter order."""
    ks = [NODE_IN, HIDDEN, EDGE_IN, HIDDEN]
    for _ in range(layers):
        ks += [3*HIDDEN, HIDDEN, 2*HIDDEN, HIDDEN]
    ks += [HIDDEN, HIDDEN]
    return [(k, HIDDEN) for k in ks]


class Network:
    def __init__(self, dim, layers, radius, rho0=1000.0, vel_scale=1.0, rho_scale=1e-3,
                 out_dx=1e-3, out_dv=1e-2, out_drho=1.0, drift_t=0.0, deg_scale=1.0, train=True, seed=1):
        cfg = GnnConfig(dim=dim, layers=layers, hidden=HIDDEN, train=1 if train else 0, radius=radius,
                        rho0=rho0, vel_scale=vel_scale, rho_scale=rho_scale, out_dx=out_dx,
                        out_dv=out_dv, out_drho=out_drho, drift_t=drift_t, deg_scale=deg_scale, seed=seed)
        h = ctypes.c_void_p()
        _check(_lib().psph_gnn_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.cfg = cfg
        self.dim = dim
        self.layers = layers
        self.dout = 2*dim + 1
        self.n = 0

    def close(self):
        if self._h:
            _lib().psph_gnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return int(_lib().psph_gnn_num_params(self._h))

    def get_params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        _check(_lib().psph_gnn_get_params(self._h, _ptr(p), len(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib().psph_gnn_set_params(self._h, _ptr(p), len(p)))

    def grads(self):
        g = np.empty(self.num_params, dtype=np.float32)
        _check(_lib().psph_gnn_get_grads(self._h, _ptr(g), len(g)))
        return g

    def load_state(self, x, v, rho, isfree):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        fr = np.ascontiguousarray(isfree, dtype=np.int32)
        self.n = len(x)
        _check(_lib().psph_gnn_load_state(self._h, _ptr(x), _ptr(v), _ptr(rho), _ptr(fr), self.n))

    def load_inputs(self, node_in, edge_in, rowptr, cols):
        node_in = np.ascontiguousarray(node_in, dtype=np.float32)
        edge_in = np.ascontiguousarray(edge_in, dtype=np.float32)
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.n = len(node_in)
        _check(_lib().psph_gnn_load_inputs(self._h, _ptr(node_in), _ptr(edge_in), _ptr(rowptr), _ptr(cols),
                                           self.n, len(cols)))

    def graph(self):
        ne = ctypes.c_int64()
        _check(_lib().psph_gnn_graph(self._h, ctypes.byref(ne), None, None, 0))
        rp = np.empty(self.n + 1, dtype=np.int32)
        cols = np.empty(max(ne.value, 1), dtype=np.int32)
        _check(_lib().psph_gnn_graph(self._h, ctypes.byref(ne), _ptr(rp), _ptr(cols), len(cols)))
        return rp, cols[:ne.value]

    def predict(self):
        y = np.empty((self.n, self.dout), dtype=np.float32)
        _check(_lib().psph_gnn_predict(self._h, _ptr(y)))
        return y

    def advance(self):
        """Applies one network step (k solver steps) to the loaded state; returns x, v, rho."""
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        rho = np.empty(self.n)
        _check(_lib().psph_gnn_advance(self._h, _ptr(x), _ptr(v), _ptr(rho)))
        return x, v, rho

    def set_target_state(self, x1, v1, rho1):
        x1 = np.ascontiguousarray(x1, dtype=np.float64).reshape(-1, 3)
        v1 = np.ascontiguousarray(v1, dtype=np.float64).reshape(-1, 3)
        rho1 = np.ascontiguousarray(rho1, dtype=np.float64)
        _check(_lib().psph_gnn_set_target_state(self._h, _ptr(x1), _ptr(v1), _ptr(rho1)))

    def set_targets(self, tgt, wgt):
        tgt = np.ascontiguousarray(tgt, dtype=np.float32)
        wgt = np.ascontiguousarray(wgt, dtype=np.float32)
        _check(_lib().psph_gnn_set_targets(self._h, _ptr(tgt), _ptr(wgt)))

    def train_step(self, lr, update=True):
        loss = ctypes.c_double()
        _check(_lib().psph_gnn_train_step(self._h, lr, 1 if update else 0, ctypes.byref(loss)))
        return loss.value


__all__ = ["Network", "GnnConfig", "layer_shapes", "GNN_EXPORTS", "SolverError"]
