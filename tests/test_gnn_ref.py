"""The FP64 reference of the surrogate network (oracle/gnn_ref.py), pinned on the CPU: its
hand-written backward against central finite differences of its own loss, permutation
equivariance of the message passing, and the closed form of a first Adam step."""
import numpy as np

from oracle import gnn_ref as R


def _problem(seed, n=7, layers=2, dim=2):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, size=(n, 3))
    x[:, 2] = 0.0
    rp, cols = R.radius_graph(x, 0.6, dim)
    recv = np.repeat(np.arange(n), np.diff(rp))
    send = cols.astype(np.int64)
    X = np.zeros((n, R.NODE_IN))
    X[:, :4] = rng.normal(size=(n, 4))
    U = np.zeros((len(cols), R.EDGE_IN))
    U[:, :dim] = (x[recv, :dim] - x[send, :dim])/0.6
    U[:, dim] = np.linalg.norm(U[:, :dim], axis=1)
    p = rng.normal(scale=0.08, size=R.num_params(layers))
    dout = 2*dim + 1
    tgt = rng.normal(size=(n, dout))
    wgt = (rng.uniform(size=n) > 0.3).astype(float)
    wgt[0] = 1.0
    return p, layers, X, U, recv, send, tgt, wgt


def test_backward_matches_finite_differences():
    p, layers, X, U, recv, send, tgt, wgt = _problem(3)
    assert len(recv) > 0
    L, g = R.gradient(p, layers, X, U, recv, send, tgt, wgt)
    assert abs(L - R.loss(p, layers, X, U, recv, send, tgt, wgt)) < 1e-14*max(1.0, L)
    rng = np.random.default_rng(0)
    # every layer's weights and biases are represented among the probed entries
    offs, o = [], 0
    for k in R.layer_inputs(layers):
        offs.append((o, k*R.H))
        offs.append((o + k*R.H, R.H))
        o += k*R.H + R.H
    idx = []
    for start, size in offs:
        idx += list(start + rng.choice(size, size=3, replace=False))
    nz = np.flatnonzero(np.abs(g) > 1e-6)
    idx += list(rng.choice(nz, size=40, replace=False))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss(pp, layers, X, U, recv, send, tgt, wgt) - R.loss(pm, layers, X, U, recv, send, tgt, wgt))/(2*h)
        assert abs(fd - g[i]) <= 1e-6*max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def test_message_passing_is_permutation_equivariant():
    p, layers, X, U, recv, send, tgt, wgt = _problem(5)
    n = len(X)
    perm = np.random.default_rng(1).permutation(n)
    inv = np.argsort(perm)
    y = R.forward(p, layers, X, U, recv, send, tgt.shape[1])
    # relabel node k as inv[k]
    y2 = R.forward(p, layers, X[perm], U, inv[recv], inv[send], tgt.shape[1])
    np.testing.assert_allclose(y2, y[perm], rtol=1e-12, atol=1e-12)


def test_first_adam_step_is_signed_lr():
    rng = np.random.default_rng(2)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, m, v = R.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-3, eps=0.0)
    np.testing.assert_allclose(p1, p - 1e-3*np.sign(g), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1*g) and np.allclose(v, 0.001*g*g)
