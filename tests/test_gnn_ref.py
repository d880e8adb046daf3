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


def test_bf16_rounding_is_nearest_even():
    """gnn_ref.bf16 keeps 8 significant bits, rounding to nearest with ties to even."""
    one = 1.0
    ulp = 2.0**-7
    vals = np.array([one, one + ulp/2, one + 3*ulp/2, one + ulp/2 + 2**-20, -one - ulp/2, 3.0e38, 1e-40],
                    dtype=np.float32)
    got = R.bf16(vals).astype(np.float64)
    want = [one, one, one + 2*ulp, one + ulp, -one, float(np.float32(3.0e38)) , None]
    for g, w, v in zip(got, want, vals):
        if w is None:
            continue
        if abs(w) > 1e38:
            assert abs(g - w)/w < 2**-8
        else:
            assert g == w, (v, g, w)


def test_bf16_forward_tracks_fp64_forward():
    """The BF16-rounding forward of 10 layers stays within BF16 precision of the FP64 one, and
    with BF16-exact weights, inputs and a single zero-width layer stack it reduces to it."""
    rng = np.random.default_rng(0)
    n, dim, layers = 150, 2, 10
    x = rng.uniform(0, 1 + 0.3*n**0.5, size=(n, 3))
    x[:, 2] = 0
    rp, cols = R.radius_graph(x, 0.6, dim)
    recv = np.repeat(np.arange(n), np.diff(rp))
    X = np.zeros((n, R.NODE_IN), np.float32)
    X[:, :4] = rng.normal(size=(n, 4))
    U = np.zeros((len(cols), R.EDGE_IN), np.float32)
    U[:, :2] = (x[recv, :2] - x[cols, :2])/0.6
    p = (rng.normal(size=R.num_params(layers))*0.05).astype(np.float32)
    y0 = R.forward(p.astype(np.float64), layers, R.bf16(X).astype(np.float64), R.bf16(U).astype(np.float64),
                   recv, cols.astype(np.int64), 2*dim + 1)
    y1 = R.forward_bf16(p, layers, X, U, recv, cols.astype(np.int64), 2*dim + 1)
    err = np.linalg.norm(y1 - y0)/np.linalg.norm(y0)
    assert 1e-4 < err < 2e-2, err
    # all weights zero: every layer gives its bias exactly, the output is the decoder's bias
    z = np.zeros_like(p)
    P = R.unpack(z.astype(np.float64), layers)
    bd2 = np.zeros(R.H)
    bd2[:2*dim + 1] = [0.5, -0.25, 0.125, 1.0, -2.0]
    z[-R.H:] = bd2
    y2 = R.forward_bf16(z, layers, X, U, recv, cols.astype(np.int64), 2*dim + 1)
    np.testing.assert_array_equal(y2, np.broadcast_to(bd2[:2*dim + 1].astype(np.float32), y2.shape))


def test_bf16_forward_sensitivity_to_accumulation_order():
    """How far two BF16-rounding forwards that differ only in the FP32 accumulation order of
    their products drift apart: a few 1e-4 of the output norm at 2 layers, a few 1e-3 at 10,
    because each BF16 rounding that the order moves changes one activation by up to 2^-8 and
    such changes compound through the residual layers.  This bounds how closely the matrix
    cores' forward (FP32 accumulation in their own order) can follow gnn_ref.forward_bf16."""
    rng = np.random.default_rng(0)
    n, dim = 300, 2
    x = rng.uniform(0, 1 + 0.3*n**0.5, size=(n, 3))
    x[:, 2] = 0
    rp, cols = R.radius_graph(x, 0.6, dim)
    recv = np.repeat(np.arange(n), np.diff(rp))
    send = cols.astype(np.int64)
    X = np.zeros((n, R.NODE_IN), np.float32)
    X[:, :4] = rng.normal(size=(n, 4))
    U = np.zeros((len(cols), R.EDGE_IN), np.float32)
    U[:, :2] = (x[recv, :2] - x[send, :2])/0.6
    U[:, 2] = np.linalg.norm(U[:, :2], axis=1)
    drift = {}
    for layers in (2, 10):
        ks = R.layer_inputs(layers)
        p = np.concatenate([np.concatenate([rng.uniform(-1, 1, k*R.H)/np.sqrt(k), np.zeros(R.H)]) for k in ks])
        a = R.forward_bf16(p, layers, X, U, recv, send, 2*dim + 1)
        b = R.forward_bf16(p, layers, X, U, recv, send, 2*dim + 1, f32_products=True)
        drift[layers] = float(np.linalg.norm(a - b)/np.linalg.norm(a))
    assert 1e-5 < drift[2] < 1e-3, drift
    assert 1e-3 < drift[10] < 1e-2, drift


def test_bf16_reference_equals_fp64_where_bf16_is_exact():
    """On problems where every stored value fits BF16 (gnn_ref.bf16_exact_problem), the BF16
    forward and gradient equal the FP64 ones (pinned by finite differences above) bit for bit,
    2 and 10 layers, and do not depend on the FP32 accumulation order.  A slip in the BF16
    backward's structure (a dropped residual, reverse-edge or mean-share term, a swapped
    sender/receiver block, a wrong ReLU mask) breaks the equality."""
    for layers, seed, dens in ((2, 0, 0.01), (2, 2, 0.01), (2, 3, 0.01), (10, 3, 0.006)):
        p, X, U, rp, cols, recv, send, tgt, wgt = R.bf16_exact_problem(seed, layers, 4, density=dens)
        X64, U64 = X.astype(np.float64), U.astype(np.float64)
        y0 = R.forward(p.astype(np.float64), layers, X64, U64, recv, send, 4)
        np.testing.assert_array_equal(R.forward_bf16(p, layers, X, U, recv, send, 4), y0)
        L0, g0 = R.gradient(p.astype(np.float64), layers, X64, U64, recv, send, tgt.astype(np.float64),
                            wgt.astype(np.float64))
        L1, g1 = R.gradient_bf16(p, layers, X, U, recv, send, tgt, wgt)
        _, g2 = R.gradient_bf16(p, layers, X, U, recv, send, tgt, wgt, f32_products=True)
        assert np.count_nonzero(g0) > 4000
        assert L1 == L0
        np.testing.assert_array_equal(g1, g0)
        np.testing.assert_array_equal(g2, g0)


def test_bf16_gradient_tracks_fp64_gradient():
    """On a generic problem the BF16-rounding gradient stays within BF16 precision of the FP64
    one: every parameter block within 0.2 of its norm, cosine above 0.98."""
    p, layers, X, U, recv, send, tgt, wgt = _problem(9, n=40, layers=2)
    L0, g0 = R.gradient(p, layers, R.bf16(X).astype(np.float64), R.bf16(U).astype(np.float64), recv, send,
                        tgt, wgt)
    L1, g1 = R.gradient_bf16(p.astype(np.float32), layers, X.astype(np.float32), U.astype(np.float32), recv,
                             send, tgt.astype(np.float32), wgt.astype(np.float32))
    assert abs(L1 - L0) < 1e-2*L0
    o = 0
    for k in R.layer_inputs(layers):
        for size in (k*R.H, R.H):
            a, b = g1[o:o + size], g0[o:o + size]
            o += size
            nb = np.linalg.norm(b)
            if nb == 0.0:
                continue
            assert np.linalg.norm(a - b)/nb < 0.2 and float(a @ b)/(np.linalg.norm(a)*nb) > 0.98
