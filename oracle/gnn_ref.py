"""FP64 numpy restatement of the learned surrogate's network (test infrastructure only).

The surrogate has no counterpart in the reference; this module is the CPU check of the HIP
forward, backward and Adam kernels (persiansph_amd/csrc/gnn_kernels.hip) on the same
parameters and inputs.  Only tests/ may import it.  It is pinned by central finite
differences of its own loss (tests/test_gnn_ref.py), by the Adam closed form, and by
permutation symmetry of the message passing.

Parameter order and shapes follow include/psph_gnn.h: for every Linear layer its weight
[k_in][128] (row-major) then its bias [128]; layers: encoder node 1, 2, encoder edge 1, 2,
then per message-passing layer edge 1, 2, node 1, 2, then decoder 1, 2.
"""
import numpy as np

H = 128
NODE_IN = 32
EDGE_IN = 32


def layer_inputs(layers):
    ks = [NODE_IN, H, EDGE_IN, H]
    for _ in range(layers):
        ks += [3*H, H, 2*H, H]
    ks += [H, H]
    return ks


def unpack(p, layers):
    """List of (W, b) views into the flat parameter vector."""
    out, o = [], 0
    for k in layer_inputs(layers):
        W = p[o:o + k*H].reshape(k, H)
        o += k*H
        b = p[o:o + H]
        o += H
        out.append((W, b))
    assert o == len(p), (o, len(p))
    return out


def num_params(layers):
    return sum(k*H + H for k in layer_inputs(layers))


def degrees(recv, n):
    return np.maximum(np.bincount(recv, minlength=n), 1).astype(np.float64)


def segment_mean(E, recv, n):
    A = np.zeros((n, E.shape[1]))
    np.add.at(A, recv, E)
    return A/degrees(recv, n)[:, None]


def forward(p, layers, X, U, recv, send, dout, keep=False):
    """Network output y [n][dout]; with keep=True also the activations for backward."""
    P = unpack(np.asarray(p, dtype=np.float64), layers)
    n = len(X)
    relu = lambda z: np.maximum(z, 0.0)
    c = {}
    (W0, b0), (W1, b1), (W2, b2), (W3, b3) = P[:4]
    c["Zn0"] = relu(X @ W0 + b0)
    h = c["Zn0"] @ W1 + b1
    c["Ze0"] = relu(U @ W2 + b2)
    e = c["Ze0"] @ W3 + b3
    c["H"], c["E"], c["ZE"], c["AG"], c["ZN"], c["CE"], c["CN"] = [h], [e], [], [], [], [], []
    for l in range(layers):
        (We1, be1), (We2, be2), (Wn1, bn1), (Wn2, bn2) = P[4 + 4*l:8 + 4*l]
        ce = np.concatenate([e, h[recv], h[send]], axis=1)
        ze = relu(ce @ We1 + be1)
        e = e + ze @ We2 + be2
        a = segment_mean(e, recv, n)
        cn = np.concatenate([h, a], axis=1)
        zn = relu(cn @ Wn1 + bn1)
        h = h + zn @ Wn2 + bn2
        c["CE"].append(ce); c["ZE"].append(ze); c["AG"].append(a); c["CN"].append(cn); c["ZN"].append(zn)
        c["H"].append(h); c["E"].append(e)
    (Wd1, bd1), (Wd2, bd2) = P[4 + 4*layers:]
    c["Zd"] = relu(h @ Wd1 + bd1)
    y = (c["Zd"] @ Wd2 + bd2)[:, :dout]
    return (y, c) if keep else y


def bf16(a):
    """Round to BF16 (nearest, ties to even) and back to FP32, as the kernels store activations."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(p, layers, X, U, recv, send, dout, f32_products=False, keep=False):
    """The forward pass with the kernels' roundings: BF16 weights and stored activations,
    products accumulated in (at least) FP32, bias, ReLU and residual adds in FP32 in the
    kernels' order (Y = bf16(act(acc + b)), Y = bf16((R + acc) + b)), the segment mean as an
    FP32 running sum in CSR order times the FP32 reciprocal of the degree, the decoder output
    in FP32.  The products here are FP64 dot products rounded to FP32 (f32_products: FP32
    products in numpy's own accumulation order); the matrix cores' FP32 accumulation differs
    from either by a few FP32 ulps, which moves a BF16 rounding only where a value sits next to
    a rounding boundary -- but such moves compound through the layers.  keep=True also returns
    the stored (BF16) activations and weights for gradient_bf16."""
    f32 = np.float32
    P = unpack(np.asarray(p, dtype=np.float64).astype(f32), layers)
    Wb = [(bf16(W).astype(f32 if f32_products else np.float64), b.astype(f32)) for W, b in P]
    n = len(X)

    def mm(A, i):
        if f32_products:
            return (np.asarray(A, dtype=f32) @ Wb[i][0]).astype(f32)
        return (np.asarray(A, dtype=np.float64) @ Wb[i][0]).astype(f32)

    def act(A, i, relu):
        z = mm(A, i) + Wb[i][1]
        return bf16(np.maximum(z, f32(0)) if relu else z)

    def resid(R, A, i):
        return bf16((R + mm(A, i)) + Wb[i][1])

    deg = np.maximum(np.bincount(recv, minlength=n), 1).astype(f32)
    inv = (f32(1)/deg).astype(f32)
    c = {"X": bf16(X), "U": bf16(U), "W": Wb, "H": [], "E": [], "ZE": [], "ZN": [], "AG": []}
    c["Zn0"] = act(c["X"], 0, True)
    h = act(c["Zn0"], 1, False)
    c["Ze0"] = act(c["U"], 2, True)
    e = act(c["Ze0"], 3, False)
    for l in range(layers):
        i0 = 4 + 4*l
        c["H"].append(h); c["E"].append(e)
        ze = act(np.concatenate([e, h[recv], h[send]], axis=1), i0, True)
        e = resid(e, ze, i0 + 1)
        s = np.zeros((n, H), dtype=f32)
        np.add.at(s, recv, e)                   # FP32 running sums in CSR order
        a = bf16(s*inv[:, None])
        zn = act(np.concatenate([h, a], axis=1), i0 + 2, True)
        h = resid(h, zn, i0 + 3)
        c["ZE"].append(ze); c["AG"].append(a); c["ZN"].append(zn)
    c["H"].append(h); c["E"].append(e)
    iD = 4 + 4*layers
    c["Zd"] = act(h, iD, True)
    y = (mm(c["Zd"], iD + 1) + Wb[iD + 1][1])[:, :dout]
    return (y, c) if keep else y


def gradient_bf16(p, layers, X, U, recv, send, tgt, wgt, f32_products=False):
    """(loss, d loss / d p) with the training kernels' roundings: the BF16 forward above; the
    loss gradient dy = 2 w (y - t) / (dout sum w) in FP32 stored as BF16; through each ReLU layer
    dZ = bf16(dY W^T) where the stored activation is positive; gradients into a layer's inputs
    dIn = dZ W^T in FP32; the FP32 running gradients of node and edge latents (decoder input,
    residual paths, the mean's share of each edge, the sums over a node's edges) mirrored as
    BF16 where a product reads them; weight gradients X^T dY and bias gradients (column sums)
    of the BF16 operands accumulated in FP32 (f32_products: every product in FP32 in numpy's
    order, as in forward_bf16)."""
    f32 = np.float32
    dout = tgt.shape[1]
    n = len(X)
    y, c = forward_bf16(p, layers, X, U, recv, send, dout, f32_products=f32_products, keep=True)
    ft = f32 if f32_products else np.float64
    Wb = c["W"]
    g = np.zeros(len(p), dtype=np.float64)
    G = unpack(g, layers)
    wsum = float(np.sum(wgt))
    inv_count = f32(1.0/(dout*wsum))
    w = np.asarray(wgt, dtype=f32)[:, None]
    d = (y - np.asarray(tgt, dtype=f32)).astype(f32)
    L = float(np.sum((w*d*d*inv_count).astype(f32), dtype=np.float64))
    dy = np.zeros((n, H), dtype=f32)
    dy[:, :dout] = bf16(f32(2)*w*d*inv_count)

    def wgrad(i, Xin, dY):
        G[i][0][...] += (np.asarray(Xin, ft).T @ np.asarray(dY, ft)).astype(f32)
        G[i][1][...] += np.asarray(dY, ft).sum(axis=0).astype(f32)

    def back_mask(dY, i, Z):
        return np.where(Z > 0, bf16((np.asarray(dY, ft) @ Wb[i][0].T).astype(f32)), f32(0))

    def back_in(dZ, i):
        return (np.asarray(dZ, ft) @ Wb[i][0].T).astype(f32)

    iD = 4 + 4*layers
    deg = np.maximum(np.bincount(recv, minlength=n), 1).astype(f32)
    wgrad(iD + 1, c["Zd"], dy)
    dZ = back_mask(dy, iD + 1, c["Zd"])
    wgrad(iD, c["H"][layers], dZ)
    dh = back_in(dZ, iD)
    dhb = bf16(dh)
    de = None
    deb = None
    # the reverse edge of (i <- j) is (j <- i): gradients reach senders through it
    key = {(int(r), int(s_)): k for k, (r, s_) in enumerate(zip(recv, send))}
    rev = np.array([key.get((int(s_), int(r)), -1) for r, s_ in zip(recv, send)], dtype=np.int64)
    for l in range(layers - 1, -1, -1):
        ie1, ie2, in1, in2 = 4 + 4*l, 5 + 4*l, 6 + 4*l, 7 + 4*l
        wgrad(in2, c["ZN"][l], dhb)
        dZ = back_mask(dhb, in2, c["ZN"][l])
        wgrad(in1, np.concatenate([c["H"][l], c["AG"][l]], axis=1), dZ)
        dInN = back_in(dZ, in1)
        det = ((de if de is not None else f32(0)) + dInN[recv, H:]/deg[recv][:, None]).astype(f32)
        detb = bf16(det)
        wgrad(ie2, c["ZE"][l], detb)
        dZ = back_mask(detb, ie2, c["ZE"][l])
        wgrad(ie1, np.concatenate([c["E"][l], c["H"][l][recv], c["H"][l][send]], axis=1), dZ)
        dInE = back_in(dZ, ie1)
        de = (dInE[:, :H] + det).astype(f32)
        deb = bf16(de)
        dh_new = (dh + dInN[:, :H]).astype(f32)
        acc = np.zeros((n, H), dtype=np.float64)
        np.add.at(acc, recv, dInE[:, H:2*H].astype(np.float64))
        ok = rev >= 0
        np.add.at(acc, recv[ok], dInE[rev[ok], 2*H:].astype(np.float64))
        dh = (dh_new + acc).astype(f32)
        dhb = bf16(dh)
    wgrad(1, c["Zn0"], dhb)
    dZ = back_mask(dhb, 1, c["Zn0"])
    wgrad(0, c["X"], dZ)
    if deb is None:
        deb = np.zeros((len(recv), H), dtype=f32)
    wgrad(3, c["Ze0"], deb)
    dZ = back_mask(deb, 3, c["Ze0"])
    wgrad(2, c["U"], dZ)
    return L, g


def bf16_exact_problem(seed, layers, dout, n=16, density=0.01):
    """A problem on which BF16 storage loses nothing: a ring graph (every node receives from
    its two neighbours, so every mean is a halving), inputs and targets in halves, biases in
    {0, +-1/2, 1}, sparse weights of +-1, half of the loss weights zero.  Where every stored
    value then fits BF16's 8 significant bits, gradient_bf16 and forward_bf16 equal the FP64
    gradient and forward bit for bit (for the seeds the tests use), and so does any FP32
    accumulation order.  Returns (p, X, U, rowptr, cols, recv, send, tgt, wgt)."""
    rng = np.random.default_rng(seed)
    rows = [sorted({(i - 1) % n, (i + 1) % n}) for i in range(n)]
    cols = np.array([j for r in rows for j in r], dtype=np.int32)
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    recv = np.repeat(np.arange(n), np.diff(rowptr))
    send = cols.astype(np.int64)
    X = np.zeros((n, NODE_IN), np.float32)
    X[:, :4] = rng.integers(-2, 3, size=(n, 4))/2
    U = np.zeros((len(cols), EDGE_IN), np.float32)
    U[:, :3] = rng.integers(-2, 3, size=(len(cols), 3))/2
    parts = []
    for k in layer_inputs(layers):
        W = np.zeros(k*H)
        m = rng.uniform(size=k*H) < density
        W[m] = rng.choice([-1.0, 1.0], size=int(m.sum()))
        b = np.zeros(H)
        mb = rng.uniform(size=H) < 0.3
        b[mb] = rng.choice([-0.5, 0.5, 1.0], size=int(mb.sum()))
        parts += [W, b]
    p = np.concatenate(parts).astype(np.float32)
    tgt = (rng.integers(-4, 5, size=(n, dout))/2).astype(np.float32)
    wgt = np.ones(n, np.float32)
    wgt[rng.permutation(n)[:n//2]] = 0.0
    return p, X, U, rowptr, cols, recv, send, tgt, wgt


def loss(p, layers, X, U, recv, send, tgt, wgt):
    dout = tgt.shape[1]
    y = forward(p, layers, X, U, recv, send, dout)
    return float(np.sum(wgt[:, None]*(y - tgt)**2)/(dout*np.sum(wgt)))


def gradient(p, layers, X, U, recv, send, tgt, wgt):
    """(loss, d loss / d p) by reverse-mode differentiation written out layer by layer."""
    p = np.asarray(p, dtype=np.float64)
    dout = tgt.shape[1]
    n = len(X)
    y, c = forward(p, layers, X, U, recv, send, dout, keep=True)
    P = unpack(p, layers)
    g = np.zeros_like(p)
    G = unpack(g, layers)
    scale = 1.0/(dout*np.sum(wgt))
    L = float(np.sum(wgt[:, None]*(y - tgt)**2)*scale)
    dy = np.zeros((n, H))
    dy[:, :dout] = 2.0*wgt[:, None]*(y - tgt)*scale

    def lin_back(i, Xin, dY, Z=None):
        """Gradient of Y = Xin W + b for layer i; returns d Xin."""
        W, _ = P[i]
        G[i][0][...] += Xin.T @ dY
        G[i][1][...] += dY.sum(axis=0)
        return dY @ W.T

    iD = 4 + 4*layers
    dZd = lin_back(iD + 1, c["Zd"], dy)*(c["Zd"] > 0)
    dh = lin_back(iD, c["H"][layers], dZd)
    de = np.zeros_like(c["E"][0])
    for l in range(layers - 1, -1, -1):
        ie1, ie2, in1, in2 = 4 + 4*l, 5 + 4*l, 6 + 4*l, 7 + 4*l
        dzn = lin_back(in2, c["ZN"][l], dh)*(c["ZN"][l] > 0)
        dcn = lin_back(in1, c["CN"][l], dzn)
        dh_next = dh + dcn[:, :H]
        da = dcn[:, H:]/degrees(recv, n)[:, None]
        det = de + da[recv]
        dze = lin_back(ie2, c["ZE"][l], det)*(c["ZE"][l] > 0)
        dce = lin_back(ie1, c["CE"][l], dze)
        de = det + dce[:, :H]
        dh = dh_next.copy()
        np.add.at(dh, recv, dce[:, H:2*H])
        np.add.at(dh, send, dce[:, 2*H:])
    dZn0 = lin_back(1, c["Zn0"], dh)*(c["Zn0"] > 0)
    lin_back(0, X, dZn0)
    dZe0 = lin_back(3, c["Ze0"], de)*(c["Ze0"] > 0)
    lin_back(2, U, dZe0)
    return L, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t >= 1); returns the new p, m, v."""
    m = b1*m + (1 - b1)*g
    v = b2*v + (1 - b2)*g*g
    mh = m/(1 - b1**t)
    vh = v/(1 - b2**t)
    return p - lr*mh/(np.sqrt(vh) + eps), m, v


def radius_graph(x, radius, dim):
    """CSR by receiver (ascending senders) of pairs closer than radius, brute force."""
    n = len(x)
    d = x[:, None, :dim] - x[None, :, :dim]
    r2 = np.sum(d*d, axis=2)
    adj = (r2 <= radius*radius) & ~np.eye(n, dtype=bool)
    rp = np.zeros(n + 1, dtype=np.int32)
    rp[1:] = np.cumsum(adj.sum(axis=1))
    cols = np.nonzero(adj)[1].astype(np.int32)
    return rp, cols
