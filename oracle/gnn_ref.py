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
