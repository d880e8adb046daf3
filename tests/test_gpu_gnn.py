"""GPU checks of the learned surrogate's HIP kernels against the FP64 numpy reference of the
same network (oracle/gnn_ref.py): the radius graph, the forward pass (BF16 matrix-core
GEMMs with FP32 accumulation), the backward pass, the Adam update and a short fit.

Tolerances: against FP64 on identical parameters and inputs, where activations are rounded to
BF16 (8 significant bits) at every layer, the forward output is held to 3% of its norm and each
layer's gradient to 0.15 of its norm; against the BF16-rounding reference (gnn_ref.forward_bf16,
gradient_bf16) to 1e-3 (forward) and 1e-2 (gradients) up to 3 layers and to the reference's own
accumulation-order drift at 10; on problems where BF16 storage is exact, bit for bit."""
import numpy as np
import pytest

from oracle import gnn_ref as R

pytestmark = pytest.mark.gpu


def _net(dim, layers, radius=0.6, train=True, seed=7):
    from persiansph_amd.gnn import Network
    return Network(dim, layers, radius, train=train, seed=seed)


def _inputs(n, dim, radius, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0 + 0.3*n**(1.0/dim), size=(n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    rp, cols = R.radius_graph(x, radius, dim)
    recv = np.repeat(np.arange(n), np.diff(rp))
    X = np.zeros((n, R.NODE_IN), dtype=np.float32)
    X[:, :dim + 2] = rng.normal(size=(n, dim + 2))
    U = np.zeros((len(cols), R.EDGE_IN), dtype=np.float32)
    U[:, :dim] = (x[recv, :dim] - x[cols, :dim])/radius
    U[:, dim] = np.linalg.norm(U[:, :dim], axis=1)
    return x, rp, cols, recv, X, U


def _bf16(a):
    """Round FP32 values to BF16 (nearest even) and back, as the library stores its inputs."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("dim", [2, 3])
def test_radius_graph_matches_brute_force(dim):
    rng = np.random.default_rng(11)
    n = 700
    x = rng.uniform(0.0, 3.0, size=(n, 3))
    if dim == 2:
        x[:, 2] = 0.0
    radius = 0.35 if dim == 3 else 0.2
    print(f"[gnn graph] dim={dim} n={n}", flush=True)
    g = _net(dim, 1, radius=radius)
    g.load_state(x, np.zeros((n, 3)), np.full(n, 1000.0), np.ones(n, np.int32))
    rp, cols = g.graph()
    rp0, cols0 = R.radius_graph(x, radius, dim)
    np.testing.assert_array_equal(rp, rp0)
    np.testing.assert_array_equal(cols, cols0)
    g.close()


@pytest.mark.parametrize("dim,layers", [(2, 2), (3, 3)])
def test_forward_matches_reference(dim, layers):
    n = 300
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 3)
    print(f"[gnn fwd] dim={dim} layers={layers} n={n} edges={len(cols)}", flush=True)
    g = _net(dim, layers)
    g.load_inputs(X, U, rp, cols)
    y = g.predict()
    p = g.get_params().astype(np.float64)
    y0 = R.forward(p, layers, _bf16(X), _bf16(U), recv, cols.astype(np.int64), 2*dim + 1)
    err = np.linalg.norm(y - y0)/np.linalg.norm(y0)
    assert err < 3e-2, err
    g.close()


@pytest.mark.parametrize("dim,layers", [(2, 2), (3, 3), (2, 10), (3, 10)])
def test_forward_matches_bf16_reference(dim, layers):
    """Against a reference that rounds to BF16 where the kernels do (gnn_ref.forward_bf16), the
    forward pass agrees to the FP32 accumulation order of the matrix cores: 1e-3 of the output
    norm up to 3 message-passing layers, 1e-2 at 10 (the FP64 reference above is held to 3e-2).
    Two CPU forwards that differ only in that order drift apart by a few 1e-3 at 10 layers
    (test_gnn_ref.test_bf16_forward_sensitivity_to_accumulation_order), so no tighter bar holds
    at that depth."""
    n = 300
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 3)
    print(f"[gnn fwd bf16] dim={dim} layers={layers} n={n} edges={len(cols)}", flush=True)
    g = _net(dim, layers, train=False)
    g.load_inputs(X, U, rp, cols)
    y = g.predict()
    p = g.get_params()
    y0 = R.forward_bf16(p, layers, X, U, recv, cols.astype(np.int64), 2*dim + 1)
    err = float(np.linalg.norm(y - y0)/np.linalg.norm(y0))
    print(f"[gnn fwd bf16] relative error {err:.3e}", flush=True)
    assert err < (1e-3 if layers <= 3 else 1e-2), err
    g.close()


@pytest.mark.parametrize("dim,layers", [(2, 2), (3, 3)])
def test_fused_inference_equals_training_forward(dim, layers):
    """Inference runs each edge update as one fused kernel (the first layer's activations stay
    in LDS); training runs the two GEMMs apart and keeps the activations.  Same operations in
    the same order: the predictions agree to the bit, and both meet the reference's bar."""
    n = 300
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 5)
    print(f"[gnn fused] dim={dim} layers={layers} n={n} edges={len(cols)}", flush=True)
    gt = _net(dim, layers, train=True)
    gi = _net(dim, layers, train=False)
    gi.set_params(gt.get_params())
    for g in (gt, gi):
        g.load_inputs(X, U, rp, cols)
    yt, yi = gt.predict(), gi.predict()
    np.testing.assert_array_equal(yi, yt)
    p = gt.get_params().astype(np.float64)
    y0 = R.forward(p, layers, _bf16(X), _bf16(U), recv, cols.astype(np.int64), 2*dim + 1)
    assert np.linalg.norm(yi - y0)/np.linalg.norm(y0) < 3e-2
    gt.close()
    gi.close()


@pytest.mark.parametrize("dim,layers", [(2, 2), (3, 2), (3, 10)])
def test_gradients_match_reference(dim, layers):
    """Gradients against the BF16-rounding reference (gnn_ref.gradient_bf16, which rounds where
    the training kernels do).  Up to 3 layers: every parameter block within 1e-2 of its norm
    (two FP32 accumulation orders of that reference differ by ~5e-4).  At 10 layers the BF16
    gradient is itself sensitive to the accumulation order (rounding flips compound through
    the layers: the reference's two orders differ by several per cent per block), so each block
    is held to three times that measured drift of the reference on the same inputs (at least
    1e-2), and the whole vector likewise.  The FP64 reference (gnn_ref.gradient) is held to the
    BF16 precision of the kernels: 0.15 per block, cosine above 0.99."""
    n = 250
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 4)
    rng = np.random.default_rng(5)
    dout = 2*dim + 1
    tgt = rng.normal(size=(n, dout)).astype(np.float32)
    wgt = (rng.uniform(size=n) > 0.2).astype(np.float32)
    print(f"[gnn grad] dim={dim} layers={layers} n={n} edges={len(cols)}", flush=True)
    g = _net(dim, layers)
    g.load_inputs(X, U, rp, cols)
    g.set_targets(tgt, wgt)
    loss = g.train_step(0.0, update=False)
    gg = g.grads().astype(np.float64)
    p = g.get_params()
    send = cols.astype(np.int64)
    Lb, gb = R.gradient_bf16(p, layers, X, U, recv, send, tgt, wgt)
    _, gb2 = R.gradient_bf16(p, layers, X, U, recv, send, tgt, wgt, f32_products=True)
    L0, g0 = R.gradient(p.astype(np.float64), layers, _bf16(X), _bf16(U), recv, send, tgt.astype(np.float64),
                        wgt.astype(np.float64))
    assert abs(loss - Lb) <= 1e-3*Lb, (loss, Lb)
    assert abs(loss - L0) <= 3e-2*L0, (loss, L0)
    deep = layers > 3
    worst = 0.0
    o = 0
    for li, k in enumerate(R.layer_inputs(layers)):
        for size in (k*R.H, R.H):
            a, b, b2, c = gg[o:o + size], gb[o:o + size], gb2[o:o + size], g0[o:o + size]
            o += size
            nb = np.linalg.norm(b)
            if nb == 0.0:
                assert np.linalg.norm(a) < 1e-6
                continue
            rel = np.linalg.norm(a - b)/nb
            bar = max(3.0*np.linalg.norm(b2 - b)/nb, 1e-2) if deep else 1e-2
            worst = max(worst, rel/bar)
            assert rel < bar, (li, size, rel, bar)
            cos = float(a @ c)/(np.linalg.norm(a)*np.linalg.norm(c))
            assert cos > 0.99 and np.linalg.norm(a - c)/np.linalg.norm(c) < 0.15, (li, size, cos)
    glob = np.linalg.norm(gg - gb)/np.linalg.norm(gb)
    gbar = max(3.0*np.linalg.norm(gb2 - gb)/np.linalg.norm(gb), 1e-2) if deep else 1e-2
    print(f"[gnn grad] vs bf16 reference: whole vector {glob:.3e} (bar {gbar:.3e}), worst block at "
          f"{worst:.2f} of its bar", flush=True)
    assert glob < gbar
    g.close()


@pytest.mark.parametrize("layers,seed,dens", [(2, 0, 0.01), (2, 4, 0.01), (10, 3, 0.006), (10, 1, 0.006)])
def test_bf16_exact_problem_matches_reference_to_the_bit(layers, seed, dens):
    """On problems where every stored value fits BF16 (gnn_ref.bf16_exact_problem: sparse +-1
    weights, inputs in halves, a ring graph), every FP32 accumulation is exact in any order, so
    the matrix-core forward and training gradients equal the BF16-rounding reference
    (gnn_ref.forward_bf16, gradient_bf16) bit for bit, 10 message-passing layers included."""
    dim = 2
    dout = 2*dim + 1
    p, X, U, rp, cols, recv, send, tgt, wgt = R.bf16_exact_problem(seed, layers, dout, density=dens)
    print(f"[gnn exact] layers={layers} seed={seed} n={len(X)} edges={len(cols)}", flush=True)
    g = _net(dim, layers)
    g.set_params(p)
    g.load_inputs(X, U, rp, cols)
    y = g.predict()
    np.testing.assert_array_equal(y, R.forward_bf16(p, layers, X, U, recv, send, dout))
    g.set_targets(tgt, wgt)
    loss = g.train_step(0.0, update=False)
    gg = g.grads()
    Lb, gb = R.gradient_bf16(p, layers, X, U, recv, send, tgt, wgt)
    assert np.count_nonzero(gb) > 4000
    assert abs(loss - Lb) <= 1e-6*Lb, (loss, Lb)
    np.testing.assert_array_equal(gg, gb.astype(np.float32))
    g.close()


def test_adam_update_matches_closed_form():
    dim, layers, n = 2, 1, 120
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 6)
    rng = np.random.default_rng(8)
    tgt = rng.normal(size=(n, 2*dim + 1)).astype(np.float32)
    g = _net(dim, layers)
    g.load_inputs(X, U, rp, cols)
    g.set_targets(tgt, np.ones(n, np.float32))
    p0 = g.get_params().astype(np.float64)
    lr = 1e-3
    g.train_step(lr, update=True)
    grad = g.grads().astype(np.float64)
    p1 = g.get_params().astype(np.float64)
    pe, _, _ = R.adam(p0, grad, np.zeros_like(p0), np.zeros_like(p0), 1, lr)
    np.testing.assert_allclose(p1, pe, rtol=0, atol=2e-6*lr + 1e-7)
    g.close()


def test_training_fits_one_sample():
    dim, layers, n = 2, 2, 200
    x, rp, cols, recv, X, U = _inputs(n, dim, 0.6, 9)
    rng = np.random.default_rng(10)
    # a smooth target: a fixed linear map of the node inputs plus a neighbour average
    T = X[:, :dim + 2] @ rng.normal(size=(dim + 2, 2*dim + 1))
    tgt = (T/np.abs(T).max()).astype(np.float32)
    g = _net(dim, layers)
    g.load_inputs(X, U, rp, cols)
    g.set_targets(tgt, np.ones(n, np.float32))
    first = g.train_step(1e-3)
    for _ in range(300):
        last = g.train_step(1e-3)
    print(f"[gnn fit] loss {first:.4e} -> {last:.4e}", flush=True)
    assert last < 0.1*first, (first, last)
    g.close()
