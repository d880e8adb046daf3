"""GPU parity on the benchmarked workload itself: the 3-D dam break of the benchmark at full size
(tank 3.2 x 1.0 x 1.0 m, column 1.2 x 0.6 x 1.0 m, dx = 0.01 m, 3 wall layers, fluid jitter
+-0.01 dx, seed 12345: 1,079,268 particles of which 720,000 fluid).

  * the cell grid equals the oracle's (CellInitiate, Domain.cpp:549-660);
  * the oriented pair set of the first step equals the oracle's (28 M pairs), compared as
    sorted arrays of P1 * N + P2;
  * one force pass is within 1e-12 of the contribution scale for a and dDensity of every fluid
    particle and for the wall pressure and NSv of every wall, as in the small cases.

The oracle runs with every host core the box gives this process (up to 16 OpenMP threads): its
pair order then differs from the single-thread reference order, which the set comparison and
the contribution-scale bar do not depend on."""
import os

import numpy as np
import pytest

from persiansph_amd import Solver, cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-12


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _check(name, g, o, scale, rel=REL):
    err = np.abs(g - o)
    tol = rel*(np.abs(o) + scale)
    bad = err > tol
    assert not bad.any(), (name, int(bad.sum()), float((err/np.maximum(tol, 1e-300)).max()))


def _codes(pairs, n):
    c = pairs[:, 0].astype(np.int64)*n + pairs[:, 1].astype(np.int64)
    c.sort()
    return c


def test_full_size_dam_break_pairs_and_force_pass():
    cfg, p, dt = cases.dam_break_3d()
    n = len(p)
    assert n == 1079268 and int((p["IsFree"] == 1).sum()) == 720000
    nt = _threads()
    print(f"[fullsize] n={n} oracle threads={nt}", flush=True)
    g = Solver(cfg, p, dt)
    o = Oracle(cfg, p, dt, nthreads=nt, track_scales=True)
    gg, og = g.grid(), o.grid()
    assert list(gg.cell_no) == list(og.cell_no) == [125, 39, 40]
    assert list(gg.cell_size) == list(og.cell_size)
    assert list(gg.blpf) == list(og.blpf)
    gp = _codes(g.pairs(), n)
    op = _codes(o.pairs(), n)
    print(f"[fullsize] pairs gpu={len(gp)} oracle={len(op)}", flush=True)
    assert len(gp) == len(op) > 20_000_000
    assert np.array_equal(gp, op)
    del gp, op
    sg, so = g.run(1, tout=0.0), o.run(1, tout=0.0)
    assert sg.stop_reason == so.stop_reason == 2 and sg.pairs == so.pairs
    assert abs(sg.deltatmin - so.deltatmin) <= 1e-12*so.deltatmin
    G, O_ = g.particles(), o.particles()
    sc = o.scales(len(O_))
    fr = O_["IsFree"] == 1
    fx = ~fr
    _check("a", G["a"][fr], O_["a"][fr], sc[fr, 0:3])
    _check("dDensity", G["dDensity"][fr], O_["dDensity"][fr], sc[fr, 3])
    _check("wall Pressure", G["Pressure"][fx], O_["Pressure"][fx], sc[fx, 4])
    _check("wall NSv", G["NSv"][fx], O_["NSv"][fx], sc[fx, 5:8])
    print(f"[fullsize] force pass ok, pairs={sg.pairs}", flush=True)
