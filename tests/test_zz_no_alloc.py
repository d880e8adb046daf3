"""The no-allocation pair query (psph_neighbour_pairs_noalloc) through the C ABI against the
allocating one (psph_neighbour_pairs): identical oriented pair lists, in the same order, on two
calls in a row whose second has a smaller input, then the size limit's error."""
import ctypes

import numpy as np
import pytest

from persiansph_amd import Solver, cases
from persiansph_amd import solver as S

pytestmark = pytest.mark.gpu


def _query(g, noalloc):
    L = S.lib()
    f = L.psph_neighbour_pairs_noalloc if noalloc else L.psph_neighbour_pairs
    n = ctypes.c_size_t()
    assert f(g._h, None, 0, ctypes.byref(n)) == 0
    out = np.full((n.value, 2), -7, dtype=np.int64)
    rc = f(g._h, out.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n))
    assert rc == 0, rc
    return out


def test_noalloc_pairs_equal_allocating_pairs():
    big = cases.dam_break(3, 0.04, (0.8, 0.6, 0.4), (0.4, 0.3, 0.4), jitter=0.01, seed=7)
    small = cases.dam_break_2d(dx=0.05)
    sizes = []
    for name, (cfg, p, dt) in (("3-D dam break", big), ("2-D dam break (smaller)", small)):
        print(f"[noalloc] {name} n={len(p)}", flush=True)
        g = Solver(cfg, p, dt)
        g.run(3)
        a = _query(g, noalloc=False)
        b = _query(g, noalloc=True)
        assert len(a) > 1000 and np.array_equal(a, b), name
        sizes.append(len(a))
        # the run continues as after the allocating query
        g.run(2)
        h = Solver(cfg, p, dt)
        h.run(3)
        h.pairs()
        h.run(2)
        np.testing.assert_array_equal(g.particles()["x"], h.particles()["x"])
        g.close()
        h.close()
    assert sizes[1] < sizes[0]


def test_noalloc_pairs_reject_more_than_the_staging_holds():
    """The 3-D dam break at dx = 0.02 m has more pairs than the 2^21 the staging array holds: the
    count comes back, the call fails with PSPH_ERR_CAPACITY, and the handle stays usable."""
    cfg, p, dt = cases.dam_break_3d(dx=0.02)
    print(f"[noalloc limit] n={len(p)}", flush=True)
    g = Solver(cfg, p, dt)
    L = S.lib()
    n = ctypes.c_size_t()
    assert L.psph_neighbour_pairs_noalloc(g._h, None, 0, ctypes.byref(n)) == 0
    assert n.value > 2**21
    out = np.zeros((n.value, 2), dtype=np.int64)
    rc = L.psph_neighbour_pairs_noalloc(g._h, out.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n))
    assert rc == -8
    buf = ctypes.create_string_buffer(512)
    L.psph_last_error(buf, 512)
    assert b"no-allocation" in buf.value
    assert len(g.pairs()) == n.value
    g.run(2)
