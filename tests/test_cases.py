"""Input generation restating AddBoxLength (Source/Domain.cpp:294-495) and the dam-break drivers."""
import numpy as np

from persiansph_amd import cases


def test_add_box_length_cubic_2d_counts_and_mass():
    # the loop test xp <= V + Lx - r runs in binary floating point: 19*0.05 > 0.95, so a
    # 1.0 m row holds 9 particles, which is why the reference drivers pad lengths by dx/10
    p = cases.add_box_length(2, (0.0, 0.0, 0.0), 1.0, 0.5, 0.0, 0.05, 1000.0, 0.13, 1, 0, False)
    assert 19*0.05 > 1.0 - 0.05
    assert len(p) == 9*5
    p = cases.add_box_length(2, (0.0, 0.0, 0.0), 1.01, 0.51, 0.0, 0.05, 1000.0, 0.13, 1, 0, False)
    assert len(p) == 10*5
    np.testing.assert_allclose(p["Mass"], 0.1*0.1*1000.0)
    np.testing.assert_allclose(p["x"][0], (0.05, 0.05, 0.0))
    np.testing.assert_allclose(p["x"][-1], (0.95, 0.45, 0.0))
    assert np.all(np.diff(p["x"][:10, 0]) > 0)          # x runs fastest, as the reference pushes


def test_fast_box_equals_loop_box_3d():
    V = (-0.03, -0.03, -0.03)
    slow = cases.add_box_length(3, V, 0.2, 0.13, 0.11, 0.005, 1000.0, 0.013, 1, 0, False)
    counts = tuple(cases._loop_count(V[c], L, 0.005) for c, L in enumerate((0.2, 0.13, 0.11)))
    fast = cases._fast_cubic_box(3, V, counts, 0.005, 1000.0, 0.013)
    assert len(slow) == len(fast)
    np.testing.assert_array_equal(slow["x"], fast["x"])
    np.testing.assert_array_equal(slow["Mass"], fast["Mass"])


def test_dam_break_2d_config():
    cfg, p, dt = cases.dam_break_2d()
    fluid = p["IsFree"] == 1
    assert fluid.sum() == 5000 and (~fluid).sum() == 1518
    assert np.isclose(dt, 0.2*0.026/(10*np.sqrt(2*9.81*2.0)))
    # hydrostatic initial density, walls at rho0, three layers per wall
    top = p["x"][fluid, 1].max()
    assert np.isclose(top, 1.99)
    assert np.all(p["Density"][~fluid] == 1000.0)
    assert np.all(p["Density"][fluid] >= 1000.0)
    assert np.all(p["NoSlip"][~fluid] == 1) and np.all(p["IsFree"][~fluid] == 0)


def test_dam_break_3d_config_counts():
    cfg, p, dt = cases.dam_break_3d(dx=0.02, jitter=0.01, seed=1)
    fluid = p["IsFree"] == 1
    assert fluid.sum() == 60*30*50
    d1 = cases.dam_break_3d(dx=0.02, jitter=0.01, seed=1)[1]
    np.testing.assert_array_equal(p["x"], d1["x"])       # seeded
    lat = cases.dam_break_3d(dx=0.02, jitter=0.0)[1]
    assert np.abs(p["x"][fluid] - lat["x"][fluid]).max() <= 0.01*0.02
    assert np.all(p["x"][~fluid] == lat["x"][~fluid])
