"""What the trained network gains over stepping without it: trained on plain-solver
trajectories of seeded 2-D dam-break variants, it predicts the k-step state change of a
held-out variant better than both predictors without a network (ballistic free fall and
constant-velocity drift), and the revert engine runs end to end with it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_network_beats_stepping_without_it():
    from persiansph_amd import surrogate as S
    dx, steps, k = 0.05, 400, 10
    trajs = []
    for aspect in (0.6, 1.0, 1.6, 2.0):
        cfg, p, dt = S.dam_break_variant(aspect, dx=dx, area=0.5)
        trajs.append(S.trajectory(cfg, p, dt, steps*dt, k*dt))
    cfg_h, p_h, dt_h = S.dam_break_variant(1.3, dx=dx, area=0.5)
    held = [S.trajectory(cfg_h, p_h, dt_h, steps*dt_h, k*dt_h)]
    radius = 2.0*1.3*dx
    sc = S.scales_from(trajs, k, dt_h, radius)
    print(f"[surrogate] {len(p_h)} particles, {sum(len(t) for t in trajs)} frames, scales {sc}", flush=True)
    net = S.make_network(2, 4, sc, seed=3)
    untrained = S.prediction_error(held, S.net_predictor(net), sc)
    losses = S.train(net, trajs, 800, lr=1e-3, seed=1)
    e_net = S.prediction_error(held, S.net_predictor(net), sc)
    e_bal = S.prediction_error(held, S.ballistic_predictor(k, dt_h), sc)
    e_drift = S.prediction_error(held, S.drift_predictor(k, dt_h), sc)
    print(f"[surrogate] loss {np.mean(losses[:20]):.3e} -> {np.mean(losses[-20:]):.3e}; held-out k-step error: "
          f"network {e_net:.3f} (untrained {untrained:.3f}), ballistic {e_bal:.3f}, drift {e_drift:.3f}",
          flush=True)
    assert np.mean(losses[-20:]) < 0.5*np.mean(losses[:20])
    assert e_net < min(e_bal, e_drift), (e_net, e_bal, e_drift)

    eng = S.HybridEngine(cfg_h, p_h, dt_h, S.net_predictor(net), k=k, M=50)
    out = eng.run(4)
    st = eng.stats
    print(f"[surrogate] engine {st}", flush=True)
    assert st["windows"] == 4 == st["accepted"] + st["reverted"]
    assert np.all(np.isfinite(out.x)) and out.t > 0.0
    net.close()
