"""Widened robot-material coupling with the interval-wrench network: the window logic (judge
and fallback), agreement of the batched coupling with the reference's single-knife loop, GPU
training from simulator trajectories, and what the trained network gains over holding the
last measured wrench (the widened exchange without a model)."""

import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import surrogate as S  # noqa: E402
from blademaster_hip.mlp import Mlp  # noqa: E402

NETWORK = os.path.join(os.path.dirname(S.__file__), "wrench_mlp.npz")


def test_window_judge_and_fallback():
    run = S.CutRun(2, seed=5, window=3)
    # Never fall back: the arm is driven by the held wrench.
    for _ in range(40):
        run.window_step("hold", threshold=np.inf)
    assert not np.array(run.log["fallback"]).any()
    np.testing.assert_array_equal(np.array(run.log["used"]), np.array(run.log["predicted"]))
    # Always fall back: the arm is re-integrated with the measured wrenches.
    saved = run.arm.save()
    steps = run.steps
    run.window_step("hold", threshold=-1.0)
    assert np.array(run.log["fallback"][-1]).all()
    measured = np.array(run.log["measured"][-3:])
    np.testing.assert_array_equal(np.array(run.log["used"][-3:]), measured)
    replay = S.ImpedanceArm(saved[0], saved[1])
    replay.load(saved)
    for k in range(3):
        origin, orientation, _, _ = run.task.at((steps + k) * S.INTERVAL)
        replay.step(origin, orientation, measured[k])
    np.testing.assert_array_equal(replay.p, run.arm.p)
    np.testing.assert_array_equal(replay.v, run.arm.v)
    # The next window starts from the last measured wrench.
    np.testing.assert_array_equal(run.prev, measured[-1])


def test_batched_full_coupling_follows_the_single_knife_loop():
    """The batched port with one instance and the plain Solver + ExternalToolPort loop of
    franka_run._exchange_wrench drive the same arm through the same material."""
    steps = 400
    cfg = S.task_configs(1, seed=4)[0]
    task = S.ThreeCutTask(1, seed=4, half_z=cfg.blade_half_size[2])
    _, plain = S.plain_run(cfg, task, steps=steps)
    run = S.CutRun(configs=[cfg], task=task, window=1)
    run.run(steps=steps)
    batched = np.array(run.log["measured"])[:, 0, :3]
    assert np.linalg.norm(plain, axis=1).max() > 0.1  # the knife is in the material
    # Cumulative impulse traces agree within 2 % of the total.
    ip, ib = np.cumsum(plain, 0), np.cumsum(batched, 0)
    assert np.abs(ip - ib).max() <= 0.02 * np.abs(ip).max(0).max()


def test_training_on_simulator_trajectories_lowers_the_validation_loss():
    traj, _ = S.generate(32, seed=21, steps=400)
    X, Y, _ = S.samples(traj)
    vtraj, _ = S.generate(8, seed=22, steps=400)
    Xv, Yv, _ = S.samples(vtraj)
    net, history = S.train(X, Y, epochs=4, batch=2048, lr=1e-3, validation=(Xv, Yv))
    assert history[-1] < history[0]
    assert min(net.validation_history) < net.validation_history[0]


@pytest.mark.skipif(not os.path.exists(NETWORK), reason="trained network not present")
def test_trained_network_beats_holding_the_wrench_on_unseen_tasks():
    net = Mlp.load(NETWORK)
    traj, _ = S.generate(8, seed=31, steps=1100)
    X, Y, _ = S.samples(traj)
    rows = len(X) // S.MAX_AGE
    gains = []
    for k in range(S.MAX_AGE):
        e = S.evaluate_predictors(net, X[k * rows:(k + 1) * rows], Y[k * rows:(k + 1) * rows])
        gains.append(e["mlp"] / e["hold"])
    # Two or more steps ahead the network's force error is at least 10 % below holding.
    assert gains[2] < 0.9 and gains[3] < 0.9, gains
    hold = S.CutRun(8, seed=32, window=4).run("hold", steps=1100)
    mlp = S.CutRun(8, seed=32, window=4).run("mlp", net, steps=1100)
    assert mlp["predictor_error"] < 0.95 * hold["predictor_error"], (mlp, hold)
    assert mlp["mean_residual"] < hold["mean_residual"], (mlp, hold)
    # At the same 5 % threshold the network falls back less often, at the same speed, with the
    # force felt by the arm within 2 % of the judged TLMPM force.
    assert mlp["fallback_rate"] < hold["fallback_rate"], (mlp, hold)
    assert mlp["sim_per_wall"] >= 0.9 * hold["sim_per_wall"], (mlp, hold)
    assert mlp["force_error"] <= 0.02, mlp
    # It accepts windows that hold rejects, so at the same threshold it feels more error; the
    # fair comparison is at equal fallback rate.  Holding reaches the network's fallback rate
    # only at a looser threshold, and there it feels at least the network's force error.
    curve = [S.CutRun(8, seed=32, window=4).run("hold", steps=1100, threshold=t) for t in (0.06, 0.07, 0.08, 0.1)]
    reach = [r for r in [hold] + curve if r["fallback_rate"] >= mlp["fallback_rate"]]
    assert reach, (mlp, curve)
    closest = min(reach, key=lambda r: r["fallback_rate"])
    assert mlp["force_error"] <= closest["force_error"], (mlp, closest)
