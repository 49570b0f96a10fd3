"""Host logic of the force-convergence driver: presets and the summary statistics
(force_convergence.py:132-164) on synthetic run records with closed-form answers."""

import numpy as np

from blademaster_hip.convergence import DESCENT_S, HOLD_S, INTERVAL, summarize
from blademaster_hip.presets import banana_config, banana_dt


def test_banana_dt_divides_the_robot_step_and_respects_cfl():
    cfg = banana_config()
    for spacing in (0.006, 0.005, 0.004, 0.003):
        dt = banana_dt(cfg, spacing, INTERVAL)
        n = INTERVAL / dt
        assert abs(n - round(n)) < 1e-9 and dt <= 0.0000625 + 1e-15
    assert banana_dt(cfg, 0.006) == 0.0000625


def test_summarize_closed_form():
    steps = round((DESCENT_S + HOLD_S) / INTERVAL)
    t = (np.arange(steps) + 1) * INTERVAL
    runs = []
    for seed, level in enumerate((1.0, 3.0)):
        fz = np.where(t <= DESCENT_S, level, 0.0)
        force = np.stack([np.zeros(steps), np.zeros(steps), fz], axis=1)
        runs.append(dict(spacing=0.004, seed=seed, x_offset_m=0.0, dt=5e-5, particles=100, force_n=force.tolist(),
                         pieces=[60, 40] if seed else [100], wall_s=2.0 * (seed + 1), gpu_errors=[0, 0, 0]))
    out = summarize(runs)["by_spacing"]["4 mm"]
    assert out["samples"] == 2 and out["seeds"] == 2 and out["severed_runs"] == 1
    assert abs(out["steady_force_n"]["mean"] - 2.0) < 1e-12
    assert abs(out["steady_force_n"]["std"] - np.std([1.0, 3.0], ddof=1)) < 1e-12
    assert abs(out["descent_impulse_ns"]["mean"] - 2.0 * DESCENT_S) < 1e-9
    assert out["peak_single_run_n"] == dict(mean=2.0, max=3.0)
    assert abs(out["wall_s_per_run"] - 3.0) < 1e-12
    assert len(out["mean_trace_n"]) == steps // 20
