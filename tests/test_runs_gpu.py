"""Run drivers on the GPU writing the reference's on-disk formats."""

import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import freefall  # noqa: E402
from blademaster_hip.runio import load_states  # noqa: E402

RESULT_KEYS = {"status", "component_sizes", "wall_s", "orientation_change_deg", "horizontal_displacement_m",
               "final_blade_height_m", "final_lowest_edge_m", "particle_count", "assigned_particles",
               "assigned_below_threshold_final", "released_particles", "directional_contact_work_j", "gpu_errors",
               "minimum_det_f_sampled"}


def test_freefall_writes_states_manifest_and_result(tmp_path):
    out = tmp_path / "drop"
    result = freefall.run(out, duration=0.04, size=(0.06, 0.04, 0.03))
    assert set(result) == RESULT_KEYS and result["status"] == "complete"
    z = load_states(out / "states.npz")
    assert {"position", "sides", "blade_position", "time", "reference"} <= set(z)
    frames = z["position"].shape[0]
    assert frames == 1 + 4 and z["position"].shape[1] == result["particle_count"]
    # The knife is released at rest above the block: it falls under gravity.
    dz = z["blade_position"][:, 0, 2]
    assert np.all(np.diff(dz) < 0)
    manifest = json.loads((out / "manifest.json").read_text())
    assert manifest["status"] == "complete" and manifest["completed_steps"] == 20
    assert manifest["config"]["free_blade"] is True
    assert sum(result["component_sizes"]) == result["particle_count"]


def test_coupled_run_records_trajectory_and_report(tmp_path):
    """The single-knife coupled loop writes the reference's run files (franka_run.py:600-690)."""
    from blademaster_hip import surrogate
    from blademaster_hip.presets import banana_config

    task = surrogate.ThreeCutTask(1, seed=0)
    cfg = surrogate.task_configs(1)[0]
    summary, forces = surrogate.plain_run(cfg, task, steps=30, record=tmp_path / "run", frame_stride=10)
    assert summary["robot_steps"] == 30 and np.isfinite(forces).all()
    with np.load(tmp_path / "run" / "trajectory.npz", allow_pickle=False) as z:
        assert z["log"].shape[0] == 30 and z["particles"].shape[0] == 3
        assert z["particles"].shape[1] == z["codes"].shape[1]
        np.testing.assert_allclose(z["log"][:, 0], 0.001 * np.arange(1, 31), atol=1e-12)
    report = json.loads((tmp_path / "run" / "report.json").read_text())
    assert report["status"] == "completed" and report["recording"] is True
    assert abs(report["sim_s"] - 0.03) < 1e-12
