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
