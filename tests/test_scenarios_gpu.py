"""Prescribed-knife scenes through ExternalToolPort (the counterpart of scenarios.run): the
three sequential cuts of the log separate it into four pieces, and the run writes the
reference's outputs."""

import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import scenarios as sc  # noqa: E402
from blademaster_hip.runio import load_states  # noqa: E402


def test_three_sequential_cuts_leave_four_pieces(tmp_path):
    meta = sc.run(tmp_path / "run", "three-cuts", log=lambda *a: None)
    assert meta["status"] == "completed" and meta["min_det_F"] > 0
    pieces = np.array(meta["components"])
    n = meta["particles"]
    assert len(pieces) == 4 and pieces.sum() == n, pieces
    # The 40 mm end pieces and the two middle pieces each hold about a quarter of the log.
    assert np.all(pieces > 0.18 * n), pieces
    saved = json.loads((tmp_path / "run" / "manifest.json").read_text())
    assert saved["components"] == meta["components"] and saved["scenario"] == "three-cuts"
    states = load_states(tmp_path / "run" / "states.npz")
    intervals = int(round(sc.keyframes("three-cuts")[-1][0] / 0.002))
    assert states["forces"].shape == (intervals, 7)
    assert states["energy"].shape[1] == 7
    assert states["reference"].shape == (n, 3)
    assert states["position"].shape[1:] == (n, 3) and len(states["position"]) == len(states["energy"])
    # The knife pushed down on the log while cutting: a net upward reaction on the knife.
    fz = states["forces"][:, 3]
    assert fz.max() > 0.01


def test_reference_scene_runs_through_the_port(tmp_path):
    meta = sc.run(tmp_path / "run", "fig8", duration=0.5, log=lambda *a: None)
    assert meta["status"] == "completed" and meta["completed_sim_s"] == pytest.approx(0.5)


@pytest.mark.parametrize("scene", ["fig7-shallow", "fig7-deep", "fig8", "fig4", "cross-shallow"])
def test_crossing_cut_scenes_stay_within_the_field_slots(tmp_path, scene):
    """The field table holds fields_per_node codes per node (16 in these scenes), where the
    reference's hash table has no cap: with 4, the crossing cuts of Fig. 7 and Fig. 4 overflow
    (2 to 28 nodes).  The crossing cuts of Fig. 7, the slices of Fig. 8, the two radial cuts and
    lift of Fig. 4 and the crossed demo cuts run their whole keyframed duration without a
    storage overflow (the port raises on any device error count), and the material stays
    valid."""
    meta = sc.run(tmp_path / "run", scene, log=lambda *a: None)
    keys = sc.keyframes(scene)
    assert meta["status"] == "completed", meta.get("error")
    assert meta["completed_sim_s"] == pytest.approx(keys[-1][0], abs=2 * max(0.002, meta["config"]["dt"]))
    assert meta["min_det_F"] > 0
    sides = load_states(tmp_path / "run" / "states.npz")["sides"][-1]
    assert np.count_nonzero(sides) > 0  # the knife labelled the material it crossed
