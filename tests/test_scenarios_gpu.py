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
