"""Run outputs in the reference's on-disk formats.

- `write`: JSON with indent 2 and no NaN (`blademaster_warp/run.py:37-38`).
- `git_revision`: {"commit", "dirty"} of the repository (`provenance.py:9-21`).
- `StateLog`: the `states.npz` archive of paper_freefall and scenarios
  (`paper_freefall.py:108, 128-131`; `scenarios.py:406-412`): every key of `Solver.state()`
  stacked over snapshots, plus `reference`.
- `edge_height`: lowest edge point and tilt of each knife pose (`paper_impact.py:65-72`).
"""

import json
import subprocess
from pathlib import Path

import numpy as np
from scipy.spatial.transform import Rotation

ROOT = Path(__file__).resolve().parents[2]


def write(path, data):
    Path(path).write_text(json.dumps(data, indent=2, allow_nan=False) + "\n")


def git_revision():
    def git(*args):
        return subprocess.run(["git", "-C", str(ROOT), *args], capture_output=True, text=True, check=True).stdout.strip()

    try:
        return {"commit": git("rev-parse", "HEAD"), "dirty": bool(git("status", "--porcelain", "--untracked-files=no"))}
    except (OSError, subprocess.CalledProcessError):
        return {"commit": None, "dirty": None}


class StateLog:
    """Snapshots of `Solver.state()`; `save` writes them as the reference's states.npz."""

    def __init__(self, first_state=None):
        self.saved = {}
        if first_state is not None:
            self.add(first_state)

    def add(self, state):
        if self.saved and set(state) != set(self.saved):
            raise ValueError("State keys changed between snapshots")
        for k, v in state.items():
            self.saved.setdefault(k, []).append(v)

    def __getitem__(self, key):
        return self.saved[key]

    def save(self, path, reference):
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in self.saved.items()}, reference=reference)


def load_states(path):
    """The arrays of a states.npz archive (no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def edge_height(position, orientation, half):
    """Lowest edge point height and tilt from vertical, per frame and knife (xyzw quaternions)."""
    position, orientation = np.asarray(position), np.asarray(orientation)
    rotation = Rotation.from_quat(orientation.reshape(-1, 4))
    ends = [rotation.apply([0.0, y, -half[2]])[:, 2] for y in (-half[1], half[1])]
    lowest = np.minimum(*ends).reshape(position.shape[:2]) + position[..., 2]
    up = rotation.apply([0.0, 0.0, 1.0])[:, 2].reshape(position.shape[:2])
    return lowest, np.degrees(np.arccos(np.clip(up, -1.0, 1.0)))
