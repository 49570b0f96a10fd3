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


# Column layout of the per-step log of a coupled run (franka_run.py:620-633): time, edge
# target, tool position, interval wrench, the arm's generalized positions (nq) and
# velocities (nv), requested and commanded arm torques (7 each) and the material's generalized
# force (nv).
def log_columns(nq, nv):
    q_end = 13 + nq
    v_end = q_end + nv
    return {"time": 0, "target": [1, 4], "tool_position": [4, 7], "wrench": [7, 13], "qpos": [13, q_end],
            "qvel": [q_end, v_end], "requested_arm_torque": [v_end, v_end + 7],
            "commanded_arm_torque": [v_end + 7, v_end + 14], "material_generalized_force": [v_end + 14, v_end + 14 + nv]}


class TrajectoryRecorder:
    """The recording of a coupled cutting run in the reference's trajectory.npz
    (franka_run.py:636-664): a log row per robot step (layout `log_columns`) and, every
    `frame_stride` steps, a frame of the arm state, the particles (positions, codes, F, Fp,
    velocities), the material energies, the commanded tool pose and the knife edge height."""

    def __init__(self, nq, nv, interval, frame_stride=10):
        self.nq, self.nv = int(nq), int(nv)
        self.interval, self.frame_stride = float(interval), int(frame_stride)
        self.rows, self.frames = [], {k: [] for k in ("qpos", "particles", "codes", "deformation",
                                                      "plastic_deformation", "velocity", "energy", "control_tcp",
                                                      "edge_heights")}
        self.steps = 0

    def step(self, time_s, target, tool_position, wrench, qpos, qvel, requested_torque, commanded_torque,
             material_force):
        row = np.concatenate([[time_s], np.ravel(target), np.ravel(tool_position), np.ravel(wrench), np.ravel(qpos),
                              np.ravel(qvel), np.ravel(requested_torque), np.ravel(commanded_torque),
                              np.ravel(material_force)]).astype(np.float64)
        width = log_columns(self.nq, self.nv)["material_generalized_force"][1]
        if row.size != width:
            raise ValueError(f"log row of {row.size} values, the layout has {width}")
        self.rows.append(row)
        self.steps += 1
        return self.steps % self.frame_stride == 0

    def frame(self, qpos, solver, energy_row, control_tcp, edge_height):
        f = self.frames
        f["qpos"].append(np.asarray(qpos, dtype=np.float64))
        f["particles"].append(solver.x.numpy())
        f["codes"].append(solver.codes.numpy())
        f["deformation"].append(solver.F.numpy())
        f["plastic_deformation"].append(solver.Fp.numpy())
        f["velocity"].append(solver.v.numpy())
        f["energy"].append(np.asarray(energy_row, dtype=np.float64))
        f["control_tcp"].append(np.asarray(control_tcp, dtype=np.float64))
        f["edge_heights"].append(float(edge_height))

    def save(self, out, origin, radius, energy_columns):
        out = Path(out)
        arrays = {k: np.asarray(v) for k, v in self.frames.items()}
        np.savez_compressed(out / "trajectory.npz", log=np.asarray(self.rows), origin=np.asarray(origin),
                            cycle_trace=np.zeros((0,)), radius=float(radius), interval=self.interval,
                            frame_stride=self.frame_stride, **arrays)
        write(out / "energy_columns.json", list(energy_columns))
        write(out / "columns.json", log_columns(self.nq, self.nv))


def run_report(status, exception, sim_s, wall_s, phase_s, port_timing, recording, warmup_s=0.0, warnings=()):
    """report.json of a coupled run, with the fields of franka_run.py:666-690: status, the
    simulated and wall time and their ratio, the phase times and the coupling port's timing."""
    return {
        "status": status, "exception": exception, "sim_s": sim_s, "wall_s": wall_s,
        "realtime_ratio": sim_s / wall_s if wall_s > 0 else None, "warmup_s": warmup_s, "recording": bool(recording),
        "display": "none", "pause_wall_s": 0.0, "live_setup_s": 0.0,
        "timing_scope": "material intervals and the arm integration; excludes graph capture and the recording",
        "time_s": dict(phase_s), "material_present": True, "warnings": list(warnings), "surface_wall_s": 0.0,
        "surface_phase_s": {}, "material_interval_phase_s": dict(port_timing),
    }
