"""On-disk formats of the run outputs (paper_freefall.py:108-155, run.py:37-38,
paper_impact.py:65-72): states.npz keys, JSON writer, knife edge height."""

import json
import math

import numpy as np
import pytest

from blademaster_hip.runio import StateLog, edge_height, load_states, write

STATE_KEYS = ["position", "velocity", "deformation", "plastic_deformation", "sides", "damage", "blade_position",
              "blade_velocity", "blade_orientation", "blade_omega", "time"]


def fake_state(n, t):
    rng = np.random.default_rng(int(t * 1000))
    return {
        "position": rng.normal(size=(n, 3)).astype(np.float32),
        "velocity": rng.normal(size=(n, 3)).astype(np.float32),
        "deformation": np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)),
        "plastic_deformation": np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)),
        "sides": rng.integers(0, 3, n).astype(np.int64),
        "damage": rng.random(n).astype(np.float32),
        "blade_position": rng.normal(size=(1, 3)).astype(np.float32),
        "blade_velocity": rng.normal(size=(1, 3)).astype(np.float32),
        "blade_orientation": np.array([[0, 0, 0, 1]], np.float32),
        "blade_omega": np.zeros((1, 3), np.float32),
        "time": t,
    }


def test_states_npz_stacks_every_state_key_and_the_reference(tmp_path):
    n = 7
    log = StateLog(fake_state(n, 0.0))
    states = [fake_state(n, 0.01 * k) for k in range(1, 4)]
    for st in states:
        log.add(st)
    ref = np.arange(3 * n, dtype=np.float64).reshape(n, 3)
    log.save(tmp_path / "states.npz", ref)
    z = load_states(tmp_path / "states.npz")
    assert sorted(z) == sorted(STATE_KEYS + ["reference"])
    assert z["position"].shape == (4, n, 3) and z["sides"].dtype == np.int64
    np.testing.assert_array_equal(z["time"], [0.0, 0.01, 0.02, 0.03])
    np.testing.assert_array_equal(z["position"][2], states[1]["position"])
    np.testing.assert_array_equal(z["reference"], ref)
    with pytest.raises(ValueError):
        log.add({"position": 0})


def test_json_writer_matches_the_reference_format(tmp_path):
    write(tmp_path / "a.json", {"b": [1, 2.5], "a": "x"})
    text = (tmp_path / "a.json").read_text()
    assert text == json.dumps({"b": [1, 2.5], "a": "x"}, indent=2) + "\n"
    with pytest.raises(ValueError):
        write(tmp_path / "b.json", {"x": float("nan")})


def test_edge_height_closed_forms():
    half = (0.0005, 0.06, 0.02)
    low, tilt = edge_height([[[0.0, 0.0, 0.1]]], [[[0.0, 0.0, 0.0, 1.0]]], half)
    assert low.shape == (1, 1) and abs(low[0, 0] - 0.08) < 1e-15 and abs(tilt[0, 0]) < 1e-12
    a = math.radians(25.0)
    q = [math.sin(a / 2), 0.0, 0.0, math.cos(a / 2)]  # about x: the edge line tilts in y-z
    low, tilt = edge_height([[[0.0, 0.0, 0.1]]], [[q]], half)
    expected = 0.1 - half[2] * math.cos(a) - half[1] * math.sin(a)
    assert abs(low[0, 0] - expected) < 1e-12
    assert abs(tilt[0, 0] - 25.0) < 1e-9


def test_drop_scenes_release_pose_and_command_values():
    """The knife drops of blademaster_hip.freefall: the release height of the tilted knife is
    the reference's closed form (paper_freefall.py:64-76: h_z cos a + h_y sin a about the
    thickness axis, h_z cos a + h_x sin a about the length axis, above the clearance); the
    8960-particle scene has its particle count; --set values parse as numbers and tuples."""
    from dataclasses import replace

    from blademaster_hip import freefall
    from blademaster_hip.geometry import points

    fig3 = freefall.SCENES["fig3"]
    hx, hy, hz = fig3.knife_half_size
    a = math.radians(25.0)
    q, h = freefall.release_pose(fig3, 0.08)
    np.testing.assert_allclose(q, [math.sin(a / 2), 0, 0, math.cos(a / 2)], atol=1e-15)
    assert abs(h - (0.08 + 0.015 + hz * math.cos(a) + hy * math.sin(a))) < 1e-15
    q, h = freefall.release_pose(replace(fig3, tilt_axis="length"), 0.08)
    np.testing.assert_allclose(q, [0, math.sin(a / 2), 0, math.cos(a / 2)], atol=1e-15)
    assert abs(h - (0.08 + 0.015 + hz * math.cos(a) + hx * math.sin(a))) < 1e-15
    with pytest.raises(ValueError):
        freefall.release_pose(replace(fig3, tilt_axis="edge"), 0.08)
    knife = freefall.SCENES["gravity-knife"]
    cfg = knife.config()
    ref, _, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    assert len(ref) == 8960 and tuple(cells + 1) == (29, 11, 5)
    assert knife.steps() == 8000 and cfg.free_blade
    q, h = freefall.release_pose(knife, cfg.size[2])
    np.testing.assert_allclose(q, [0, 0, 0, 1])
    assert abs(h - (cfg.size[2] + 0.002 + cfg.blade_half_size[2])) < 1e-15
    assert freefall._value("0.2") == 0.2 and freefall._value("3") == 3
    assert freefall._value("0.06,0.04,0.03") == (0.06, 0.04, 0.03)
    assert freefall._value("true") is True and freefall._value("length") == "length"
    with pytest.raises(ValueError):
        replace(knife, duration=0.00001).steps()


TRAJECTORY_KEYS = {"log", "qpos", "particles", "codes", "origin", "deformation", "plastic_deformation", "velocity",
                   "energy", "control_tcp", "cycle_trace", "edge_heights", "radius", "interval", "frame_stride"}
REPORT_KEYS = {"status", "exception", "sim_s", "wall_s", "realtime_ratio", "warmup_s", "recording", "display",
               "pause_wall_s", "live_setup_s", "timing_scope", "time_s", "material_present", "warnings",
               "surface_wall_s", "surface_phase_s", "material_interval_phase_s"}


class _Arrays:
    def __init__(self, a):
        self.a = a

    def numpy(self):
        return self.a


def test_trajectory_and_report_formats(tmp_path):
    """trajectory.npz, columns.json and report.json carry the keys of franka_run.py:600-690;
    the log columns tile the row (nq = 7 pose values and nv = 6 twist values for the arm)."""
    from blademaster_hip.runio import TrajectoryRecorder, log_columns, run_report

    n = 5
    solver = type("S", (), {})()
    solver.x, solver.v = _Arrays(np.zeros((n, 3))), _Arrays(np.ones((n, 3)))
    solver.codes = _Arrays(np.zeros(n, np.int64))
    solver.F = solver.Fp = _Arrays(np.tile(np.eye(3), (n, 1, 1)))
    rec = TrajectoryRecorder(nq=7, nv=6, interval=0.001, frame_stride=2)
    frames = 0
    for k in range(4):
        if rec.step((k + 1) * 0.001, np.zeros(3), np.ones(3), np.arange(6), np.arange(7), np.zeros(6),
                    np.full(7, np.nan), np.full(7, np.nan), np.arange(6)):
            rec.frame(np.arange(7), solver, np.zeros(8), np.zeros(7), 0.01)
            frames += 1
    with pytest.raises(ValueError):
        rec.step(0.0, np.zeros(3), np.zeros(3), np.zeros(6), np.zeros(6), np.zeros(6), np.zeros(7), np.zeros(7),
                 np.zeros(6))
    rec.save(tmp_path, np.zeros(3), 0.0015, ["time_s", "a", "b"])
    with np.load(tmp_path / "trajectory.npz", allow_pickle=False) as z:
        assert set(z.files) == TRAJECTORY_KEYS
        assert z["log"].shape == (4, log_columns(7, 6)["material_generalized_force"][1])
        assert z["particles"].shape == (frames, n, 3) and frames == 2
    cols = json.loads((tmp_path / "columns.json").read_text())
    spans = sorted([v] if isinstance(v, int) else list(range(*v)) for v in cols.values())
    assert sorted(i for s in spans for i in s) == list(range(13 + 7 + 6 + 14 + 6))
    report = run_report("completed", None, 0.004, 0.002, {"material": 0.001}, {"submission": 0.0}, True)
    assert set(report) == REPORT_KEYS and report["realtime_ratio"] == 2.0
