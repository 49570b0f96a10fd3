"""Scene keyframes and the prescribed-tool command (scenarios.py:110-307) on the host."""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from blademaster_hip import scenarios as sc
from blademaster_hip.geometry import points


@pytest.mark.parametrize("name", sc.SCENARIOS)
def test_scene_keyframes_are_ordered_and_configs_sample(name):
    keys = sc.keyframes(name)
    times = np.array([k[0] for k in keys])
    assert np.all(np.diff(times) > 0)
    events = [k[3] for k in keys if k[3] >= 0]
    assert events == sorted(events) and events[0] == 0
    cfg = sc.scene_config(name)
    ref, _, _ = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    assert len(ref) > 0


def test_command_interpolates_pose_and_keeps_the_largest_cut_index():
    keys = [(0.0, [0, 0, 0.1], [0, 0, 0], 0), (1.0, [0.2, 0, 0.1], [90, 0, 0], 0),
            (2.0, [0.2, 0, 0.0], [90, 0, 0], -1), (3.0, [0.2, 0.1, 0.0], [90, 0, 0], 1)]
    tool, active, index = sc.command_at(0.25, keys)
    np.testing.assert_allclose(tool.position, [0.05, 0, 0.1])
    np.testing.assert_allclose(tool.linear_velocity, [0.2, 0, 0])
    expected = Rotation.from_euler("z", 22.5, degrees=True).as_quat()
    np.testing.assert_allclose(np.abs(tool.orientation_xyzw @ expected), 1.0, atol=1e-12)
    np.testing.assert_allclose(tool.angular_velocity, [0, 0, np.pi / 2], atol=1e-12)
    assert active and index == 0
    _, active, index = sc.command_at(2.5, keys)
    assert not active and index == 0
    _, active, index = sc.command_at(10.0, keys)  # past the end: the last segment, clamped
    assert not active and index == 0
    tool, _, _ = sc.command_at(10.0, keys)
    np.testing.assert_allclose(tool.position, [0.2, 0.1, 0.0])


def test_three_cuts_scene_matches_its_statement():
    cfg = sc.scene_config("three-cuts")
    ref, _, _ = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    assert len(ref) == 11984
    # Cut planes 40, 80 and 120 mm from the -x end of the log, plunged at 50 mm/s.
    keys = sc.keyframes("three-cuts")
    end = ref[:, 0].min() - cfg.grid_spacing / 4
    for k, from_end in enumerate(sc.LOG_CUTS_FROM_END):
        cut = [key for key in keys if key[3] == k]
        assert len({key[1][0] for key in cut}) == 1
        assert cut[0][1][0] - end == pytest.approx(from_end, abs=1e-3)
        down = [key for key in cut]
        speed = (down[1][1][2] - down[2][1][2]) / (down[2][0] - down[1][0])
        assert speed == pytest.approx(sc.LOG_PLUNGE_SPEED)
    # Each cut is followed by a 10 mm push towards -x at board level.
    pushes = [(a, b) for a, b in zip(keys, keys[1:]) if a[3] == b[3] == -1 and a[1][2] == b[1][2]
              and b[1][0] != a[1][0]]
    assert len(pushes) == 3 and all(b[1][0] - a[1][0] == pytest.approx(-sc.LOG_PUSH) for a, b in pushes)
