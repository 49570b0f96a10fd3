"""Several free knives (extra_blade, contact.py:500-566; paper Fig. 6) on the GPU: the
reference's tests/test_multi_tool.py and substeps against the oracle.

Bars: substeps from identical inputs match the float32 oracle (codes bit-exact; x, v, F and
every knife's pose and damage within 1e-5 relative); mirror symmetry and slot swaps within the
reference's float-atomic bars (1e-7 m, 1e-5 m/s, 1e-6 m).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver  # noqa: E402
from oracle import OracleSim  # noqa: E402


def two_knives(count=2, **kwargs):
    cfg = Config(size=(0.06, 0.03, 0.03), grid_spacing=0.0075, dt=0.0005, free_blade=True, tool_count=count,
                 blade_half_size=(0.0005, 0.025, 0.02), **kwargs)
    return Solver(cfg)


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_only_free_knives_may_be_several():
    for kwargs in ({}, {"pose_servo": True}, {"external_tool": True}):
        with pytest.raises(ValueError):
            Config(tool_count=2, **kwargs)
    with pytest.raises(ValueError):
        Config(tool_count=0)


def dropped(xs, steps, graph=False):
    s = two_knives()
    z = 0.03 + 0.02 + 0.002
    s.tool_pos.assign([[x, 0.0, z] for x in xs])
    s.tool_vel.assign([[0.0, 0.0, -0.6]] * 2)
    if graph:
        s.advance(np.zeros((steps, 3)), cutting=True)
    else:
        for _ in range(steps):
            s.step(cutting=True)
    return s, s.state()


def test_knife_index_does_not_matter():
    _, a = dropped([-0.015, 0.015], 100)
    _, b = dropped([0.015, -0.015], 100)
    np.testing.assert_allclose(a["blade_position"], b["blade_position"][::-1], atol=1.0e-7)
    np.testing.assert_allclose(a["position"], b["position"], atol=1.0e-6)
    np.testing.assert_allclose(a["tool_damage"], b["tool_damage"][::-1], atol=1.0e-6)


@pytest.mark.parametrize("graph", [False, True])
def test_mirrored_knives_move_alike_and_label_their_own_digits(graph):
    steps = 60
    s, state = dropped([-0.015, 0.015], steps, graph)
    p, v = state["blade_position"], state["blade_velocity"]
    np.testing.assert_allclose(p[0] * [-1, 1, 1], p[1], atol=1.0e-7)
    np.testing.assert_allclose(v[0] * [-1, 1, 1], v[1], atol=1.0e-5)
    assert 0.0 < p[0, 2] - 0.02 < 0.028
    assert v[0, 2] > -0.6 - 9.81 * steps * 0.0005
    codes = state["sides"]
    first, second = codes % 3, (codes // 3) % 3
    assert set(np.unique(first)) == {0, 1, 2} and set(np.unique(second)) == {0, 1, 2}
    assert np.all(codes < 9)
    ref = s.reference
    assert ref[first == 1, 0].max() < -0.015 < ref[first == 2, 0].min()
    assert ref[second == 1, 0].max() < 0.015 < ref[second == 2, 0].min()
    assert state["tool_damage"].shape == (2, s.n)


def test_knife_beside_the_block_falls_freely():
    s = two_knives()
    z0 = 0.2
    s.tool_pos.assign([[-0.015, 0.0, 0.03 + 0.02 + 0.002], [0.2, 0.0, z0]])
    n = 100
    for _ in range(n):
        s.step(cutting=True)
    state = s.state()
    dt, g = 0.0005, 9.81
    assert state["blade_velocity"][1, 2] == pytest.approx(-g * n * dt, rel=1.0e-5)
    assert state["blade_position"][1, 2] == pytest.approx(z0 - g * dt * dt * n * (n + 1) / 2, abs=1.0e-6)
    np.testing.assert_allclose(state["blade_position"][1, :2], [0.2, 0.0], atol=1.0e-9)


@pytest.mark.parametrize("first", [0, 1])
def test_each_knife_pays_its_own_cutting_work(first):
    s = two_knives(cutting_energy=1.0)
    x = s.reference.copy()
    x[:, 0] += 10
    x[0] = [0, 0, 0.039]
    s.x.assign(x)
    positions = np.array([[3.0, 0, 0.06], [3.0, 0, 0.06]])
    positions[first, 0] = 0
    s.tool_pos.assign(positions)
    s.tool_vel.assign([[0, 0, -1], [0, 0, -1]])
    s.rigid_contacts(True)
    initial_impulse = s.reaction.numpy()[first].copy()
    assert np.abs(initial_impulse).max() > 0
    assert s.tool_damage[first].numpy()[0] == 1
    assert s.tool_damage[1 - first].numpy()[0] == 0
    positions[:, 0] = positions[::-1, 0]
    s.tool_pos.assign(positions)
    s.reaction.zero_()
    s.torque.zero_()
    s.rigid_contacts(True)
    np.testing.assert_allclose(s.reaction.numpy()[1 - first], initial_impulse, rtol=1e-6)
    assert s.tool_damage[1 - first].numpy()[0] == 1
    s._set_cut(2)
    assert all(not damage.numpy().any() for damage in s.tool_damage)


def test_cut_indices_leave_room_for_every_knife():
    s = two_knives(count=3)
    with pytest.raises(ValueError):
        s._set_cut(s.max_cut_index - 1)
    assert s._set_cut(s.max_cut_index - 2) == s.max_cut_index - 2


@pytest.mark.parametrize("count", [2, 3])
def test_substeps_match_oracle(count):
    """Knives at x = -15, 0 (, +15) mm dropping into the block at 0.6 m/s: each substep from
    identical inputs against the float32 oracle."""
    s = two_knives(count=count)
    xs = [-0.015, 0.015, 0.0][:count]
    z = 0.03 + 0.02 + 0.002
    s.tool_pos.assign([[x, 0.0, z] for x in xs])
    s.tool_vel.assign([[0.0, 0.0, -0.6]] * count)
    o = OracleSim(s.config, s.reference, s.node_numpy, s.weight_numpy, s._gradient_numpy, s.node_count, 32,
                  s.max_cut_index)
    cut = 0
    for step in range(40):
        for name in ("x", "v", "F", "Fp", "stress", "damage", "histories", "tool_pos", "tool_vel", "tool_rot",
                     "tool_omega"):
            o.set(name, getattr(s, name).numpy())
        o.set("codes", s.codes.numpy())
        o.set("extra_histories", s.extra_histories.numpy())
        o.set("extra_damage", np.stack([d.numpy() for d in s.tool_damage[1:]]))
        s.step(cutting=True)
        o.substep(cutting=True)
        np.testing.assert_array_equal(s.codes.numpy(), o.get("codes"), err_msg=f"step {step}")
        for name in ("x", "v", "F"):
            assert rel(getattr(s, name).numpy(), o.get(name)) < 1e-5, (step, name)
        for name in ("tool_pos", "tool_vel"):
            assert rel(getattr(s, name).numpy(), o.get(name)) < 1e-5, (step, name)
        np.testing.assert_allclose(np.stack([d.numpy() for d in s.tool_damage[1:]]).ravel(), o.get("extra_damage"),
                                   rtol=1e-5, atol=1e-7)
        cut = max(cut, int(np.count_nonzero(s.codes.numpy() // 3)))
    assert cut > 0  # the extra knives labelled particles
