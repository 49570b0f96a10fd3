"""Single operators on caller arrays (bm_probe_resist_cut, bm_couple, bm_g2p) against the
reference's own known answers and the oracle.

- resist_cut: tests/test_cut_resistance.py:45-67 (impulses, damage [1, 1, 1, .5, .54, 1],
  work 4e-5), for a level and a 35-degree knife;
- couple: the five-field node of tests/test_mechanics.py:80-105 (codes 0, 1, 2, 4, 7) within
  4e-7 of the float64 oracle, momentum conserved within 8e-7, and random multi-node tables;
- g2p: tests/test_particle_transport.py:10-45 (a resting grid leaves positions unchanged;
  v = [2 + 6 f, 4, 6 - 4 f], x = dt new_v), and random tables against the oracle.
"""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402
from blademaster_hip import operators  # noqa: E402


@pytest.mark.parametrize("angle", [0, 35])
def test_resist_cut_known_answer(angle):
    rotation = Rotation.from_euler("y", angle, degrees=True)
    q = rotation.as_quat()
    edge = np.array([0.0, 0.0, 0.02]) + rotation.apply([0.0, 0.0, -0.02])
    x = edge + rotation.apply([[0, 0, -0.001]] * 5 + [[0, 0, 0.001]])
    tool_velocity = rotation.apply([[0, 0, -speed] for speed in [0.01, 0, -0.01, -0.01, 0.01, 0]])
    j = rotation.apply([0.002, 0.001, -0.004])
    out, damage, work = operators.resist_cut(j, x, np.zeros((6, 3)), [0.0, 0.0, 0.02], tool_velocity, q, np.zeros(3),
                                             [0.0005, 0.025, 0.02], 0.0015, 0.006, [1, 1, 1, 0.5, 0.5, 1], 0.001)
    expected = rotation.apply([[0.002, 0.001, 0]] * 3 + [[0.002, 0.001, -0.004]] * 3)
    np.testing.assert_allclose(out, expected, atol=2e-9)
    np.testing.assert_allclose(damage, [1, 1, 1, 0.5, 0.54, 1], atol=1e-7)
    assert work == pytest.approx(0.004 * 0.01, rel=1e-6)


def test_couple_five_field_node():
    codes = np.array([0, 1, 2, 4, 7], np.int64)
    masses = np.array([0.2, 0.5, 0.7, 0.3, 0.4], np.float32)
    states = np.random.default_rng(3).normal(size=(5, 3)).astype(np.float32)
    cap = 16
    keys = np.full(cap, -1, np.int64)
    keys[:5] = codes
    mass = np.zeros(cap, np.float32)
    mass[:5] = masses
    state = np.zeros((cap, 3), np.float32)
    state[:5] = states
    heads = np.array([0], np.int32)
    links = np.array([1, 2, 3, 4, -1] + [-1] * 11, np.int32)
    result = operators.couple(keys, heads, links, mass, state, 1)
    want = oracle.couple(keys, heads, links, mass, state, 1)
    np.testing.assert_allclose(result[:5], want[:5], atol=4e-7)
    np.testing.assert_allclose(result[:5].sum(0), states.sum(0), atol=8e-7)
    assert np.all(result[5:] == 0.0)  # unused fields are not written


def test_couple_random_tables_match_oracle():
    rng = np.random.default_rng(11)
    node_count, count = 40, 160
    node = rng.integers(0, node_count, count)
    digits = rng.integers(0, 3, (count, 3))
    code = digits @ np.array([1, 3, 9])
    keys = (code * node_count + node).astype(np.int64)
    keys[rng.random(count) < 0.05] = -1
    heads = np.full(node_count, -1, np.int32)
    links = np.full(count, -1, np.int32)
    for a in range(count):  # push-front lists, as registration builds them
        links[a] = heads[node[a]]
        heads[node[a]] = a
    mass = rng.uniform(0.0, 1.0, count).astype(np.float32)
    mass[rng.random(count) < 0.05] = 0.0
    state = rng.normal(size=(count, 3)).astype(np.float32)
    result = operators.couple(keys, heads, links, mass, state, node_count)
    want = oracle.couple(keys, heads, links, mass, state, node_count)
    active = (keys >= 0) & (mass > 0)
    np.testing.assert_allclose(result[active], want[active], atol=2e-6)
    assert np.all(result[~active] == 0.0)


@pytest.mark.parametrize("flip", [0.0, 0.95, 1.0])
def test_g2p_resting_grid_keeps_positions(flip):
    initial = np.array([[0.0, 0.0, 0.0], [0.003, 0.0, 0.0]], np.float32)
    velocity = np.array([[10.0, 0.0, 0.0], [-10.0, 0.0, 0.0]], np.float32)
    slots = np.tile(np.arange(8), (2, 1))
    weights = np.full((2, 8), 1 / 8)
    grid = np.zeros((8, 3))
    x, v = operators.g2p(initial, velocity, slots, weights, grid, grid, 0.001, flip)
    np.testing.assert_array_equal(x, initial)
    np.testing.assert_allclose(v, flip * velocity, atol=1e-6)


@pytest.mark.parametrize("flip", [0.0, 0.95, 1.0])
def test_g2p_uniform_grid_acceleration(flip):
    old = np.tile([1.0, 2.0, 3.0], (8, 1))
    new = np.tile([2.0, 4.0, 6.0], (8, 1))
    x, v = operators.g2p(np.zeros((1, 3)), [[7.0, 2.0, -1.0]], np.arange(8).reshape(1, 8), np.full((1, 8), 1 / 8),
                         old, new, 0.001, flip)
    np.testing.assert_allclose(x, [[0.002, 0.004, 0.006]], rtol=1e-6)
    np.testing.assert_allclose(v, [[2.0 + 6 * flip, 4.0, 6.0 - 4 * flip]], rtol=1e-6)


def test_g2p_random_tables_match_oracle():
    rng = np.random.default_rng(5)
    n, grid = 500, 300
    x0 = rng.normal(size=(n, 3))
    v0 = rng.normal(size=(n, 3))
    slots = rng.integers(0, grid, (n, 8))
    weights = rng.random((n, 8))
    weights /= weights.sum(1, keepdims=True)
    old, new = rng.normal(size=(grid, 3)), rng.normal(size=(grid, 3))
    x, v = operators.g2p(x0, v0, slots, weights, old, new, 0.002, 0.95)
    xo, vo = oracle.g2p(x0, v0, slots, weights, old, new, 0.002, 0.95)
    np.testing.assert_allclose(x, xo, atol=1e-6)
    np.testing.assert_allclose(v, vo, atol=1e-5)
