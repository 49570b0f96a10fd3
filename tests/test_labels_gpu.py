"""GPU piece labels (union-find over the (node, code) graph) against the host restatement of
topology.components (topology.py:17-57): labels and sizes equal bit for bit."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver  # noqa: E402
from blademaster_hip.topology import components  # noqa: E402


def solver(**kw):
    base = dict(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=0.0002, cache_topology=True)
    base.update(kw)
    return Solver(Config(**base))


def check(s, codes):
    s.codes.assign(np.asarray(codes, dtype=np.int64))
    labels, sizes = s.piece_labels()
    ref_labels, ref_sizes = components(s.node_numpy, s.weight_numpy, codes)
    np.testing.assert_array_equal(labels, ref_labels)
    np.testing.assert_array_equal(sizes, ref_sizes)
    return sizes


def test_uncut_block_is_one_piece():
    s = solver()
    sizes = check(s, np.zeros(s.n, np.int64))
    assert sizes.tolist() == [s.n]


def test_planar_cut_splits_the_block_and_an_uncut_bridge_joins_it():
    s = solver()
    x = s.reference
    codes = np.where(x[:, 0] < 0, 1, 2)
    assert len(check(s, codes)) == 2
    bridge = codes.copy()
    bridge[(np.abs(x[:, 0]) < 0.004) & (x[:, 2] < 0.004)] = 0  # uncut material under the cut
    assert len(check(s, bridge)) == 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_two_cut_labels_match(seed):
    s = solver(sampling_seed=seed, size=(0.06, 0.036, 0.024))
    rng = np.random.default_rng(seed)
    x = s.reference
    c0 = np.where(x[:, 0] < rng.uniform(-0.01, 0.01), 1, 2)
    c1 = np.where(x[:, 1] < rng.uniform(-0.01, 0.01), 1, 2)
    codes = c0 + 3 * c1
    # Uncut patches and single-digit particles scattered through the block.
    codes[rng.random(s.n) < 0.05] = 0
    single = rng.random(s.n) < 0.05
    codes[single] = c0[single]
    sizes = check(s, codes)
    assert sizes.sum() == s.n


def test_labels_after_substeps_with_a_cut():
    s = solver(cutting_energy=0.05)
    cfg = s.config
    s.tool_pos.assign([[0.0, 0.0, cfg.size[2] + cfg.blade_half_size[2] - 0.004]])
    for _ in range(5):
        s.step((0.0, 0.0, 0.0), cutting=True)
    codes = s.codes.numpy()
    labels, sizes = s.piece_labels()
    ref_labels, ref_sizes = components(s.node_numpy, s.weight_numpy, codes)
    np.testing.assert_array_equal(labels, ref_labels)
    np.testing.assert_array_equal(sizes, ref_sizes)
