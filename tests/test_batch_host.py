"""Host logic of the batched instances: packing keeps each instance's lattice on the shared
grid and keeps instances a knife apart; mismatched shared parameters are rejected."""

import numpy as np
import pytest

from blademaster_hip import Config
from blademaster_hip.batch import BatchSolver, pack
from blademaster_hip.geometry import points


def configs():
    base = dict(grid_spacing=0.006, dt=0.0002, external_tool=True, tool_substep="predict",
                blade_half_size=(0.0005, 0.02, 0.015))
    return [Config(size=(0.048, 0.024, 0.018), young_modulus=1e5, **base),
            Config(size=(0.03, 0.036, 0.024), young_modulus=3e5, poisson_ratio=0.35, sampling_seed=2, **base),
            Config(size=(0.06, 0.018, 0.012), young_modulus=2e4, shape="block", **base)]


def test_pack_keeps_lattices_and_gaps():
    cfgs = configs()
    layout, offsets, slices = pack(cfgs)
    h = cfgs[0].grid_spacing
    ref = layout["reference"]
    assert len(ref) == sum(s.stop - s.start for s in slices) == len(layout["instance"])
    boxes = []
    for i, cfg in enumerate(cfgs):
        own, own_origin, own_cells = points(cfg.size, h, cfg.shape, cfg.sampling_seed)
        mine = ref[slices[i]]
        np.testing.assert_allclose(mine - offsets[i], own, rtol=0, atol=1e-12)
        # same cell-local fractions on the packed grid as on the instance's own grid
        f_packed = (mine - layout["origin"]) / h
        f_own = (own - own_origin) / h
        np.testing.assert_allclose(f_packed - np.floor(f_packed), f_own - np.floor(f_own), atol=1e-9)
        assert np.all(layout["instance"][slices[i]] == i)
        lo = np.floor(f_packed.min(0)).astype(int)
        hi = np.floor(f_packed.max(0)).astype(int)
        assert np.all(lo >= 0) and np.all(hi < layout["cells"])
        boxes.append((mine.min(0), mine.max(0)))
    gap = 2 * np.linalg.norm(cfgs[0].blade_half_size)
    for i in range(len(boxes)):
        for j in range(i + 1, len(boxes)):
            (a0, a1), (b0, b1) = boxes[i], boxes[j]
            sep = np.maximum(np.maximum(b0 - a1, a0 - b1), 0.0)[:2]
            assert sep.max() >= gap


def test_batch_rejects_mismatched_shared_parameters():
    cfgs = configs()
    cfgs[1] = Config(size=(0.03, 0.036, 0.024), grid_spacing=0.004, dt=0.0002, external_tool=True,
                     tool_substep="predict", blade_half_size=(0.0005, 0.02, 0.015))
    with pytest.raises(ValueError, match="grid_spacing"):
        BatchSolver(cfgs)
    with pytest.raises(ValueError, match="predict"):
        BatchSolver([Config(size=(0.03, 0.03, 0.03))])
