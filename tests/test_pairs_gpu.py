"""Pair search modes (solver.py:260-263): every pair tested ("all"), partner lists rebuilt from
the cell buckets each step ("hash") and Verlet partner lists ("list", "auto") find the same
contacts.  The two-particle cases restate tests/test_pair_list.py:11-46 of the reference:
exactly one pair, equal impulses; a short cut then runs the same under every mode."""

from dataclasses import replace

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver  # noqa: E402

MODES = ("all", "hash", "list", "auto")


@pytest.mark.parametrize("dt,speed,skin", [(0.001, 1.0, 0.0), (0.002, 1.0, 0.0), (0.0000625, 16.0, 0.0),
                                           (0.001, 1.0, 0.0001)])
@pytest.mark.parametrize("graph", [False, True])
def test_search_modes_find_the_same_single_pair(dt, speed, skin, graph):
    cfg = Config(size=(0.018, 0.018, 0.018), grid_spacing=0.006, dt=dt, grid_storage="compact", pair_skin=skin)
    results = []
    for mode in MODES:
        s = Solver(replace(cfg, pair_search=mode))
        x = np.zeros((s.n, 3), np.float32)
        x[:, 0] = np.arange(s.n) * 0.1 + 2
        x[:, 2] = 0.1
        x[:2] = [[0, 0, 0.1], [0.0048, 0, 0.1]]
        v = np.zeros_like(x)
        v[:2] = [[speed, 0, 0], [-speed, 0, 0]]
        codes = np.zeros(s.n, np.int64)
        codes[:2] = [1, 2]
        s.x.assign(x)
        s.v.assign(v)
        s.codes.assign(codes)
        s.speed.fill_(speed)
        if graph:
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(g):
                s.particle_contacts()
            g.replay()
        else:
            s.particle_contacts()
        assert s.pair_count[1].numpy()[0] == 1, mode
        results.append(s.impulses.numpy())
    for result in results[1:]:
        np.testing.assert_allclose(result, results[0], rtol=1e-6, atol=1e-12)
    assert np.abs(results[0][:2]).max() > 0


def test_search_modes_agree_over_a_cut():
    cfg = Config(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=0.0002, cache_topology=True,
                 cutting_energy=0.05)
    finals = {}
    for mode in MODES:
        s = Solver(replace(cfg, pair_search=mode))
        s.tool_pos.assign([[0.0, 0.0, cfg.size[2] + cfg.blade_half_size[2] - 0.002]])
        s.advance(np.tile([0.0, 0.0, 0.004], (30, 1)), cutting=True)
        finals[mode] = (s.x.numpy(), s.codes.numpy(), int(s.pair_count[s.step_count % 2].numpy()[0]))
    x0, c0, n0 = finals["all"]
    assert (c0 != 0).sum() > 0
    for mode in MODES[1:]:
        x, c, n = finals[mode]
        np.testing.assert_array_equal(c, c0)
        assert n == n0, (mode, n, n0)
        assert np.abs(x - x0).max() <= 1e-5 * np.abs(x0).max(), mode


def test_all_pairs_refuses_large_solvers():
    with pytest.raises(ValueError, match="pair_search='all'"):
        Solver(Config(size=(0.3, 0.15, 0.075), grid_spacing=0.003, pair_search="all"))
