"""Coupling ports on one solver: each port replays only graphs captured for its own buffers."""

import gc

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from blademaster_hip import Config, ExternalToolPort, Solver, ToolState  # noqa: E402
from blademaster_hip.batch import BatchPort, BatchSolver  # noqa: E402


def _cfg(**kw):
    base = dict(size=(0.036, 0.018, 0.012), grid_spacing=0.006, dt=0.0002, external_tool=True,
                tool_substep="predict", cache_topology=True)
    base.update(kw)
    return Config(**base)


def _owners(solver):
    return {k[-1][:2] for k in solver._batches if isinstance(k[-1], tuple)}


def test_ports_in_sequence_capture_their_own_graphs():
    cfg = _cfg()
    s = Solver(cfg)
    snap = s.snapshot()
    state = ToolState([0.0, 0.0, cfg.size[2] + cfg.blade_half_size[2] - 0.002], [0, 0, 0, 1], [0, 0, -0.05],
                      [0, 0, 0])
    first = ExternalToolPort(s)
    first.advance(state, 4 * cfg.dt)
    owner = first._owner
    assert owner in _owners(s)
    del first
    gc.collect()
    # The collected port's graphs (raw pointers into its freed buffers) left the cache.
    assert owner not in _owners(s)
    s.restore(snap)
    second = ExternalToolPort(s)
    assert second._owner != owner
    r_graph = second.advance(state, 4 * cfg.dt)
    assert second._owner in _owners(s)
    s.restore(snap)
    r_eager = ExternalToolPort(s).advance(state, 4 * cfg.dt, graph=False)
    scale = max(np.abs(r_eager.impulse_ns).max(), 1e-12)
    assert np.abs(r_graph.impulse_ns - r_eager.impulse_ns).max() <= 1e-4 * scale
    assert np.abs(r_graph.impulse_ns).max() > 0


def test_batch_ports_release_their_graphs():
    cfgs = [_cfg(size=(0.024, 0.018, 0.012)), _cfg(size=(0.030, 0.012, 0.012), young_modulus=2e5)]
    s = BatchSolver(cfgs)
    states = [ToolState([0.0, 0.0, c.size[2] + c.blade_half_size[2] - 0.001], [0, 0, 0, 1], [0, 0, -0.05],
                        [0, 0, 0]) for c in cfgs]
    port = BatchPort(s)
    port.advance(states, 2 * cfgs[0].dt)
    owner = port._owner
    assert owner in _owners(s)
    del port
    gc.collect()
    assert owner not in _owners(s)
    again = BatchPort(s)
    reactions = again.advance(states, 2 * cfgs[0].dt)
    assert len(reactions) == 2 and all(np.isfinite(r.impulse_ns).all() for r in reactions)
