"""Held-blade grid constraint (blade_constraint.py) on the GPU against the oracle.

Bar: one held substep from identical inputs — particle state within 1e-5 relative, the
interval impulse within 1e-5 of the block's momentum scale m n |v| (the two faces' impulses
cancel), face sides identical; the unresolved inward residual stays below the
reference's 1e-6 m/s acceptance (coupling.py:187-188).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver, ExternalToolPort, ToolState  # noqa: E402
from oracle import OracleSim  # noqa: E402


def face_setup(s, squeeze):
    """Columns next to the blade plane sit 50 um off the faces and move into them."""
    hx = s.config.blade_half_size[0]
    x = s.reference.astype(np.float32).copy()
    near = np.abs(x[:, 0]) < 0.002
    x[near, 0] = np.sign(x[near, 0]) * (hx + 5.0e-5)
    v = np.zeros((s.n, 3), np.float32)
    v[:, 0] = -np.sign(s.reference[:, 0]) * squeeze
    s.x.assign(x)
    s.v.assign(v)
    return x, v


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("squeeze", [0.05, 0.2])
def test_held_blade_substep_matches_oracle(squeeze):
    cfg = Config(size=(0.036, 0.018, 0.018), grid_spacing=0.006, dt=0.0005, external_tool=True,
                 cache_topology=True, blade_half_size=(0.0005, 0.025, 0.02))
    s = Solver(cfg)
    port = ExternalToolPort(s)
    x, v = face_setup(s, squeeze)
    hz = cfg.blade_half_size[2]
    state = ToolState([0.0, 0.0, cfg.size[2] + hz - 0.009], [0, 0, 0, 1], [0, 0, 0], [0, 0, 0])
    o = OracleSim(cfg, s.reference, s.node_numpy, s.weight_numpy, s._gradient_numpy, s.node_count, 32,
                  s.max_cut_index)
    o.set("x", x)
    o.set("v", v)
    o.set("tool_pos", state.position)
    o.set("tool_rot", state.orientation_xyzw)
    r = port.advance(state, cfg.dt, cutting=True)
    o.substep(cutting=True)
    np.testing.assert_array_equal(s.codes.numpy(), o.get("codes"))
    for name in ("x", "v", "F"):
        assert rel(getattr(s, name).numpy(), o.get(name)) < 1e-5, name
    np.testing.assert_array_equal(s.blade_constraint.side.numpy(), o.get("bc_side"))
    active = np.abs(o.get("bc_side")) > 0
    assert active.sum() > 0
    # The two faces' impulses cancel in x: compare on the momentum scale of the block.
    scale = s.mass * s.n * max(squeeze, 1e-3)
    np.testing.assert_allclose(r.impulse_ns, o.get("interval_impulse"), rtol=0, atol=1e-5 * scale)
    assert s.blade_constraint.max_violation.numpy()[0] <= 1e-6
    assert o.get("bc_max_violation")[0] <= 1e-6


def test_held_blade_stops_inward_grid_motion():
    """After the constraint the advecting grid velocity of face particles does not enter the
    blade (blade_constraint.py docstring; reference tests/test_blade_constraint.py:57-101)."""
    cfg = Config(size=(0.036, 0.018, 0.018), grid_spacing=0.006, dt=0.0005, external_tool=True,
                 cache_topology=True, blade_half_size=(0.0005, 0.025, 0.02))
    s = Solver(cfg)
    face_setup(s, 0.2)
    hz = cfg.blade_half_size[2]
    s.tool_pos.assign([[0.0, 0.0, cfg.size[2] + hz - 0.009]])
    s.rigid_contacts(cutting=False)
    s._call("bm_register")
    s._call("bm_p2g")
    s._call("bm_couple_and_velocity")
    before = s.new_v.numpy().copy()
    mass = s.grid_mass.numpy().astype(np.float64)
    s.reaction.zero_()
    s.torque.zero_()
    s.blade_constraint.solve()
    after = s.new_v.numpy().astype(np.float64)
    side = s.blade_constraint.side.numpy()
    active = np.flatnonzero(side != 0)
    assert len(active) > 0
    slots = s.slots.numpy()
    w = s.weight_numpy
    flat = after.reshape(-1, 3)
    advect = np.sum(w[active, :, None] * flat[slots[active]], axis=1)
    gap = np.abs(s.x.numpy()[active, 0]) - cfg.blade_half_size[0]
    # velocity into the face is at most what closes the present gap
    assert np.all(side[active] * advect[:, 0] >= -np.maximum(gap, 0) / cfg.dt - 2e-5)
    material_impulse = np.sum(mass[..., None] * (after - before), axis=(0, 1))
    np.testing.assert_allclose(material_impulse + s.reaction.numpy()[0], 0, atol=2e-9)
    energy_before = 0.5 * np.sum(mass[..., None] * before.astype(np.float64) ** 2)
    energy_after = 0.5 * np.sum(mass[..., None] * after ** 2)
    assert energy_after <= energy_before * (1 + 1e-6)
