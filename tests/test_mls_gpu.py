"""MLS-MPM baseline on the GPU against the oracle (mlsmpm.py:293-393).

Bar: one substep from identical inputs within 1e-5 relative (x, v, F, C); damage flags equal;
behaviour of the reference tests/test_mlsmpm.py (a block settles on the board without
leaving the grid; a diverged particle is reported).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from blademaster_hip import Config, ToolState  # noqa: E402
from blademaster_hip.mlsmpm import MlsSolver, TopoCutParams  # noqa: E402
from oracle import OracleMls  # noqa: E402


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def cfg(**kw):
    base = dict(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=0.0001, yield_stress=2000.0)
    base.update(kw)
    return Config(**base)


@pytest.mark.parametrize("plastic", [False, True])
def test_mls_substeps_match_oracle(plastic):
    c = cfg(yield_stress=300.0 if plastic else 0.0)
    params = TopoCutParams(softening=1e4, damaged_shear=0.5)
    s = MlsSolver(c, params)
    rng = np.random.default_rng(3)
    s.v.assign(rng.normal(0, 0.05, (s.n, 3)).astype(np.float32))
    F = np.eye(3) + 0.01 * rng.normal(size=(s.n, 3, 3))
    s.F.assign(F.astype(np.float32))
    o = OracleMls(c, params, s.reference, s.origin, s.dims, 32)
    state = ToolState([0.0, 0.0, c.size[2] + c.blade_half_size[2] - 0.002], [0, 0, 0, 1], [0, 0, -0.05], [0, 0, 0])
    for step in range(5):
        for name in ("x", "v", "F", "C", "damage", "yield_stress", "damage_reference"):
            o.set(name, getattr(s, name).numpy())
        o.set("tool_pos", state.position)
        o.set("tool_vel", state.linear_velocity)
        o.set("tool_rot", state.orientation_xyzw)
        o.set("tool_omega", state.angular_velocity)
        o.set("histories", s._t["histories"].cpu().numpy())
        o.set("impulse", np.zeros(3))
        o.set("angular_impulse", np.zeros(3))
        r = s.advance(state, c.dt, graph=step % 2 == 1)
        o.substep()
        for name in ("x", "v", "F"):
            assert rel(getattr(s, name).numpy(), o.get(name)) < 1e-5, (step, name)
        assert rel(s.C.numpy(), o.get("C")) < 1e-4
        np.testing.assert_array_equal(s.damage.numpy() > 0, o.get("damage") > 0)
        np.testing.assert_allclose(r.impulse_ns, o.get("impulse"), rtol=1e-4, atol=1e-12)


def test_mls_block_settles_and_reports_divergence():
    c = cfg(dt=0.0002)
    s = MlsSolver(c, damage_from_start=False)
    far = ToolState([0.0, 0.0, 0.3], [0, 0, 0, 1], [0, 0, 0], [0, 0, 0])
    for _ in range(10):
        s.advance(far, 20 * c.dt)
    st = s.state()
    assert st["position"][:, 2].min() > 0.0
    assert np.abs(st["velocity"]).max() < 0.5
    s.start_damage()
    x = s.x.numpy()
    x[0] = [np.nan, 0, 0]
    s.x.assign(x)
    with pytest.raises(RuntimeError, match="diverged"):
        s.advance(far, c.dt)
