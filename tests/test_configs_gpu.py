"""The workloads of the measurement plan on the GPU, each against the oracle or against the
single-instance solver:

- the 8960-particle gravity-knife drop (56 x 20 x 8 particles on a 3 mm lattice, 6 mm grid,
  20 g knife released at rest 2 mm above, 62.5 us substeps): substeps against the float32
  oracle from identical inputs, and the 0.5 s drop through the run driver;
- the batched sweep of 256 uneven instances: one coupling interval with per-instance checks,
  three instances against their own solvers and one against the oracle;
- the MLS-MPM baseline on the 1M-particle block against its oracle.

Bars: codes bit-exact; x, v, F within 1e-5 relative after each substep from identical inputs
(a batched instance: x within 1e-5 of its extent plus the rounding of its packed
coordinates, v within 1e-3); over whole intervals, knife impulse and velocities within 1 %.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, ExternalToolPort, Solver, ToolState, freefall  # noqa: E402
from blademaster_hip.batch import BatchPort, BatchSolver, sweep_instances  # noqa: E402
from blademaster_hip.mlsmpm import MlsSolver, TopoCutParams  # noqa: E402
from oracle import OracleMls, OracleSim  # noqa: E402


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


SYNC = ("x", "v", "F", "Fp", "stress", "damage", "histories", "tool_pos", "tool_vel", "tool_rot", "tool_omega")


def test_gravity_knife_drop_substeps_match_oracle():
    scene = freefall.SCENES["gravity-knife"]
    cfg = scene.config()
    s = Solver(cfg)
    assert s.n == 8960 and s.node_count == 29 * 11 * 5
    q, height = freefall.release_pose(scene, cfg.size[2])
    s.tool_rot.assign([q])
    s.tool_pos.assign([[0.0, 0.0, height]])
    o = OracleSim(cfg, s.reference, s.node_numpy, s.weight_numpy, s._gradient_numpy, s.node_count, 32,
                  s.max_cut_index)
    worst = {}
    for step in range(8):
        if step == 4:
            # Past the free fall: the knife 1 mm into the block, moving down at 0.3 m/s.
            s.tool_pos.assign([[0.0, 0.0, height - scene.clearance - 0.001]])
            s.tool_vel.assign([[0.0, 0.0, -0.3]])
        for name in SYNC:
            o.set(name, getattr(s, name).numpy())
        o.set("codes", s.codes.numpy())
        s.step((0.0, 0.0, 0.0), cutting=True)
        o.substep(cutting=True)
        np.testing.assert_array_equal(s.codes.numpy(), o.get("codes"))
        for name in ("x", "v", "F", "tool_pos", "tool_vel"):
            worst[name] = max(worst.get(name, 0.0), rel(getattr(s, name).numpy(), o.get(name)))
    assert all(v < 1e-5 for v in worst.values()), worst


def test_gravity_knife_drop_runs_half_a_second(tmp_path):
    """The whole drop (8000 substeps): the knife falls 2 mm, lands on the block and rests on
    it; the material stays valid and in one piece."""
    result = freefall.run(tmp_path / "drop", "gravity-knife")
    scene = freefall.SCENES["gravity-knife"]
    assert result["status"] == "complete" and result["gpu_errors"] == [0, 0, 0]
    assert result["particle_count"] == 8960
    top = scene.size[2]
    assert top - 0.006 < result["final_lowest_edge_m"] < top + 0.001, result["final_lowest_edge_m"]
    assert result["minimum_det_f_sampled"] > 0.5
    assert sum(result["component_sizes"]) == 8960
    from blademaster_hip.runio import load_states

    z = load_states(tmp_path / "drop" / "states.npz")
    assert len(z["time"]) == 1 + 50
    speed = np.linalg.norm(z["blade_velocity"][-1][0])
    assert speed < 0.05, speed  # at rest on the block
    # It fell freely first: 10 ms in, v = -g t.
    np.testing.assert_allclose(z["blade_velocity"][1][0][2], -9.81 * 0.01, rtol=2e-3)


def _into_block(knives):
    """Start times at which every knife's edge is 1 mm into its block (they start 2 mm above)."""
    return [0.003 / -kn(0.0).linear_velocity[2] for kn in knives]


def test_sweep_of_256_instances_follows_single_solvers():
    configs, knives, gap = sweep_instances(256, seed=0)
    batch = BatchSolver(configs, gap=gap)
    port = BatchPort(batch)
    dt = configs[0].dt
    t0 = _into_block(knives)
    interval = 40 * dt
    sample = [0, 97, 255]
    singles = {i: Solver(configs[i]) for i in sample}
    ports = {i: ExternalToolPort(singles[i]) for i in sample}
    for k in range(2):
        states = [kn(t0[i] + k * interval) for i, kn in enumerate(knives)]
        reactions = port.advance(states, interval, cutting=True)
        assert len(reactions) == 256
        assert all(np.isfinite(r.impulse_ns).all() and np.isfinite(r.angular_impulse_nms).all() for r in reactions)
        for i in sample:
            r = ports[i].advance(states[i], interval, cutting=True)
            scale = max(np.abs(r.impulse_ns).max(), 1e-12)
            assert np.abs(reactions[i].impulse_ns - r.impulse_ns).max() <= 0.01 * scale + 1e-12, (k, i)
    x_all = batch.x.numpy()
    assert np.isfinite(x_all).all() and np.isfinite(batch.v.numpy()).all()
    assert batch.errors.numpy().tolist() == [0, 0, 0]
    cut = [int((batch.instance_state(i, "codes") != 0).sum()) for i in range(256)]
    assert sum(c > 0 for c in cut) >= 32  # knives are labelling particles in many instances
    # Packed instances sit up to metres from the packed origin: each substep rounds their
    # float32 positions to the spacing of those coordinates.
    ulp = float(np.spacing(np.float32(np.abs(x_all).max())))
    for i in sample:
        s = singles[i]
        np.testing.assert_array_equal(batch.instance_state(i, "codes"), s.codes.numpy())
        x = s.x.numpy()
        assert np.abs(batch.instance_state(i, "x") - x).max() <= 1e-5 * np.abs(x).max() + 80 * ulp, i
        # Over 80 substeps of knife contact the rounding of packed positions (1e-7 m) moves
        # penalty impulses by up to about 1 %: velocities are held to the impulse bar.
        v = s.v.numpy()
        assert np.abs(batch.instance_state(i, "v") - v).max() <= 1e-2 * max(np.abs(v).max(), 1e-6), i


def test_sweep_instance_substeps_match_oracle():
    """Instance 0 of the 256-instance sweep, one substep at a time from identical inputs,
    against the oracle run of its own configuration."""
    configs, knives, gap = sweep_instances(256, seed=0)
    batch = BatchSolver(configs, gap=gap)
    port = BatchPort(batch)
    dt = configs[0].dt
    t0 = _into_block(knives)
    i = 0
    ref = batch.instance_state(i, "ref")
    single = Solver(configs[i])  # the instance's stencil and constants, as the oracle needs them
    assert np.abs(single.reference - ref).max() < 1e-6
    o = OracleSim(configs[i], single.reference, single.node_numpy, single.weight_numpy, single._gradient_numpy,
                  single.node_count, 32, single.max_cut_index)
    sl = batch.slices[i]
    for step in range(4):
        states = [kn(t0[j] + step * dt) for j, kn in enumerate(knives)]
        st = states[i]
        o.set("x", batch.instance_state(i, "x"))
        for name in ("v", "F", "Fp", "stress", "damage", "codes"):
            o.set(name, batch.instance_state(i, name))
        o.set("histories", batch.histories.numpy()[:, sl])
        o.set("tool_pos", st.position)
        o.set("tool_rot", st.orientation_xyzw)
        o.set("tool_vel", st.linear_velocity)
        o.set("tool_omega", st.angular_velocity)
        o.set("tool_mobility", np.zeros(36))
        port.advance(states, dt, cutting=True)
        o.substep(cutting=True)
        np.testing.assert_array_equal(batch.instance_state(i, "codes"), o.get("codes"))
        x, xo = batch.instance_state(i, "x"), o.get("x").reshape(-1, 3)
        ulp = float(np.spacing(np.float32(np.abs(batch.x.numpy()).max())))
        assert np.abs(x - xo).max() <= 1e-5 * np.abs(xo).max() + 2 * ulp, step
        v, vo = batch.instance_state(i, "v"), o.get("v").reshape(-1, 3)
        assert np.abs(v - vo).max() <= 1e-3 * max(np.abs(vo).max(), 1e-6), step
        assert rel(batch.instance_state(i, "F"), o.get("F")) < 1e-5, step


def test_mls_block_of_one_million_particles_matches_oracle():
    rng = np.random.default_rng(0)
    E = float(10 ** rng.uniform(4, 6))
    cfg = Config(size=(0.3, 0.15, 0.075), grid_spacing=0.003, dt=25e-6, young_modulus=E, poisson_ratio=0.3,
                 external_tool=True, tool_substep="predict", blade_half_size=(0.0005, 0.08, 0.03))
    params = TopoCutParams()
    s = MlsSolver(cfg, params)
    assert s.n == 1_000_000
    rng = np.random.default_rng(4)
    s.v.assign(rng.normal(0, 0.01, (s.n, 3)).astype(np.float32))
    s.F.assign((np.eye(3) + 0.003 * rng.normal(size=(s.n, 3, 3))).astype(np.float32))
    tilt = np.deg2rad(10.0)
    q = [np.sin(tilt / 2), 0.0, 0.0, np.cos(tilt / 2)]
    hy, hz = cfg.blade_half_size[1], cfg.blade_half_size[2]
    z = cfg.size[2] - 0.003 + hy * np.sin(tilt) + hz * np.cos(tilt)  # edge 3 mm into the block
    state = ToolState([0.0, 0.0, z], q, [0, 0, -0.1], [0, 0, 0])
    o = OracleMls(cfg, params, s.reference, s.origin, s.dims, 32)
    for step in range(2):
        for name in ("x", "v", "F", "C", "damage", "yield_stress", "damage_reference"):
            o.set(name, getattr(s, name).numpy())
        o.set("tool_pos", state.position)
        o.set("tool_vel", state.linear_velocity)
        o.set("tool_rot", state.orientation_xyzw)
        o.set("tool_omega", state.angular_velocity)
        o.set("histories", s._t["histories"].cpu().numpy())
        o.set("impulse", np.zeros(3))
        o.set("angular_impulse", np.zeros(3))
        r = s.advance(state, cfg.dt, graph=False)
        o.substep()
        for name in ("x", "v", "F"):
            assert rel(getattr(s, name).numpy(), o.get(name)) < 1e-5, (step, name)
        assert rel(s.C.numpy(), o.get("C")) < 1e-4, step
        # Damage flags are equal except where J sits within float rounding of a threshold.
        J = np.linalg.det(o.get("F").reshape(-1, 3, 3))
        edge = np.minimum(np.abs(J - s.compressed), np.abs(J - s.stretched)) < 1e-5
        differ = (s.damage.numpy() > 0) != (o.get("damage") > 0)
        assert not np.any(differ & ~edge), (step, int(differ.sum()))
        assert (o.get("damage") > 0).sum() > 1000
        impulse = o.get("impulse")
        assert np.abs(impulse).max() > 0  # the knife is in contact
        np.testing.assert_allclose(r.impulse_ns, impulse, rtol=0, atol=1e-4 * np.abs(impulse).max())
