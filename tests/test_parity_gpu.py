"""Parity of the HIP substep with the CPU oracle (float32 mirror) on the same seeded inputs.

Bar: cut codes and field keys bit-exact; particle state (x, v, F) and the blade/reaction
state within 1e-5 relative to each array's magnitude after one substep from identical
inputs; multi-step runs within looser bounds set by float-atomic order (the reference
itself differs run to run by ~1e-7 m, tests/test_multi_tool.py:42-44 of the reference).
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from blademaster_hip import Config, Solver  # noqa: E402
from oracle import OracleSim  # noqa: E402

TOL = 1.0e-5


def rel(a, b, floor=1e-30):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), floor)) if a.size else 0.0


def oracle_for(s, precision=32):
    o = OracleSim(s.config, s.reference, s.node_numpy, s.weight_numpy, s._gradient_numpy, s.node_count, precision,
                  s.max_cut_index)
    sync_to_oracle(s, o)
    return o


def sync_to_oracle(s, o):
    for name, src in [("x", s.x), ("v", s.v), ("F", s.F), ("Fp", s.Fp), ("stress", s.stress),
                      ("damage", s.damage), ("histories", s.histories), ("tool_pos", s.tool_pos),
                      ("tool_vel", s.tool_vel), ("tool_rot", s.tool_rot), ("tool_omega", s.tool_omega),
                      ("tool_mobility", s.tool_mobility), ("impulses", s.impulses)]:
        o.set(name, src.numpy())
    o.set("codes", s.codes.numpy())


def small_cfg(**kw):
    base = dict(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=0.0002, cache_topology=True)
    base.update(kw)
    return Config(**base)


def blade_touching(s, depth=0.001, x=0.0):
    cfg = s.config
    s.tool_pos.assign([[x, 0.0, cfg.size[2] + cfg.blade_half_size[2] - depth]])


@pytest.mark.parametrize("cut_hold", [False, True])
def test_one_substep_matches_oracle(cut_hold):
    cfg = small_cfg(cut_hold=cut_hold, cutting_energy=0.05)
    s = Solver(cfg)
    rng = np.random.default_rng(0)
    s.v.assign(rng.normal(0, 0.01, (s.n, 3)).astype(np.float32))
    blade_touching(s, 0.002)
    s.tool_vel.assign([[0.0, 0.0, -0.05]])
    o = oracle_for(s)
    target = (0.0, 0.0, 0.0)
    s.step(target, cutting=True)
    o.substep(target, cutting=True)
    assert np.array_equal(s.codes.numpy(), o.get("codes"))
    for name in ("x", "v", "F"):
        err = rel(getattr(s, name).numpy().ravel(), o.get(name))
        assert err < TOL, (name, err)
    np.testing.assert_allclose(s.tool_pos.numpy().ravel(), o.get("tool_pos"), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(s.damage.numpy(), o.get("damage"), atol=1e-6)
    assert abs(s.work.numpy()[0] - o.get("work")[0]) <= 1e-5 * max(abs(o.get("work")[0]), 1e-12)


def test_material_step_with_incompatible_codes_matches_oracle():
    cfg = small_cfg(fields_per_node=32)
    s = Solver(cfg)
    rng = np.random.default_rng(23)
    s.v.assign(rng.normal(0, 0.01, (s.n, 3)).astype(np.float32))
    s.F.assign(np.tile(np.diag([1.01, 0.995, 1.0]), (s.n, 1, 1)).astype(np.float32))
    s.refresh_stress()
    s.codes.assign(rng.integers(0, 27, s.n))
    s.impulses.assign(rng.normal(0, 1.0e-7, (s.n, 3)).astype(np.float32))
    o = oracle_for(s)
    s.material_step()
    o.stage("material")
    # Field keys: the same set, and every corner maps to the key code * N + node.
    keys = np.sort(s.grid_keys.numpy())
    np.testing.assert_array_equal(keys, np.sort(o.grid_keys()))
    codes = s.codes.numpy().astype(np.int64)
    np.testing.assert_array_equal(s.corner_keys(), codes[:, None] * s.node_count + s.node_numpy)
    for name in ("x", "v", "F"):
        err = rel(getattr(s, name).numpy().ravel(), o.get(name))
        assert err < TOL, (name, err)
    assert not s.errors.numpy().any()


def test_rigid_contact_impulses_match_oracle():
    cfg = small_cfg(free_blade=True)
    s = Solver(cfg)
    theta = 0.3
    s.tool_rot.assign([[0, np.sin(theta / 2), 0, np.cos(theta / 2)]])
    s.tool_pos.assign([[0.0, 0.0, 0.02]])
    rng = np.random.default_rng(5)
    s.v.assign(rng.normal(0, 0.02, (s.n, 3)).astype(np.float32))
    o = oracle_for(s)
    s.rigid_contacts(cutting=True)
    o.stage("rigid", 1)
    err = rel(s.impulses.numpy().ravel(), o.get("impulses"))
    assert err < TOL, err
    np.testing.assert_allclose(s.reaction.numpy().ravel(), o.get("reaction"), rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(s.histories.numpy().ravel(), o.get("histories"), rtol=1e-5, atol=1e-10)
    # Momentum balance of the contact (reference tests/test_mechanics.py:423-437).
    gravity = np.array([0.0, 0.0, -9.81 * s.mass * cfg.dt]) * s.n
    material = s.impulses.numpy().astype(np.float64).sum(0)
    np.testing.assert_allclose(material + s.reaction.numpy()[0] - s.board.numpy()[0] - gravity, 0, atol=3e-9)


def test_pair_contacts_across_a_cut_match_oracle():
    cfg = small_cfg(dt=0.00025)
    s = Solver(cfg)
    x = s.x.numpy()
    codes = np.where(s.reference[:, 0] < 0, 1, 2)
    # Push the two halves into each other so incompatible pairs touch.
    x[:, 0] += np.where(codes == 1, 0.0006, -0.0006)
    s.x.assign(x)
    s.codes.assign(codes)
    rng = np.random.default_rng(9)
    s.v.assign(rng.normal(0, 0.01, (s.n, 3)).astype(np.float32))
    o = oracle_for(s)
    s.rigid_contacts(cutting=False)
    s.particle_contacts()
    o.stage("rigid", 0)
    o.stage("pairs")
    count = int(s.pair_count[(s.step_count + 1) % 2].numpy()[0])
    assert count == o.scalars()["pair_count"] > 0
    err = rel(s.impulses.numpy().ravel(), o.get("impulses"))
    assert err < TOL, err


def test_assign_labels_are_bit_exact():
    s = Solver(small_cfg())
    rng = np.random.default_rng(1)
    s.v.assign(rng.normal(0, 0.01, (s.n, 3)).astype(np.float32))
    s.tool_pos.assign([[0.001, 0.0, 0.025]])
    s.tool_vel.assign([[0.0, 0.0, -0.2]])
    o = oracle_for(s)
    s.assign_sides(0, True)
    o.stage("assign", 1, 1)
    a = s.codes.numpy()
    assert set(np.unique(a)) == {0, 1, 2}
    np.testing.assert_array_equal(a, o.get("codes"))


@pytest.mark.parametrize("mode", ["guided", "free", "servo"])
def test_short_run_tracks_oracle(mode):
    kw = dict(guided={}, free=dict(free_blade=True), servo=dict(pose_servo=True))[mode]
    cfg = small_cfg(yield_stress=2000.0, **kw)
    s = Solver(cfg)
    blade_touching(s, 0.0005)
    o = oracle_for(s)
    target = (0.0, 0.0, 0.008)
    worst = {}
    for step in range(20):
        s.step(target, cutting=True)
        o.substep(target, cutting=True)
        np.testing.assert_array_equal(s.codes.numpy(), o.get("codes"))
        for name in ("x", "v", "F"):
            worst[name] = max(worst.get(name, 0.0), rel(getattr(s, name).numpy().ravel(), o.get(name)))
        # Per-substep error from identical inputs: resync and compare the next step alone.
        sync_to_oracle(s, o)
    assert worst["x"] < TOL and worst["F"] < TOL, worst
    assert worst["v"] < 1e-4, worst


def test_friction_histories_track_contact_on_and_off():
    """Friction history rows (board, blade) against the oracle while board contacts begin and
    blade contacts end (the knife is lifted away after four substeps): a particle without
    contact holds an exactly zero history, including one whose contact just ended (the contact
    kernel skips rewriting histories known to be zero)."""
    cfg = small_cfg(cutting_energy=0.05)
    s = Solver(cfg)
    blade_touching(s, 0.0015)
    s.v.assign(np.tile([0.0, 0.0, -0.02], (s.n, 1)).astype(np.float32))
    o = oracle_for(s)
    lifted = [0.0, 0.0, cfg.size[2] + cfg.blade_half_size[2] + 0.05]
    began = ended = 0
    before = None
    for step in range(8):
        if step == 4:
            s.tool_pos.assign([lifted])
            o.set("tool_pos", lifted)
        target = (0.0, 0.0, cfg.size[2] - 0.002) if step < 4 else (0.0, 0.0, cfg.size[2] + 0.05)
        s.step(target, cutting=False)
        o.substep(target, cutting=False)
        h, ho = s.histories.numpy().reshape(2, -1, 3), o.get("histories").reshape(2, -1, 3)
        zero = np.all(ho == 0.0, axis=2)
        assert np.all(h[zero] == 0.0), ("history left nonzero", step)
        for row in range(2):
            scale = max(np.abs(ho[row]).max(), 1e-30)
            assert np.abs(h[row] - ho[row]).max() <= 1e-4 * scale, (row, step)
        if before is not None:
            began += int((before & ~zero).sum())
            ended += int((~before & zero).sum())
        before = zero
        sync_to_oracle(s, o)
    assert began > 0 and ended > 0


def test_graph_replay_matches_eager():
    cfg = small_cfg(yield_stress=2000.0)
    eager, graph = Solver(cfg), Solver(cfg)
    for s in (eager, graph):
        blade_touching(s, 0.0005)
    for n, cut, on, xoff in [(5, 0, True, 0.0), (6, 0, True, 0.001), (5, 0, False, 0.0), (6, 1, True, -0.001)]:
        commands = np.column_stack([np.full(n, xoff), np.zeros(n), np.linspace(0.012, 0.011, n)])
        eager.advance(commands, on, cut, graph=False)
        graph.advance(commands, on, cut, graph=True)
        a, b = eager.state(), graph.state()
        assert eager.step_count == graph.step_count
        np.testing.assert_allclose(a["position"], b["position"], atol=2e-7)
        np.testing.assert_allclose(a["velocity"], b["velocity"], atol=2e-5)
        np.testing.assert_allclose(a["deformation"], b["deformation"], atol=2e-5)
        np.testing.assert_array_equal(a["sides"], b["sides"])


def test_relabelled_pair_lists_match_a_full_rebuild():
    """After labels change on the device (as assign does), the pair lists of the relabelled
    particles and their neighbours are rebuilt at the last build positions; the contacts of
    the next substep equal those of a full rebuild from the current positions."""
    cfg = small_cfg(dt=0.00025)

    def make():
        s = Solver(cfg)
        x = s.x.numpy()
        codes = np.where(s.reference[:, 0] < 0, 1, 2)
        x[:, 0] += np.where(codes == 1, 0.0006, -0.0006)
        s.x.assign(x)
        s.codes.assign(codes)
        s.v.assign(np.random.default_rng(4).normal(0, 0.01, (s.n, 3)).astype(np.float32))
        s.step(cutting=False)
        return s

    a, b, control = make(), make(), make()
    assert int(a.pair_count[a.step_count % 2].numpy()[0]) > 0
    # Relabel part of the right half to the left half's code.
    ref = a.reference
    moved = np.nonzero((ref[:, 0] > 0) & (ref[:, 1] > 0))[0].astype(np.int32)
    for s in (a, b):
        T = s._t
        T["codes"][torch.from_numpy(moved).long().to(T["codes"].device)] = 1
        T["topology_dirty"].fill_(1)
    a._t["changed"][: len(moved)] = torch.from_numpy(moved).to(a._t["changed"].device)
    a._t["changed_count"].fill_(len(moved))
    a._call("bm_mark_relabelled")  # what assign does after recording the relabels
    b._t["pair_dirty"].fill_(1)
    builds = int(a.pair_list_builds.numpy()[0])
    for s in (a, b, control):
        s.step(cutting=False)
    assert int(a.pair_list_builds.numpy()[0]) == builds  # no full rebuild in a
    count_a, count_b, unchanged = (int(s.pair_count[s.step_count % 2].numpy()[0]) for s in (a, b, control))
    # The relabelled particles lose their cross-cut pairs and gain pairs with their former
    # same-side neighbours: the contact set differs from the unrelabelled run.
    assert count_a == count_b and count_a > 0 and count_a != unchanged
    for name in ("x", "v"):
        assert rel(getattr(a, name).numpy().ravel(), getattr(b, name).numpy().ravel()) < 1e-6, name


def field_tables_hold_the_codes(s):
    """Every particle finds its transfer code in its slot at each of its 8 nodes, and each
    node's used slots hold exactly the distinct codes of the particles around it."""
    T = s._t
    codes = s.codes.numpy()
    fc = T["field_code"].cpu().numpy()
    nf = T["node_nf"].cpu().numpy()
    fi = T["field_idx"].cpu().numpy().astype(np.int64)
    dz, dyz = int(s.dims[2]), int(s.dims[1] * s.dims[2])
    offsets = np.array([((a >> 2) & 1) * dyz + ((a >> 1) & 1) * dz + (a & 1) for a in range(8)])
    nodes = T["base_node"].cpu().numpy().astype(np.int64)[:, None] + offsets[None, :]
    np.testing.assert_array_equal(fc[fi, nodes], np.repeat(codes[:, None], 8, 1))
    expected = np.unique(np.stack([nodes.ravel(), np.repeat(codes, 8)], 1), axis=0)
    touched = np.unique(nodes)
    slots = [(node, fc[f, node]) for node in touched for f in range(nf[node])]
    np.testing.assert_array_equal(np.array(sorted(slots)), expected)


def test_incremental_registration_matches_a_full_one():
    """Labels changed on the device (as assign does): registering again only around the
    relabelled particles gives the field tables of a full registration (material.py:50-65)
    and the same substep."""
    cfg = small_cfg(dt=0.00025)

    def make():
        s = Solver(cfg)
        s.codes.assign(np.where(s.reference[:, 0] < 0, 1, 2))
        s.v.assign(np.random.default_rng(5).normal(0, 0.01, (s.n, 3)).astype(np.float32))
        s.step(cutting=False)
        return s

    a, b = make(), make()
    ref = a.reference
    moved = np.nonzero((ref[:, 0] > 0) & (ref[:, 1] > 0))[0].astype(np.int32)
    for s in (a, b):
        T = s._t
        T["codes"][torch.from_numpy(moved).long().to(T["codes"].device)] = 1
        T["changed"][: len(moved)] = torch.from_numpy(moved).to(T["changed"].device)
        T["changed_count"].fill_(len(moved))
        s._call("bm_mark_relabelled")  # what assign does after recording the relabels
    a._t["topology_dirty"].fill_(2)  # only around the relabelled particles
    b._t["topology_dirty"].fill_(1)  # every particle
    for s in (a, b):
        s.step(cutting=False)
        field_tables_hold_the_codes(s)
        assert int(s._t["topology_dirty"].item()) == 0
    for name in ("x", "v"):
        assert rel(getattr(a, name).numpy().ravel(), getattr(b, name).numpy().ravel()) < 1e-6, name


def test_deferred_update_matches_eager_across_relabels():
    """A recorded interval leaves each substep's F update to the next substep's P2G, except
    for the particles around relabelled ones (updated before they register again) and the
    last substep: the run equals substep-by-substep stepping while the knife relabels."""
    cfg = small_cfg(yield_stress=2000.0, cutting_energy=0.05)
    eager, graph = Solver(cfg), Solver(cfg)
    for s in (eager, graph):
        blade_touching(s, 0.002)  # the edge below the top particle layer: it is relabelled
    edge = cfg.size[2] - 0.002
    commands = np.column_stack([np.zeros(16), np.zeros(16), np.linspace(edge, edge - 0.0005, 16)])
    before = eager.codes.numpy().copy()
    eager.advance(commands, True, 0, graph=False)
    graph.advance(commands, True, 0, graph=True)
    after = eager.codes.numpy()
    assert np.count_nonzero(after != before) > 0  # labels changed inside the interval
    np.testing.assert_array_equal(graph.codes.numpy(), after)
    for name in ("x", "v", "F"):
        assert rel(getattr(graph, name).numpy().ravel(), getattr(eager, name).numpy().ravel()) < 1e-6, name
    assert int(graph._t["update_pending"].item()) == 0 and not graph._t["update_done"].any()


def test_deferred_update_matches_eager_after_overflowing_relabels():
    """With a changed list of two entries the relabels of a substep overflow it, so the next
    substep re-registers every particle: inside a recorded interval the deferred update then
    updates every particle beside the node kernel, and the run still equals substep-by-substep
    stepping."""

    class Tiny(Solver):
        CHANGED_CAPACITY = 2

    cfg = small_cfg(yield_stress=2000.0, cutting_energy=0.05)
    eager, graph, probe = Tiny(cfg), Tiny(cfg), Tiny(cfg)
    for s in (eager, graph, probe):
        blade_touching(s, 0.002)
    edge = cfg.size[2] - 0.002
    commands = np.column_stack([np.zeros(16), np.zeros(16), np.linspace(edge, edge - 0.0005, 16)])
    # The same substeps one at a time: the relabel count of each (all of them are counted,
    # also those beyond the list) shows that some substep overflows the list.
    counts = []
    for row in commands:
        probe.advance(row[None, :], True, 0, graph=False)
        counts.append(int(probe._t["changed_count"].item()))
    assert max(counts) > Tiny.CHANGED_CAPACITY, counts
    before = eager.codes.numpy().copy()
    eager.advance(commands, True, 0, graph=False)
    graph.advance(commands, True, 0, graph=True)
    after = eager.codes.numpy()
    assert np.count_nonzero(after != before) > 0
    np.testing.assert_array_equal(graph.codes.numpy(), after)
    for name in ("x", "v", "F"):
        assert rel(getattr(graph, name).numpy().ravel(), getattr(eager, name).numpy().ravel()) < 1e-6, name
    assert int(graph._t["update_pending"].item()) == 0 and not graph._t["update_done"].any()
