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


@pytest.mark.parametrize("capacity", [256, 2])
def test_refreshed_lists_find_the_contacts_of_full_rebuilds(capacity):
    """A particle that moves more than half the skin from its build position refreshes only its
    own partner list (one solver, list mode, at least 2^17 buckets).  The contacts and
    the motion equal those of a run that rebuilds every list each time instead, also when the
    movers or displaced particles overflow a small capacity (a full rebuild then)."""
    cfg = Config(size=(0.096, 0.048, 0.048), grid_spacing=0.003, dt=0.0001, cache_topology=True, pair_search="list")

    class Refresh(Solver):
        MOVER_CAPACITY = capacity

    class Full(Solver):
        PAIR_REFRESH = False

    runs = []
    for cls in (Refresh, Full):
        s = cls(cfg)
        assert s.n >= 65536
        ref = s.reference
        codes = np.where(ref[:, 0] < 0, 1, 2).astype(np.int64)
        x = s.x.numpy()
        x[:, 0] += np.where(codes == 1, 0.0004, -0.0004)  # the halves pressed together across the cut
        rng = np.random.default_rng(7)
        v = rng.normal(0, 0.01, (s.n, 3)).astype(np.float32)
        # A few particles at the cut flung at 2 m/s: they leave their build positions within a
        # few steps, again and again.
        fast = np.nonzero((np.abs(ref[:, 0]) < 0.001) & (np.abs(ref[:, 1]) < 0.003))[0][:6]
        v[fast] = rng.normal(0, 1.2, (len(fast), 3)).astype(np.float32)
        s.x.assign(x)
        s.codes.assign(codes)
        s.v.assign(v)
        counts, refreshed = [], []
        for _ in range(40):
            s.step(cutting=False)
            counts.append(int(s.pair_count[s.step_count % 2].numpy()[0]))
            refreshed.append(int(s._t["refresh_state"][2].item()))
        runs.append((s, counts, refreshed))
    (a, count_a, refreshed_a), (b, count_b, _) = runs
    assert a.mover_capacity > 0 and b.mover_capacity == 0
    assert any(refreshed_a), "no refresh happened"
    assert int(a.pair_list_builds.numpy()[0]) < int(b.pair_list_builds.numpy()[0])
    assert max(count_a) > 0
    assert count_a == count_b
    for name in ("x", "v"):
        xa, xb = getattr(a, name).numpy(), getattr(b, name).numpy()
        assert np.abs(xa - xb).max() <= 1e-5 * np.abs(xb).max(), name


def test_refreshes_beside_relabels_match_full_rebuilds():
    """Relabels (recorded as assign records them) and list refreshes in the same substeps: the
    relabel walks leave the movers' pairs to the movers' refreshes, which run beside them.  The
    contacts and the motion equal those of a run that rebuilds every list instead of refreshing."""
    cfg = Config(size=(0.096, 0.048, 0.048), grid_spacing=0.003, dt=0.0001, cache_topology=True, pair_search="list")

    class Full(Solver):
        PAIR_REFRESH = False

    runs = []
    for cls in (Solver, Full):
        s = cls(cfg)
        ref = s.reference
        codes = np.where(ref[:, 0] < 0, 1, 2).astype(np.int64)
        x = s.x.numpy()
        x[:, 0] += np.where(codes == 1, 0.0004, -0.0004)
        rng = np.random.default_rng(11)
        v = rng.normal(0, 0.01, (s.n, 3)).astype(np.float32)
        fast = np.nonzero((np.abs(ref[:, 0]) < 0.001) & (np.abs(ref[:, 1]) < 0.003))[0][:6]
        v[fast] = rng.normal(0, 1.2, (len(fast), 3)).astype(np.float32)
        s.x.assign(x)
        s.codes.assign(codes)
        s.v.assign(v)
        T = s._t
        dev = T["codes"].device
        counts = []
        for step in range(36):
            if step in (6, 14, 22, 30):
                # Relabel a strip of the right half next to the cut (and next to the fast
                # particles) to the left half's code, as assign does in G2P.
                lo = 0.0005 + 0.0015 * (step // 8)
                ids = np.nonzero((ref[:, 0] > lo) & (ref[:, 0] < lo + 0.0016) & (np.abs(ref[:, 1]) < 0.006))[0]
                ids = ids.astype(np.int32)
                T["codes"][torch.from_numpy(ids).long().to(dev)] = 1
                T["changed"][: len(ids)] = torch.from_numpy(ids).to(dev)
                T["changed_count"].fill_(len(ids))
                T["topology_dirty"].fill_(2)
            s.step(cutting=False)
            counts.append(int(s.pair_count[s.step_count % 2].numpy()[0]))
        runs.append((s, counts))
    (a, count_a), (b, count_b) = runs
    assert a.mover_capacity > 0 and int(a._t["refresh_state"][2].item()) > 0, "no refresh happened"
    assert max(count_a) > 0
    assert count_a == count_b
    for name in ("x", "v"):
        xa, xb = getattr(a, name).numpy(), getattr(b, name).numpy()
        assert np.abs(xa - xb).max() <= 1e-5 * np.abs(xb).max(), name
