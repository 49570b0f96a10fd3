"""Pins of the CPU oracle (oracle/tlmpm_oracle.cpp) before it is trusted as the checker.

Each oracle function is checked against the reference's own known answers and golden
vectors, or against a property the mathematics fixes, chosen so that a dropped term, a
wrong sign or index fails:

  blade_query      exact bevel-face results (reference tests/test_mechanics.py:151-180);
                   flat faces against a brute-force closest point on a dense surface sample
  friction         slip magnitude mu_d * j_n (test_mechanics.py:148); stick case closed form
  resist_cut       damage [1,1,1,.5,.54,1], work 4e-5, returned impulses
                   (tests/test_cut_resistance.py:45-67)
  couple (Eq. 8)   momentum conservation; all-compatible fields equalise to m_a S / M;
                   incompatible fields untouched (test_mechanics.py:80-105 fixture)
  g2p              known answers of tests/test_particle_transport.py
  svd3             reconstruction, orthogonality, singular values of numpy.linalg.svd
  stress           P = 2 mu (F - R) + lam (J - 1) J F^-T with R from scipy.linalg.polar
  von Mises        det Fp = 1, deviatoric norm on the yield surface (test_mechanics.py:303-314)
  rigid contact    material + reaction - board - gravity = 0 (test_mechanics.py:423-437)
  effective mass   articulated / prescribed = 1 / (1 + m u^T M u) (tests/test_coupling.py:234-257)
  pair contact     equal-opposite impulses, compatible neighbours excluded, exactly one
                   pair for the reference's two-particle fixture (tests/test_pair_list.py:11-46)
  assign           codes in {0, 1, 2}, digit persistence (test_mechanics.py:228-300)
  tool.step        free-fall closed form (tests/test_multi_tool.py:68-79)
  predict mode     x = 0.01 t, q_z = sin(0.1 t) (tests/test_coupling.py:322-339)
  compatible       tests/test_topology.py:10-12
  p2g              dense tent-function sums over every (particle, node) pair; total mass,
                   momentum and impulse; first moment of the impulse = sum j X^T - dt V sum P
  update_F         an affine velocity field v = A X + b gives F = I + dt A, v and x exact
                   (a swapped outer product, a missing dt or a transposed stress fails)
  MLS transfers    APIC carries an affine velocity field exactly (v = A x + b + g dt, C = A)
  MLS damage       J / J_ref thresholds (1 - eps_c)^m, (1 + eps_s)^m, off for J_ref = 0;
                   singular-value clamp at 0.05; Eq. 7 softening by gamma (|dev e| - limit)
                   with damage at zero yield stress, F returned to the yield surface
  blade constraint face particles leave a held-blade substep without velocity into the
                   faces beyond their gap (tests/test_blade_constraint.py:57-101)
  Maxwell branch   a held deformation relaxes as the closed-form exp(-n dt / tau) (float64
                   1e-12; float32 2e-5; tau = 3.5e5 steps within 1 %); the instantaneous
                   response is the elastic model's; a volume change does not relax; Fv stays
                   isochoric (tests/test_maxwell.py:90-136)
  viscosity        node velocities G X + b on a block at uniform F: no impulse at interior
                   nodes, -dt P N A on each face with P = eta J (L + L^T) F^-T, L = G F^-1; a
                   rigid spin gives none; a viscous material step conserves momentum
  extra knives     mirror-image knives stay mirror images and label their own digits; slot
                   swaps swap results; a knife off the block falls as -g t; each knife pays
                   its own cutting work (tests/test_multi_tool.py)
  threads          the OpenMP loops of the CPU baseline give the one-thread results bit for bit

Parity unpinned: the exact float32 bits of the reference's wp.svd3 and of Warp's quaternion
code generation (no Warp in this environment); the oracle's SVD is evaluated in float64.
"""

from dataclasses import replace

import numpy as np
import pytest
from scipy.linalg import polar
from scipy.spatial.transform import Rotation

import oracle
from oracle import OracleSim
from blademaster_hip.config import Config
from blademaster_hip.geometry import points, stencil


def sim(cfg, precision=64):
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, weights, gradients = stencil(ref, origin, cfg.grid_spacing, cells)
    return OracleSim(cfg, ref, nodes, weights, gradients, int(np.prod(cells + 1)), precision), ref


def test_compatible_known_cases():
    assert not oracle.compatible(1, 2)
    assert not oracle.compatible(4, 7)
    assert oracle.compatible(1, 4)
    assert oracle.compatible(0, 3**30 * 2)
    assert not oracle.compatible(3**33 * 2 + 1, 3**33 * 1 + 1)


@pytest.mark.parametrize("precision", [32, 64])
def test_bevel_face_is_exact_near_contact(precision):
    half = np.array([0.0005, 0.04, 0.02], np.float32)
    x = np.array(
        [[s * (half[0] + gap), y, z] for s in (-1, 1) for gap in (1e-8, 1e-7, 1e-6) for y in (-0.03, 0.01)
         for z in (-0.01476572, -0.005, 0.00126425, 0.019)],
        np.float32,
    )
    for p in x:
        c, nrm, d = oracle.blade_query(p, half, 0.005, precision)
        if precision == 32:
            assert np.float32(c[0]) == np.sign(p[0]) * half[0] and c[1] == p[1] and c[2] == p[2]
            assert np.float32(d) == np.abs(p[0]) - half[0]
        np.testing.assert_allclose(nrm, [np.sign(p[0]), 0, 0], atol=1e-7)


@pytest.mark.parametrize("bevel", [0.0, 0.005])
def test_blade_distance_against_brute_force_surface(bevel):
    """Points outside the cross-section: distance to the polygon boundary by point-segment projection."""
    half = np.array([0.0005, 0.025, 0.02])
    rng = np.random.default_rng(2)
    # dense surface sample of the cross-section polygon in (x, z), extruded in y
    if bevel:
        poly = np.array([[0, -half[2]], [half[0], -half[2] + bevel], [half[0], half[2]], [-half[0], half[2]],
                         [-half[0], -half[2] + bevel], [0, -half[2]]])
    else:
        poly = np.array([[-half[0], -half[2]], [half[0], -half[2]], [half[0], half[2]], [-half[0], half[2]],
                         [-half[0], -half[2]]])
    def segment_distance(q):
        best = np.inf
        for a, b in zip(poly[:-1], poly[1:]):
            t = np.clip(np.dot(q - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
            best = min(best, np.linalg.norm(q - (a + t * (b - a))))
        return best

    checked = 0
    while checked < 60:
        p = rng.uniform([-0.004, -0.02, -0.025], [0.004, 0.02, 0.025])
        inside_xz = np.all(np.abs(p[[0, 2]]) <= half[[0, 2]])
        if inside_xz:
            continue
        c, nrm, d = oracle.blade_query(p, half, bevel, 64)
        brute = segment_distance(p[[0, 2]])
        assert d == pytest.approx(brute, abs=1e-12)
        np.testing.assert_allclose(np.linalg.norm(nrm), 1.0, atol=1e-12)
        np.testing.assert_allclose(p - c, d * nrm, atol=1e-10)
        checked += 1


def test_friction_slip_magnitude_and_stick_history():
    args = dict(dt=0.001, mus=0.3, mud=0.2, gk=0.05, gc=0.1)
    n = np.array([0.0, 0.0, 1.0])
    j, h = oracle.friction(n, [1.0, 0.1, 0.3], 0.003, [0.0002, 0.0003, 0.0], 0.02, **args)
    assert np.linalg.norm(j) == pytest.approx(0.2 * 0.003, rel=2e-6)
    assert j[2] == 0.0
    # slip: the returned history reproduces the applied impulse through the spring-damper
    k = 0.05 * 0.02 / 0.001**2
    damping = 2 * 0.1 * np.sqrt(0.02 * k)
    vt = np.array([1.0, 0.1, 0.0])
    np.testing.assert_allclose((-k * h - damping * vt) * 0.001, j, rtol=1e-12, atol=1e-18)
    # stick: small sliding speed, no history
    j, h = oracle.friction(n, [0.001, 0.0, 0.0], 0.003, [0.0, 0.0, 0.0], 0.02, **args)
    np.testing.assert_allclose(h, [0.001 * 0.001, 0.0, 0.0], rtol=1e-12)
    np.testing.assert_allclose(j, [-damping * 0.001 * 0.001, 0.0, 0.0], rtol=1e-12)
    # normal history and velocity components are projected out
    j, h = oracle.friction(n, [0.0, 0.0, 5.0], 0.003, [0.0, 0.0, 0.2], 0.02, **args)
    np.testing.assert_array_equal(j, 0.0)
    np.testing.assert_array_equal(h, 0.0)


@pytest.mark.parametrize("angle", [0, 35])
def test_resist_cut_known_answer(angle):
    cfg = Config(size=(0.012, 0.012, 0.012), grid_spacing=0.006, blade_half_size=(0.0005, 0.025, 0.02))
    rot = Rotation.from_euler("y", angle, degrees=True)
    edge = np.array([0.0, 0.0, 0.02]) + rot.apply([0.0, 0.0, -0.02])
    positions = edge + rot.apply([[0, 0, -0.001]] * 5 + [[0, 0, 0.001]])
    velocities = rot.apply([[0, 0, -s] for s in [0.01, 0, -0.01, -0.01, 0.01, 0]])
    o = _resist_probe_sim(cfg, threshold=0.001)
    x = np.zeros((o.n, 3))
    x[:6] = positions
    damage = np.zeros(o.n)
    damage[:6] = [1, 1, 1, 0.5, 0.5, 1]
    o.set("x", x)
    o.set("v", np.zeros_like(x))
    o.set("damage", damage)
    q = rot.as_quat()
    j = rot.apply([0.002, 0.001, -0.004])
    results = [o.resist_cut(p, j, [0, 0, 0.02], velocities[p], q, [0, 0, 0], 0.0015, 1) for p in range(6)]
    expected = rot.apply([[0.002, 0.001, 0]] * 3 + [[0.002, 0.001, -0.004]] * 3)
    np.testing.assert_allclose(results, expected, atol=2e-9)
    np.testing.assert_allclose(o.get("damage")[:6], [1, 1, 1, 0.5, 0.54, 1], atol=1e-7)
    assert o.get("work")[0] == pytest.approx(0.004 * 0.01, rel=1e-6)


def _resist_probe_sim(cfg, threshold):
    """Oracle simulation whose damage threshold is replaced by the probe's value."""
    import ctypes

    o, _ = sim(cfg, 64)
    o.c.threshold = threshold
    o.lib.oracle_destroy(o.handle)
    o.handle = o.lib.oracle_create(64, ctypes.byref(o.c))
    return o


def _couple_fixture():
    codes = np.array([0, 1, 2, 4, 7], np.int64)
    masses = np.array([0.2, 0.5, 0.7, 0.3, 0.4])
    states = np.random.default_rng(3).normal(size=(5, 3))
    heads = np.array([0], np.int32)
    links = np.array([1, 2, 3, 4, -1], np.int32)
    return codes, masses, states, heads, links


def test_couple_conserves_momentum_and_respects_compatibility():
    codes, masses, states, heads, links = _couple_fixture()
    out = oracle.couple(codes, heads, links, masses, states, 1)
    np.testing.assert_allclose(out.sum(0), states.sum(0), atol=1e-14)
    # field 0 (uncut) is compatible with all: it gains from every other field
    assert np.abs(out[0] - states[0]).max() > 1e-3


def test_couple_all_compatible_equalises_and_incompatible_is_untouched():
    masses = np.array([0.2, 0.5, 0.7])
    states = np.random.default_rng(4).normal(size=(3, 3))
    heads = np.array([0], np.int32)
    links = np.array([1, 2, -1], np.int32)
    out = oracle.couple(np.zeros(3, np.int64), heads, links, masses, states, 1)
    np.testing.assert_allclose(out, masses[:, None] * states.sum(0) / masses.sum(), rtol=1e-13)
    out = oracle.couple(np.array([4, 5, 7], np.int64), heads, links, masses, states, 1)
    np.testing.assert_allclose(out, states, rtol=0, atol=0)


@pytest.mark.parametrize("flip", [0.0, 0.95, 1.0])
def test_g2p_known_answers(flip):
    slots = np.tile(np.arange(8), (2, 1))
    weights = np.full((2, 8), 1 / 8)
    x0 = np.array([[0.0, 0.0, 0.0], [0.003, 0.0, 0.0]])
    v0 = np.array([[10.0, 0.0, 0.0], [-10.0, 0.0, 0.0]])
    rest = np.zeros((8, 3))
    x, v = oracle.g2p(x0, v0, slots, weights, rest, rest, 0.001, flip)
    np.testing.assert_array_equal(x, x0)
    np.testing.assert_allclose(v, flip * v0, atol=1e-12)
    old = np.tile([1.0, 2.0, 3.0], (8, 1))
    new = np.tile([2.0, 4.0, 6.0], (8, 1))
    x, v = oracle.g2p(np.zeros((1, 3)), [[7.0, 2.0, -1.0]], slots[:1], weights[:1], old, new, 0.001, flip)
    np.testing.assert_allclose(x, [[0.002, 0.004, 0.006]], rtol=1e-12)
    np.testing.assert_allclose(v, [[2.0 + 6 * flip, 4.0, 6.0 - 4 * flip]], rtol=1e-12)


def test_svd3_against_numpy():
    rng = np.random.default_rng(7)
    mats = [np.eye(3), np.diag([1.3, 1 / 1.3, 1.0]), np.diag([2.0, 2.0, 0.5]), -np.eye(3)]
    mats += [np.eye(3) + 0.2 * rng.normal(size=(3, 3)) for _ in range(200)]
    mats += [rng.normal(size=(3, 3)) for _ in range(200)]
    for A in mats:
        U, s, V = oracle.svd3(A)
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, A, atol=1e-12)
        np.testing.assert_allclose(U @ U.T, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(V @ V.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        np.testing.assert_allclose(np.abs(s), np.linalg.svd(A, compute_uv=False), rtol=1e-12, atol=1e-14)
        assert s[0] >= 0 and s[1] >= 0
        assert np.sign(s.prod()) == np.sign(np.linalg.det(A)) or abs(np.linalg.det(A)) < 1e-14


def test_corotated_stress_closed_form():
    cfg = Config()
    o, _ = sim(cfg, 64)
    rng = np.random.default_rng(11)
    F = np.eye(3) + 0.05 * rng.normal(size=(o.n, 3, 3))
    o.set("F", F)
    o.stage("refresh_stress")
    P = o.get("stress").reshape(-1, 3, 3)
    k = o.k
    for f, p in zip(F[:50], P[:50]):
        R, _ = polar(f)
        J = np.linalg.det(f)
        expected = 2 * k["mu"] * (f - R) + k["lam"] * (J - 1) * J * np.linalg.inv(f).T
        np.testing.assert_allclose(p, expected, rtol=1e-9, atol=1e-9 * np.abs(expected).max())


def test_von_mises_projection():
    cfg = Config(yield_stress=100.0)
    o, _ = sim(cfg, 32)
    total = np.tile(np.diag([1.3, 1 / 1.3, 1.0]), (o.n, 1, 1)).astype(np.float32)
    o.set("F", total)
    o.stage("plastic_project")
    fp = o.get("Fp").reshape(-1, 3, 3)
    np.testing.assert_array_equal(o.get("F").reshape(-1, 3, 3), total)
    np.testing.assert_allclose(np.linalg.det(fp), 1, atol=3e-6)
    fe = total @ np.linalg.inv(fp)
    logs = np.log(np.linalg.svd(fe, compute_uv=False))
    dev = logs - logs.mean(1)[:, None]
    np.testing.assert_allclose(np.linalg.norm(dev, axis=1), 100 / (2 * o.k["mu"]) * np.sqrt(2 / 3), atol=2e-6)


def test_rigid_contact_momentum_balance_with_rotated_blade():
    cfg = Config(free_blade=True)
    o, _ = sim(cfg, 64)
    theta = 0.3
    o.set("tool_rot", [0, np.sin(theta / 2), 0, np.cos(theta / 2)])
    o.set("tool_pos", [0.0, 0.0, 0.02])
    o.stage("rigid", 0)
    gravity = np.array([0.0, 0.0, -9.81 * o.k["mass"] * cfg.dt]) * o.n
    material = o.get("impulses").reshape(-1, 3).sum(0)
    np.testing.assert_allclose(material + o.get("reaction") - o.get("board") - gravity, 0, atol=1e-15)
    assert np.linalg.norm(o.get("reaction")) > 0


def test_articulated_effective_mass_ratio():
    cfg = Config(external_tool=True)
    o, _ = sim(cfg, 64)
    o.set("tool_pos", [0, 0, 0.15])
    x = o.get("x").reshape(-1, 3)
    x[0] = [cfg.blade_half_size[0] + o.k["radius"] / 2, 0.01, 0.15]
    o.set("x", x)
    mobility = np.diag([1000, 2000, 3000, 4000, 5000, 6000.0])

    def impulse(a):
        o.set("tool_mobility", a)
        o.set("reaction", np.zeros(3))
        o.set("torque", np.zeros(3))
        o.set("histories", np.zeros(6 * o.n))
        o.stage("rigid", 0)
        return o.get("reaction")[0]

    prescribed = impulse(np.zeros((6, 6)))
    articulated = impulse(mobility)
    direction = np.array([1, 0, 0, 0, 0, -0.01])
    assert prescribed < 0
    assert articulated / prescribed == pytest.approx(1 / (1 + o.k["mass"] * (direction @ mobility @ direction)),
                                                     rel=1e-12)


@pytest.mark.parametrize("split", [False, True])
def test_pair_contact_equal_opposite_and_reference_exclusion(split):
    cfg = Config(size=(0.018, 0.018, 0.018), grid_spacing=0.006, dt=0.00025)
    o, _ = sim(cfg, 64)
    x = np.zeros((o.n, 3))
    x[:, 0] = np.arange(o.n) * 0.1 + 2
    x[:, 2] = 0.1
    x[:2] = [[0, 0, 0.1], [0.0048, 0, 0.1]]
    v = np.zeros_like(x)
    v[:2] = [[16.0, 0, 0], [-16.0, 0, 0]]
    codes = np.zeros(o.n, np.int64)
    codes[:2] = [1, 2] if split else [0, 0]
    # the two particles are reference neighbours (< 2h apart)
    ref = o.get("ref").reshape(-1, 3)
    assert np.linalg.norm(ref[0] - ref[1]) < 2 * cfg.grid_spacing
    o.set("x", x)
    o.set("v", v)
    o.set("codes", codes)
    o.set("impulses", np.zeros((o.n, 3)))
    o.set("speed", [16.0])
    o.stage("pairs")
    imp = o.get("impulses").reshape(-1, 3)
    assert o.scalars()["pair_count"] == (1 if split else 0)
    np.testing.assert_allclose(imp.sum(0), 0, atol=1e-18)
    assert bool(np.any(imp)) == split
    if split:
        assert imp[0, 0] < 0 < imp[1, 0]


def test_assign_labels_persist():
    cfg = Config()
    o, ref = sim(cfg, 32)
    o.set("tool_pos", [0.0, 0.0, 0.03])
    o.stage("assign", 1, 1)
    first = o.get("codes")
    assert set(np.unique(first)) == {0, 1, 2}
    assert ref[first == 1, 0].max() < 0 < ref[first == 2, 0].min()
    o.set("tool_pos", [0.02, 0.0, 0.1])
    o.stage("assign", 1, 1)
    np.testing.assert_array_equal(o.get("codes"), first)
    o.set("tool_pos", [0.012, 0.0, 0.03])
    o.stage("assign", 1, 3)
    np.testing.assert_array_equal(o.get("codes") % 3, first)
    assert o.get("codes").max() > 2


def test_free_knife_beside_the_block_falls_freely():
    cfg = Config(size=(0.06, 0.03, 0.03), grid_spacing=0.0075, dt=0.0005, free_blade=True,
                 blade_half_size=(0.0005, 0.025, 0.02))
    o, _ = sim(cfg, 32)
    z0 = 0.2
    o.set("tool_pos", [0.2, 0.0, z0])
    n = 100
    for _ in range(n):
        o.substep(cutting=True)
    dt, g = cfg.dt, 9.81
    assert o.get("tool_vel")[2] == pytest.approx(-g * n * dt, rel=1.0e-5)
    assert o.get("tool_pos")[2] == pytest.approx(z0 - g * dt * dt * n * (n + 1) / 2, abs=1.0e-6)
    np.testing.assert_allclose(o.get("tool_pos")[:2], [0.2, 0.0], atol=1e-9)


def test_predict_mode_integrates_the_twist():
    cfg = Config(external_tool=True, tool_substep="predict")
    o, _ = sim(cfg, 64)
    o.set("tool_pos", [0, 0, 0.15])
    o.set("tool_vel", [0.01, 0, 0])
    o.set("tool_omega", [0, 0, 0.2])
    o.set("tool_rot", [0, 0, 0, 1])
    steps = 7
    for _ in range(steps):
        o.stage("finish_external")
    t = steps * cfg.dt
    assert o.get("tool_pos")[0] == pytest.approx(0.01 * t, rel=1e-12)
    assert o.get("tool_rot")[2] == pytest.approx(np.sin(0.2 * t / 2), rel=1e-12)


def test_max_cut_index_capacity():
    from blademaster_hip.config import max_cut_index

    assert max_cut_index(225) == 33
    assert 225 * 3**34 - 1 <= np.iinfo(np.int64).max < 225 * 3**35 - 1


def test_mls_apic_reproduces_an_affine_velocity_field():
    """APIC transfers with quadratic B-splines carry an affine velocity field v = A x + b
    exactly (Jiang et al. 2015): after one substep from F = I (no stress) every particle has
    v = A x + b + g dt and C = A.  A wrong weight, D^-1 scale, dt or gravity sign fails."""
    from blademaster_hip.mlsmpm import TopoCutParams

    cfg = Config(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=1.0e-4)
    prm = TopoCutParams()
    ref, _, _ = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    ref = ref + np.array([0.0, 0.0, 0.03])  # clear of the board
    dx = cfg.grid_spacing
    low = ref.min(0) - prm.margin
    high = ref.max(0) + prm.margin
    low[2] = -2 * dx
    high[2] = ref[:, 2].max() + prm.headroom
    dims = np.ceil((high - low) / dx).astype(int) + 1
    o = oracle.OracleMls(cfg, prm, ref, low, dims, precision=64)
    A = np.array([[0.0, -0.2, 0.05], [0.2, 0.0, 0.0], [0.0, 0.03, 0.1]])
    b = np.array([0.01, -0.02, 0.005])
    v = ref @ A.T + b
    o.set("v", v)
    o.set("C", np.tile(A, (len(ref), 1, 1)))
    o.set("tool_pos", [0.0, 0.0, 1.0])  # knife far away
    o.substep(1)
    g = np.array([0.0, 0.0, -9.81]) * cfg.dt
    np.testing.assert_allclose(o.get("v").reshape(-1, 3), v + g, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("C").reshape(-1, 3, 3), np.tile(A, (len(ref), 1, 1)), rtol=0, atol=1e-9)
    np.testing.assert_allclose(o.get("x").reshape(-1, 3), ref + cfg.dt * (v + g), rtol=0, atol=1e-12)


def test_held_blade_face_particles_do_not_advance_into_the_faces():
    """With the knife held (external tool, hold mode) and pure PIC transfer, particles squeezed
    against both blade faces leave the substep with no velocity into a face beyond what closes
    their present gap (blade_constraint.py; reference tests/test_blade_constraint.py:57-101),
    while without the constraint they would move in at the squeeze speed."""
    cfg = Config(size=(0.036, 0.018, 0.018), grid_spacing=0.006, dt=0.0005, external_tool=True,
                 cache_topology=True, blade_half_size=(0.0005, 0.025, 0.02), flip_fraction=0.0)
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, w, gr = stencil(ref, origin, cfg.grid_spacing, cells)
    o = OracleSim(cfg, ref, nodes, w, gr, int(np.prod(cells + 1)), 64, 0)
    hx, hz = cfg.blade_half_size[0], cfg.blade_half_size[2]
    x = ref.copy()
    near = np.abs(x[:, 0]) < 0.002
    x[near, 0] = np.sign(x[near, 0]) * (hx + 5.0e-5)
    squeeze = 0.2
    v = np.zeros_like(x)
    v[:, 0] = -np.sign(ref[:, 0]) * squeeze
    o.set("x", x)
    o.set("v", v)
    o.set("tool_pos", [0.0, 0.0, cfg.size[2] + hz - 0.009])
    o.substep(cutting=False)
    side = o.get("bc_side")
    active = np.flatnonzero(side != 0)
    assert len(active) > 0
    vx = o.get("v").reshape(-1, 3)[active, 0]
    gap = np.abs(x[active, 0]) - hx
    assert np.all(side[active] * vx >= -np.maximum(gap, 0) / cfg.dt - 1e-9)
    assert np.max(np.abs(vx)) < 0.5 * squeeze
    assert o.get("bc_max_violation")[0] <= 1e-6


def _node_positions(origin, cells, h):
    dims = np.asarray(cells) + 1
    ix, iy, iz = np.meshgrid(*(np.arange(d) for d in dims), indexing="ij")
    return np.asarray(origin) + h * np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1)


def _hat_stencil(ref, node_x, h):
    """Trilinear weights and reference gradients of every node at every particle, brute force
    from the tent function N(u) = max(0, 1 - |u|) per axis (dense n x N)."""
    u = (ref[:, None, :] - node_x[None, :, :]) / h
    hat = np.clip(1.0 - np.abs(u), 0.0, None)
    dhat = np.where(np.abs(u) < 1.0, -np.sign(u) / h, 0.0)
    w = hat.prod(-1)
    grad = np.stack([dhat[..., 0] * hat[..., 1] * hat[..., 2], hat[..., 0] * dhat[..., 1] * hat[..., 2],
                     hat[..., 0] * hat[..., 1] * dhat[..., 2]], -1)
    return w, grad


def test_p2g_against_brute_force_tent_functions():
    """P2G (material.py:127-152) in float64 against a dense sum over every (particle, node)
    pair of tent-function weights: grid mass m w, momentum m w v, impulse w j - dt V P grad w.
    Moment identities of the trilinear basis then pin the terms independently of the dense
    sum: sum_a w_a = 1 gives the total mass and momentum, sum_a grad w_a = 0 leaves only the
    contact impulses in the total impulse, and sum_a x_a grad w_a^T = I turns the first moment
    of the impulse into sum_p j_p X_p^T - dt V sum_p P_p (a sign, a missing dt V or a
    transposed stress fails)."""
    cfg = Config(size=(0.036, 0.024, 0.018), grid_spacing=0.006, dt=2.0e-4, cache_topology=True)
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, weights, gradients = stencil(ref, origin, cfg.grid_spacing, cells)
    N = int(np.prod(np.asarray(cells) + 1))
    o = OracleSim(cfg, ref, nodes, weights, gradients, N, 64)
    rng = np.random.default_rng(11)
    n = len(ref)
    v = rng.normal(0.0, 0.1, (n, 3))
    imp = rng.normal(0.0, 1e-6, (n, 3))
    P = rng.normal(0.0, 50.0, (n, 3, 3))  # not symmetric
    o.set("v", v)
    o.set("impulses", imp)
    o.set("stress", P)
    o.stage("p2g")
    active = o.scalars()["active_count"]
    keys = o.get("grid_keys", 8 * n)[:active]
    assert np.array_equal(np.sort(keys), np.unique(nodes))  # one field per node, code 0
    mass = o.get("grid_mass", 8 * n)[:active]
    mom = o.get("momentum", 24 * n).reshape(-1, 3)[:active]
    jmp = o.get("grid_impulse", 24 * n).reshape(-1, 3)[:active]
    k = oracle.constants(cfg)
    m, V, dt, h = k["mass"], k["volume"], cfg.dt, cfg.grid_spacing
    node_x = _node_positions(origin, cells, h)
    w, grad = _hat_stencil(ref, node_x, h)
    want_mass = m * w.sum(0)
    want_mom = m * w.T @ v
    want_jmp = w.T @ imp - dt * V * np.einsum("pij,pnj->ni", P, grad)
    np.testing.assert_allclose(mass, want_mass[keys], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(mom, want_mom[keys], rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(jmp, want_jmp[keys], rtol=1e-10, atol=1e-15)
    # Moment identities.
    xk = node_x[keys]
    assert abs(mass.sum() - n * m) <= 1e-12 * n * m
    np.testing.assert_allclose(mom.sum(0), m * v.sum(0), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(jmp.sum(0), imp.sum(0), rtol=0, atol=1e-12 * np.abs(P).max() * dt * V * n / h)
    np.testing.assert_allclose(mom.T @ xk, m * v.T @ ref, rtol=1e-10, atol=1e-15)
    first = jmp.T @ xk
    np.testing.assert_allclose(first, imp.T @ ref - dt * V * P.sum(0), rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("precision", [64, 32])
def test_affine_velocity_field_updates_F_by_dt_A(precision):
    """One material step (material.py:127-475) of an affine velocity field v = A X + b from
    F = I, zero stress and no contact.  A node whose eight cells are full of the regular
    lattice gets the mass-weighted velocity A x + b, so a particle whose eight corners are
    such nodes (cells 1 .. c-2) keeps v = A X + b and moves to X + dt (A X + b).  The F update
    reads the re-scattered velocities, which are A x + b one layer further in, so a particle
    in cells 2 .. c-3 gets F = I + dt A by sum_a v_a grad w_a^T = A (an outer product with its
    arguments swapped gives I + dt A^T, a missing dt gives I + A)."""
    cfg = Config(size=(0.048, 0.042, 0.036), grid_spacing=0.006, dt=2.0e-4, cache_topology=True)
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, weights, gradients = stencil(ref, origin, cfg.grid_spacing, cells)
    N = int(np.prod(np.asarray(cells) + 1))
    o = OracleSim(cfg, ref, nodes, weights, gradients, N, precision)
    A = np.array([[0.3, -1.1, 0.4], [0.7, -0.2, 0.0], [-0.5, 0.9, 0.1]])
    b = np.array([0.02, -0.01, 0.03])
    v = ref @ A.T + b
    o.set("v", v)
    o.stage("material")
    h = cfg.grid_spacing
    cell = np.floor((ref - origin) / h).astype(int)
    top = np.asarray(cells)
    moved = np.all((cell >= 1) & (cell <= top - 2), axis=1)
    inner = np.all((cell >= 2) & (cell <= top - 3), axis=1)
    assert inner.sum() >= 100
    tol = 1e-12 if precision == 64 else 2e-6
    F = o.get("F").reshape(-1, 3, 3)[inner]
    np.testing.assert_allclose(F, np.tile(np.eye(3) + cfg.dt * A, (inner.sum(), 1, 1)), rtol=0, atol=tol)
    np.testing.assert_allclose(o.get("v").reshape(-1, 3)[moved], v[moved], rtol=0, atol=tol)
    x = o.get("x").reshape(-1, 3)[moved]
    np.testing.assert_allclose(x, ref[moved] + cfg.dt * v[moved], rtol=0, atol=tol * 0.1)


def _mls_block(cfg, prm, lift=0.03):
    ref, _, _ = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    ref = ref + np.array([0.0, 0.0, lift])  # clear of the board
    dx = cfg.grid_spacing
    low = ref.min(0) - prm.margin
    high = ref.max(0) + prm.margin
    low[2] = -2 * dx
    high[2] = ref[:, 2].max() + prm.headroom
    dims = np.ceil((high - low) / dx).astype(int) + 1
    return ref, low, dims


def test_mls_damage_threshold_and_singular_value_clamp():
    """MLS-MPM g2p damage (mlsmpm.py:196-229) on a resting block of negligible stiffness, so
    the substep leaves every F as it was: a particle is damaged exactly when
    det F / J_ref <= (1 - eps_c)^m or >= (1 + eps_s)^m with J_ref > 0, never with J_ref = 0,
    and a singular value below 0.05 is clamped to 0.05 (which then counts as compression)."""
    from blademaster_hip.mlsmpm import TopoCutParams

    cfg = Config(size=(0.036, 0.024, 0.018), grid_spacing=0.006, dt=1.0e-4, young_modulus=1.0e-6)
    prm = TopoCutParams(compression=0.025, stretch=0.01, sensitivity=2.0)
    ref, low, dims = _mls_block(cfg, prm)
    n = len(ref)
    o = oracle.OracleMls(cfg, prm, ref, low, dims, precision=64)
    rng = np.random.default_rng(5)
    J = rng.uniform(0.9, 1.05, n)
    J_ref = rng.choice([0.0, 1.0, 1.02], n)
    squashed = rng.random(n) < 0.05
    F = np.zeros((n, 3, 3))
    for p in range(n):
        R = Rotation.from_rotvec(rng.normal(0, 0.3, 3)).as_matrix()
        s = np.array([J[p] ** (1 / 3)] * 3)
        if squashed[p]:
            s = np.array([0.01, 1.0, 1.0])
        F[p] = R @ np.diag(s) @ R.T
    o.set("F", F)
    o.set("damage_reference", J_ref)
    o.set("tool_pos", [0.0, 0.0, 1.0])
    o.substep(1)
    out = o.get("F").reshape(-1, 3, 3)
    sv = np.linalg.svd(out, compute_uv=False)
    np.testing.assert_allclose(sv[squashed].min(1), 0.05, rtol=1e-9)
    np.testing.assert_allclose(out[~squashed], F[~squashed], rtol=0, atol=1e-9)
    det = np.where(squashed, 0.05, J)
    lo, hi = (1 - prm.compression) ** prm.sensitivity, (1 + prm.stretch) ** prm.sensitivity
    ratio = np.divide(det, J_ref, out=np.ones(n), where=J_ref > 0)
    # stay clear of the thresholds by more than the substep's change of F
    margin = np.minimum(np.abs(ratio - lo), np.abs(ratio - hi)) > 1e-7
    want = (J_ref > 0) & ((ratio <= lo) | (ratio >= hi))
    got = o.get("damage") > 0
    assert np.array_equal(got[margin], want[margin])
    assert want.sum() > 20 and (~want & (J_ref > 0)).sum() > 20 and squashed.sum() > 0


def test_mls_softening_follows_the_plastic_increment():
    """Hencky return map with Eq. 7 softening (mlsmpm.py:206-223): a trial strain with
    deviatoric norm |dev e| above the yield limit sigma_y / (2 mu) sqrt(2/3) is returned to the
    limit, and the yield stress drops by gamma (|dev e| - limit), floored at 0, where the
    particle is damaged; a particle below the limit keeps its F and yield stress."""
    from blademaster_hip.mlsmpm import TopoCutParams

    E, nu = 1.0e-6, 0.3
    mu = E / (2 * (1 + nu))
    cfg = Config(size=(0.036, 0.024, 0.018), grid_spacing=0.006, dt=1.0e-4, young_modulus=E, poisson_ratio=nu,
                 yield_stress=2.0e-8)
    gamma = 1.0e-6
    prm = TopoCutParams(softening=gamma, compression=0.5, stretch=0.5)
    ref, low, dims = _mls_block(cfg, prm)
    n = len(ref)
    o = oracle.OracleMls(cfg, prm, ref, low, dims, precision=64)
    rng = np.random.default_rng(9)
    stretch = rng.uniform(0.0, 0.08, n)
    F = np.zeros((n, 3, 3))
    for p in range(n):
        R = Rotation.from_rotvec(rng.normal(0, 0.3, 3)).as_matrix()
        s = np.exp(np.array([stretch[p], -stretch[p] / 2, -stretch[p] / 2]))  # volume preserving
        F[p] = R @ np.diag(s) @ R.T
    o.set("F", F)
    o.set("tool_pos", [0.0, 0.0, 1.0])
    o.substep(1)
    e = np.stack([stretch, -stretch / 2, -stretch / 2], 1)
    norm = np.linalg.norm(e - e.mean(1, keepdims=True), axis=1)
    limit = cfg.yield_stress / (2 * mu) * np.sqrt(2 / 3)
    clear = np.abs(norm - limit) > 1e-6
    yielded = norm > limit
    ys = o.get("yield_stress")
    want_ys = np.where(yielded, np.maximum(0.0, cfg.yield_stress - gamma * (norm - limit)), cfg.yield_stress)
    np.testing.assert_allclose(ys[clear], want_ys[clear], rtol=1e-6, atol=1e-14)
    out = o.get("F").reshape(-1, 3, 3)
    sv = np.linalg.svd(out, compute_uv=False)
    dev_out = np.log(sv) - np.log(sv).mean(1, keepdims=True)
    np.testing.assert_allclose(np.linalg.norm(dev_out, axis=1)[yielded & clear], limit, rtol=1e-6)
    np.testing.assert_allclose(out[~yielded & clear], F[~yielded & clear], rtol=0, atol=1e-9)
    dmg = o.get("damage") > 0
    assert np.array_equal(dmg[clear], (want_ys <= 0)[clear])
    assert yielded.sum() > 50 and (~yielded).sum() > 10 and (want_ys <= 0).sum() > 5


# ------------------------------------------------------------------ Maxwell branch, viscosity
def _rotation(axis, angle):
    return Rotation.from_rotvec(angle * np.asarray(axis, float) / np.linalg.norm(axis)).as_matrix()


# Finite stretch with shear and a 1 % volume change, rotated off the axes.
HELD_F = _rotation([1, 2, 3], 0.4) @ np.diag([1.08, 0.95, 0.985]) @ _rotation([0, 1, 1], -0.3)


def maxwell_closed_form(F, mu, lam, mu_m, factor):
    """First Piola stress of the standard linear solid at a held F when the Maxwell branch's
    deviatoric Hencky strain has decayed to `factor` of its initial value: fixed corotated
    2 mu_eq (F - R) + lam_eq (J - 1) J F^-T (mu_eq = mu - mu_m, lam_eq = lam + 2/3 mu_m), plus
    the Kirchhoff stress 2 mu_m factor dev(log V) pulled back by F^-T (V = R-rotated stretch,
    from numpy's SVD)."""
    u, s, vt = np.linalg.svd(F)
    R, J = u @ vt, np.prod(s)
    inv_t = np.linalg.inv(F).T
    equilibrium = 2 * (mu - mu_m) * (F - R) + (lam + 2 / 3 * mu_m) * (J - 1) * J * inv_t
    e = np.log(s)
    return equilibrium + u @ np.diag(2 * mu_m * factor * (e - e.mean())) @ u.T @ inv_t


def _held(cfg, F, precision):
    """Oracle with every particle at deformation F, registered, and no grid motion (zero grid
    mass: the F update only relaxes the Maxwell branch)."""
    o, _ = sim(cfg, precision)
    o.set("F", np.tile(F, (o.n, 1, 1)))
    o.stage("register")
    o.set("grid_mass", np.zeros(8 * o.n))
    o.stage("refresh_stress")
    return o


SMALL = Config(size=(0.02, 0.02, 0.02), grid_spacing=0.01)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tau_steps", [0.01, 1.0, 40.0])
def test_maxwell_held_deformation_relaxes_exactly(tau_steps, precision):
    """A held deformation relaxes the Maxwell branch by exp(-dt / tau) per step exactly, for
    any dt / tau (the reference's tests/test_maxwell.py:90-105); Fv stays isochoric and F
    unchanged.  A relaxation of the full Hencky strain, of mu instead of mu_eq, or a missing
    F^-T fails at step 0 or 1."""
    cfg = replace(SMALL, maxwell_tau=tau_steps * SMALL.dt, maxwell_fraction=0.6)
    k = oracle.constants(cfg)
    o = _held(cfg, HELD_F, precision)
    tol = 1e-12 if precision == 64 else 2e-5
    for step in range(121):
        if step:
            o.stage("update_F")
        if step in (0, 1, 5, 40, 120):
            want = maxwell_closed_form(HELD_F, k["mu"], k["lam"], k["mu_maxwell"], np.exp(-step / tau_steps))
            got = o.get("stress").reshape(-1, 3, 3)
            np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), atol=tol * np.abs(want).max(),
                                       err_msg=f"step {step}")
    assert not o.get("errors").any()
    np.testing.assert_allclose(o.get("F").reshape(-1, 3, 3), np.broadcast_to(HELD_F, (o.n, 3, 3)), atol=1e-7)
    det = np.linalg.det(np.eye(3) + o.get("viscous").reshape(-1, 3, 3))
    np.testing.assert_allclose(det, 1.0, atol=1e-12 if precision == 64 else 2e-6)


def test_maxwell_slow_relaxation_is_resolved_in_float32():
    """tau = 3.5e5 steps: the per-step change of Fv is below float32 resolution, and the
    float64 Fv keeps the float32 mirror within 1 % of the relaxation (the reference's
    tests/test_maxwell.py:108-119, where a float32 Fv was 39 % off)."""
    tau_steps, steps = 3.5e5, 20000
    cfg = replace(Config(size=(0.01, 0.01, 0.01), grid_spacing=0.005), maxwell_tau=tau_steps * SMALL.dt,
                  maxwell_fraction=0.6)
    k = oracle.constants(cfg)
    o = _held(cfg, HELD_F, 32)
    start = o.get("stress").reshape(-1, 3, 3)[0]
    for _ in range(steps):
        o.stage("update_F")
    want = maxwell_closed_form(HELD_F, k["mu"], k["lam"], k["mu_maxwell"], np.exp(-steps / tau_steps))
    error = np.abs(o.get("stress").reshape(-1, 3, 3)[0] - want).max()
    assert error < 0.01 * np.abs(start - want).max()


def test_maxwell_instantaneous_and_volumetric_response():
    """The instantaneous small-strain response equals the elastic model's (mu_eq and lam_eq
    are chosen for it); a pure volume change loads no Maxwell stress and does not relax
    (tests/test_maxwell.py:122-136)."""
    cfg = replace(SMALL, maxwell_tau=10 * SMALL.dt, maxwell_fraction=0.7)
    small = np.eye(3) + 1e-3 * np.array([[0.5, 0.8, 0.0], [0.1, -0.3, 0.2], [0.0, 0.4, 0.6]])
    elastic = _held(SMALL, small, 64).get("stress").reshape(-1, 3, 3)[0]
    viscoelastic = _held(cfg, small, 64).get("stress").reshape(-1, 3, 3)[0]
    np.testing.assert_allclose(viscoelastic, elastic, atol=3e-3 * np.abs(elastic).max())
    o = _held(cfg, 0.98 * np.eye(3), 64)
    first = o.get("stress").reshape(-1, 3, 3)[0].copy()
    for _ in range(50):
        o.stage("update_F")
    scale = np.abs(first).max()
    np.testing.assert_allclose(o.get("stress").reshape(-1, 3, 3)[0], first, atol=1e-12 * scale)
    np.testing.assert_allclose(first, np.diag(np.diag(first)), atol=1e-12 * scale)
    # At zero fraction or zero tau the branch is off: the stress is the elastic one.
    for off in (replace(SMALL, maxwell_tau=0.0, maxwell_fraction=0.5), replace(SMALL, maxwell_tau=0.1)):
        assert not oracle.constants(off)["maxwell"]
        np.testing.assert_array_equal(_held(off, HELD_F, 64).get("stress"), _held(SMALL, HELD_F, 64).get("stress"))


def _viscous_fixture(F):
    """Registered oracle (float64) of a 4 x 3 x 2-cell block at uniform deformation F, with
    the node position of every field slot."""
    cfg = Config(size=(0.04, 0.03, 0.02), grid_spacing=0.01, dynamic_viscosity=2.0)
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, weights, gradients = stencil(ref, origin, cfg.grid_spacing, cells)
    dims = np.asarray(cells) + 1
    N = int(np.prod(dims))
    o = OracleSim(cfg, ref, nodes, weights, gradients, N, 64)
    o.set("F", np.tile(F, (o.n, 1, 1)))
    o.stage("register")
    count = o.scalars()["active_count"]
    node = o.get("grid_keys", 8 * o.n)[:count] % N
    index = np.stack([node // (dims[1] * dims[2]), (node // dims[2]) % dims[1], node % dims[2]], 1)
    return cfg, o, count, index, np.asarray(origin) + cfg.grid_spacing * index, np.asarray(cells)


def _viscous_impulses(o, count, old_v):
    full = np.zeros((8 * o.n, 3))
    full[:count] = old_v
    o.set("old_v", full)
    o.set("grid_impulse", np.zeros(24 * o.n))
    o.stage("viscous")
    return o.get("grid_impulse").reshape(-1, 3)[:count]


@pytest.mark.parametrize("deformed", [False, True])
def test_viscous_impulse_face_tractions(deformed):
    """viscous_impulse (material.py:321-340) of the node velocities v = G X + b (X the
    reference node positions, so dF/dt = G) on a block at uniform F: the Newtonian Cauchy
    stress 2 eta sym(L), L = G F^-1, pulled back to P = J sigma F^-T.  The divergence of a
    uniform stress vanishes, so interior nodes get nothing, and the nodes of each face
    together get -dt P N A (N its outward normal, A its reference area)."""
    F = np.diag([1.1, 0.95, 1.02]) @ _rotation([1, -1, 2], 0.3) if deformed else np.eye(3)
    cfg, o, count, index, X, cells = _viscous_fixture(F)
    G = np.array([[0.3, 1.2, -0.4], [0.5, -0.2, 0.9], [0.1, 0.7, 0.25]])
    imp = _viscous_impulses(o, count, X @ G.T + [0.1, -0.2, 0.3])
    L = G @ np.linalg.inv(F)
    P = cfg.dynamic_viscosity * np.linalg.det(F) * (L + L.T) @ np.linalg.inv(F).T
    scale = cfg.dt * np.abs(P).max() * cfg.size[0] * cfg.size[1]
    interior = np.all((index > 0) & (index < cells), axis=1)
    assert interior.sum() > 0
    np.testing.assert_allclose(imp[interior], 0.0, atol=1e-12 * scale)
    for axis in range(3):
        area = np.prod([cfg.size[k] for k in range(3) if k != axis])
        for side, sign in ((0, -1.0), (cells[axis], 1.0)):
            face = index[:, axis] == side
            normal = np.zeros(3)
            normal[axis] = sign
            np.testing.assert_allclose(imp[face].sum(0), -cfg.dt * P @ normal * area, rtol=1e-10,
                                       atol=1e-12 * scale)


def test_viscous_impulse_of_a_rigid_spin_is_zero():
    """A rigid spin of a deformed block (dF/dt = W F, W skew) has L = W: no viscous stress.
    A velocity gradient taken as F^-1 G (a swapped product) is not skew for a stretched F and
    fails."""
    F = np.diag([1.2, 0.9, 1.05]) @ _rotation([0, 1, 1], 0.5)
    cfg, o, count, index, X, cells = _viscous_fixture(F)
    W = np.array([[0.0, -0.7, 0.4], [0.7, 0.0, -1.1], [-0.4, 1.1, 0.0]])
    G = W @ F
    imp = _viscous_impulses(o, count, X @ G.T)
    moving = _viscous_impulses(o, count, X @ (np.diag([0.3, -0.2, 0.1]) @ F).T)
    assert np.abs(moving).max() > 1e-9
    np.testing.assert_allclose(imp, 0.0, atol=1e-12 * np.abs(moving).max())


def test_viscous_material_step_conserves_momentum():
    """A material step with viscosity (solver.py:433-458): the viscous impulses sum to zero
    (the stencil gradients of a particle sum to zero), so the particles' momentum changes by
    the contact impulses only (tests/test_mechanics.py:183-230, in float64)."""
    cfg = Config(size=(0.04, 0.03, 0.02), grid_spacing=0.005, dynamic_viscosity=2.0)
    o, ref = sim(cfg, 64)
    rng = np.random.default_rng(4)
    v = rng.normal(scale=0.01, size=(o.n, 3))
    impulse = rng.normal(scale=1e-8, size=(o.n, 3))
    o.set("v", v)
    o.set("F", np.tile(np.diag([1.01, 0.995, 1.0]), (o.n, 1, 1)))
    o.stage("refresh_stress")
    o.set("codes", np.where(ref[:, 0] < 0, 1, 2))
    o.set("impulses", impulse)
    mass = oracle.constants(cfg)["mass"]
    o.stage("material")
    after = mass * o.get("v").reshape(-1, 3).sum(0)
    np.testing.assert_allclose(after, mass * v.sum(0) + impulse.sum(0), rtol=0, atol=1e-15)
    # and the viscosity acted: the same step without it differs
    o2, _ = sim(replace(cfg, dynamic_viscosity=0.0), 64)
    o2.set("v", v)
    o2.set("F", np.tile(np.diag([1.01, 0.995, 1.0]), (o.n, 1, 1)))
    o2.stage("refresh_stress")
    o2.set("codes", np.where(ref[:, 0] < 0, 1, 2))
    o2.set("impulses", impulse)
    o2.stage("material")
    assert np.abs(o2.get("v") - o.get("v")).max() > 1e-9


# ------------------------------------------------------------------ several free knives
def two_knives(precision=64, **fields):
    cfg = Config(size=(0.06, 0.03, 0.03), grid_spacing=0.0075, dt=0.0005, free_blade=True, tool_count=2,
                 blade_half_size=(0.0005, 0.025, 0.02), **fields)
    o, ref = sim(cfg, precision)
    return cfg, o, ref


def _dropped(xs, steps, precision=64):
    cfg, o, ref = two_knives(precision)
    z = 0.03 + 0.02 + 0.002
    o.set("tool_pos", [[x, 0.0, z] for x in xs])
    o.set("tool_vel", [[0.0, 0.0, -0.6]] * 2)
    for _ in range(steps):
        o.substep(cutting=True)
    return o, ref


def test_second_knife_beside_the_block_falls_freely():
    """Knife 1 well away from the block falls as -g t (tests/test_multi_tool.py:68-79) while
    knife 0 cuts; neither disturbs the other."""
    cfg, o, _ = two_knives(32)
    z0 = 0.2
    o.set("tool_pos", [[-0.015, 0.0, 0.03 + 0.02 + 0.002], [0.2, 0.0, z0]])
    n = 100
    for _ in range(n):
        o.substep(cutting=True)
    dt, g = cfg.dt, 9.81
    vel, pos = o.get("tool_vel").reshape(2, 3), o.get("tool_pos").reshape(2, 3)
    assert vel[1, 2] == pytest.approx(-g * n * dt, rel=1.0e-5)
    assert pos[1, 2] == pytest.approx(z0 - g * dt * dt * n * (n + 1) / 2, abs=1.0e-6)
    np.testing.assert_allclose(pos[1, :2], [0.2, 0.0], atol=1e-9)
    assert pos[0, 2] > 0.02 and vel[0, 2] > -0.6  # knife 0 met the block
    assert not o.get("extra_damage").any()  # knife 1 damaged nothing


def test_mirrored_knives_move_alike_and_label_their_own_digits():
    """Two knives dropped at x = -15 and +15 mm into a block symmetric about x = 0 move as mirror
    images and each labels its own ternary digit on its own side split
    (tests/test_multi_tool.py:46-65).  A knife that read another's pose, history or damage
    breaks the mirror."""
    steps = 60
    o, ref = _dropped([-0.015, 0.015], steps)
    p, v = o.get("tool_pos").reshape(2, 3), o.get("tool_vel").reshape(2, 3)
    np.testing.assert_allclose(p[0] * [-1, 1, 1], p[1], atol=1.0e-10)
    np.testing.assert_allclose(v[0] * [-1, 1, 1], v[1], atol=1.0e-8)
    assert 0.0 < p[0, 2] - 0.02 < 0.028
    assert v[0, 2] > -0.6 - 9.81 * steps * 0.0005
    codes = o.get("codes")
    first, second = codes % 3, (codes // 3) % 3
    assert set(np.unique(first)) == {0, 1, 2} and set(np.unique(second)) == {0, 1, 2}
    assert np.all(codes < 9)
    assert ref[first == 1, 0].max() < -0.015 < ref[first == 2, 0].min()
    assert ref[second == 1, 0].max() < 0.015 < ref[second == 2, 0].min()


def test_knife_index_does_not_matter():
    """Swapping the knives' slots swaps their results (tests/test_multi_tool.py:34-43)."""
    a, _ = _dropped([-0.015, 0.015], 40)
    b, _ = _dropped([0.015, -0.015], 40)
    np.testing.assert_allclose(a.get("tool_pos").reshape(2, 3), b.get("tool_pos").reshape(2, 3)[::-1], atol=1e-10)
    np.testing.assert_allclose(a.get("x"), b.get("x"), atol=1e-10)


@pytest.mark.parametrize("first", [0, 1])
def test_each_knife_pays_its_own_cutting_work(first):
    """One particle under the edge of knife `first`: that knife's damage of it reaches 1 and its
    reaction is recorded; moved under the other knife, the other knife's damage starts from
    zero and its reaction repeats the first (tests/test_multi_tool.py:82-104)."""
    cfg, o, ref = two_knives(64, cutting_energy=1.0)
    x = ref.copy()
    x[:, 0] += 10
    x[0] = [0, 0, 0.039]
    o.set("x", x)
    positions = np.array([[3.0, 0, 0.06], [3.0, 0, 0.06]])
    positions[first, 0] = 0
    o.set("tool_pos", positions)
    o.set("tool_vel", [[0, 0, -1], [0, 0, -1]])
    o.stage("rigid", 1)
    reaction = o.get("reaction").reshape(2, 3)
    assert np.abs(reaction[first]).max() > 0 and not reaction[1 - first].any()
    damage = [o.get("damage")[0], o.get("extra_damage")[0]]
    assert damage[first] == 1 and damage[1 - first] == 0
    positions[:, 0] = positions[::-1, 0]
    o.set("tool_pos", positions)
    o.set("reaction", np.zeros(6))
    o.set("torque", np.zeros(6))
    o.stage("rigid", 1)
    np.testing.assert_allclose(o.get("reaction").reshape(2, 3)[1 - first], reaction[first], rtol=1e-12)
    assert [o.get("damage")[0], o.get("extra_damage")[0]][1 - first] == 1


@pytest.mark.parametrize("tools", [1, 2])
def test_threaded_oracle_equals_one_thread(tools):
    """The OpenMP loops (CPU baseline) keep every sum's order: a cutting run with pair
    contacts, relabelling, plasticity and one or two knives is bit-identical on 1 and 4
    threads."""
    extra = dict(free_blade=True, tool_count=2) if tools == 2 else {}
    cfg = Config(size=(0.048, 0.024, 0.018), grid_spacing=0.006, dt=0.0002, cutting_energy=0.05,
                 yield_stress=2000.0, **extra)
    runs = []
    for threads in (1, 4):
        o, ref = sim(cfg, 32)
        assert o.set_threads(threads) == threads
        rng = np.random.default_rng(0)
        o.set("v", rng.normal(0, 0.01, (o.n, 3)))
        hz = cfg.blade_half_size[2]
        o.set("tool_pos", [[x, 0.0, cfg.size[2] + hz - 0.002] for x in (0.0, 0.012)[:tools]])
        o.set("tool_vel", [[0.0, 0.0, -0.05]] * tools)
        for _ in range(12):
            o.substep(cutting=True)
        runs.append({name: o.get(name) for name in ("x", "v", "F", "Fp", "stress", "damage", "histories", "codes",
                                                    "tool_pos", "tool_vel", "reaction", "work", "errors")})
    assert np.count_nonzero(runs[0]["codes"]) > 0
    for name in runs[0]:
        np.testing.assert_array_equal(runs[0][name], runs[1][name], err_msg=name)
