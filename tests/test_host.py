"""Host-side logic without a GPU: configuration checks, sampling and stencil, piece labels,
the coupling value types, and the C-ABI library's exported symbols."""

import ctypes
import os
import re

import numpy as np
import pytest

from blademaster_hip.config import Config, derived, max_cut_index
from blademaster_hip.geometry import points, stencil, compact_stencil, levelset, BANANA_DEMO_SIZE_M
from blademaster_hip.topology import components, compatible
from blademaster_hip.coupling import ToolState, Reaction

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- Config (solver.py:87-149)
@pytest.mark.parametrize(
    "kwargs",
    [
        dict(grid_storage="dense"),
        dict(pair_search="tree"),
        dict(pair_skin=-1.0),
        dict(cut_work="absolute"),
        dict(tool_substep="predict"),
        dict(dt=0.0),
        dict(poisson_ratio=0.5),
        dict(friction_dynamic=0.4),
        dict(normal_scale=1.0, contact_reference_dt=0.0001),
        dict(board_normal_scale=2.0),
        dict(blade_bevel_height=0.05),
        dict(free_blade=True, pose_servo=True),
        dict(tool_count=2),
        dict(tool_count=0),
        dict(grid_spacing=0.02),
        dict(young_modulus=float("nan")),
        dict(fields_per_node=0),
    ],
)
def test_config_rejects_invalid(kwargs):
    with pytest.raises(ValueError):
        Config(**kwargs)


def test_config_contact_rescaling_and_derived_constants():
    cfg = Config(dt=0.0002, contact_reference_dt=0.0001, normal_scale=0.1, tangent_stiffness=0.05)
    assert cfg.contact_scales() == pytest.approx((0.4, 0.2))
    d = derived(Config())
    h = 0.0075
    assert d["volume"] == pytest.approx((h / 2) ** 3)
    assert d["mass"] == pytest.approx(1100 * (h / 2) ** 3)
    assert d["radius"] == pytest.approx(h / 4)
    E, nu = 6000.0, 0.35
    assert d["mu"] == pytest.approx(E / (2 * (1 + nu)))
    assert d["lam"] == pytest.approx(E * nu / ((1 + nu) * (1 - 2 * nu)))
    # blade contact offset: max(radius, radius - half thickness) (solver.py:184)
    assert d["blade_offset"] == pytest.approx(h / 4)
    assert max_cut_index(225) == 33


# ---------------------------------------------------------------- sampling (geometry.py)
def test_particle_and_node_counts():
    x, _, cells = points((0.168, 0.06, 0.024), 0.006, "block")
    assert len(x) == 8960 and tuple(cells) == (28, 10, 4)
    x, _, cells = points((0.3, 0.15, 0.075), 0.003, "block")
    assert len(x) == 1_000_000 and int(np.prod(cells + 1)) == 133926
    x, _, _ = points((0.072, 0.06, 0.03), 0.006, "banana")
    assert len(x) == 892
    assert (levelset(x, (0.072, 0.06, 0.03), "banana") <= 0).all()
    for spacing, count in [(0.006, 1452), (0.004, 4892), (0.003, 11520)]:
        assert len(points(BANANA_DEMO_SIZE_M, spacing, "banana")[0]) == count
    for spacing, nodes in [(0.015, 45), (0.0075, 225), (0.00375, 1377)]:
        _, _, cells = points((0.06, 0.03, 0.03), spacing, "block")
        assert int(np.prod(cells + 1)) == nodes


def test_banana_clipping_is_rejected():
    with pytest.raises(ValueError, match="clipped"):
        points((0.072, 0.024, 0.03), 0.006, "banana")


@pytest.mark.parametrize("shape", ["banana", "block"])
def test_jittered_particles_start_clear_of_the_board(shape):
    size = BANANA_DEMO_SIZE_M if shape == "banana" else (0.06, 0.03, 0.03)
    for seed in (1, 2):
        x, _, _ = points(size, 0.006, shape, seed)
        assert x[:, 2].min() >= 0.006 / 4


@pytest.mark.parametrize("seed", [0, 3])
def test_stencil_partition_of_unity_and_compact_form(seed):
    cfg = Config(sampling_seed=seed)
    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, w, g = stencil(ref, origin, cfg.grid_spacing, cells)
    np.testing.assert_allclose(w.sum(1), 1, atol=1e-14)
    np.testing.assert_allclose(g.sum(1), 0, atol=1e-10)
    # gradient = d weight / d x, checked by central differences of the trilinear weights
    eps = 1e-7
    for axis in range(3):
        d = np.zeros(3)
        d[axis] = eps
        _, wp, _ = stencil(ref[:50] + d, origin, cfg.grid_spacing, cells)
        _, wm, _ = stencil(ref[:50] - d, origin, cfg.grid_spacing, cells)
        np.testing.assert_allclose((wp - wm) / (2 * eps), g[:50, :, axis], rtol=1e-5, atol=1e-3)
    base, local, _ = compact_stencil(ref, origin, cfg.grid_spacing, cells)
    dims = cells + 1
    np.testing.assert_array_equal(nodes[:, 0], base)
    np.testing.assert_array_equal(nodes[:, 7], base + dims[1] * dims[2] + dims[2] + 1)
    fx = np.stack([1 - local.astype(np.float64), local.astype(np.float64)], axis=-1)
    rebuilt = np.stack([fx[:, 0, (a >> 2) & 1] * fx[:, 1, (a >> 1) & 1] * fx[:, 2, a & 1] for a in range(8)], 1)
    np.testing.assert_allclose(rebuilt, w, atol=1e-7)


# ---------------------------------------------------------------- piece labels (topology.py)
def test_uncut_bridges_and_persistent_conflicts():
    nodes = np.zeros((3, 1), dtype=int)
    weights = np.ones((3, 1))
    assert components(nodes, weights, [1, 0, 2])[1].tolist() == [3]
    assert components(nodes, weights, [1, 1, 2])[1].tolist() == [2, 1]
    assert not compatible(1, 2)
    assert not compatible(4, 7)
    assert compatible(1, 4)


def test_zero_weight_and_spatially_disjoint_pieces():
    nodes = np.array([[0, 1], [1, 2], [3, 4]])
    weights = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 0.0]])
    assert components(nodes, weights, [0, 0, 0])[1].tolist() == [1, 1, 1]


def test_components_against_union_find_brute_force():
    """Random multi-cut codes on a real stencil against a direct union-find restatement."""
    rng = np.random.default_rng(5)
    ref, origin, cells = points((0.048, 0.024, 0.018), 0.006, "block")
    nodes, weights, _ = stencil(ref, origin, 0.006, cells)
    codes = np.where(ref[:, 0] < -0.01, 1, np.where(ref[:, 0] > 0.01, 2, 0)) + 3 * rng.integers(0, 3, len(ref))
    labels, sizes = components(nodes, weights, codes)
    n = len(ref)
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    by_node = {}
    for p in range(n):
        for a in range(8):
            by_node.setdefault(int(nodes[p, a]), []).append(p)
    for plist in by_node.values():
        for i in range(len(plist)):
            for j in range(i + 1, len(plist)):
                p, q = plist[i], plist[j]
                # p, q share a node: connected if their codes meet through compatible fields
                if compatible(codes[p], codes[q]):
                    parent[find(p)] = find(q)
    # fields with equal code link their particles; compatible distinct codes link too
    roots = np.array([find(p) for p in range(n)])
    _, first_index, inverse = np.unique(roots, return_index=True, return_inverse=True)
    order = np.argsort(first_index)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    np.testing.assert_array_equal(labels, rank[inverse])
    assert sizes.sum() == n


# ---------------------------------------------------------------- coupling value types
def test_tool_state_validation():
    ToolState([0, 0, 0], [0, 0, 0, 1], [0, 0, 0], [0, 0, 0]).checked()
    with pytest.raises(ValueError):
        ToolState([0, 0, 0], [0, 0, 0, 2], [0, 0, 0], [0, 0, 0]).checked()
    for invalid in [np.eye(5), -np.eye(6), np.full((6, 6), np.nan), np.triu(np.ones((6, 6)))]:
        with pytest.raises(ValueError, match="Mobility"):
            ToolState([0, 0, 0], [0, 0, 0, 1], [0, 0, 0], [0, 0, 0], invalid).checked()


def test_moment_shift_and_power_invariance():
    r = Reaction(0.02, np.array([0.1, 0.2, 0.3]), np.array([0.01, -0.02, 0.03]), np.array([1.0, 2.0, 3.0]))
    origin = np.array([-0.5, 0.3, 2.0])
    shifted = r.about(origin)
    velocity = np.array([0.2, -0.1, 0.7])
    omega = np.array([0.3, 0.5, -0.2])
    v_new = velocity + np.cross(omega, origin - r.origin_world_m)
    assert np.dot(r.force_n, velocity) + np.dot(r.torque_nm, omega) == pytest.approx(
        np.dot(shifted.force_n, v_new) + np.dot(shifted.torque_nm, omega)
    )


# ---------------------------------------------------------------- C-ABI library
def _declared_symbols():
    symbols = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            symbols |= set(re.findall(r"^(?:int|int64_t)\s+(bm_\w+)\s*\(", text, flags=re.M))
    return symbols


def test_library_loads_and_exports_every_declared_symbol():
    from blademaster_hip import _lib

    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in sorted(declared):
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES or name == "bm_piece_labels_static_bytes", f"{name} has no ctypes signature"
    assert lib.bm_version() > 0


def test_struct_layouts_match_the_header():
    """Every field of bm_params / bm_state in the header appears, in order, in the ctypes mirror."""
    from blademaster_hip import _lib

    text = open(os.path.join(ROOT, "include", "bm_hip.h")).read()
    for struct, mirror in (("bm_params", _lib.Params), ("bm_state", _lib.State), ("bm_labels", _lib.Labels),
                           ("bm_mls_params", _lib.MlsParams), ("bm_mls_state", _lib.MlsState),
                           ("bm_mlp", _lib.Mlp)):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, flags=re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            declarators = re.sub(r"^(const\s+)?\w+\s*", "", decl)
            for item in declarators.split(","):
                item = item.strip().lstrip("*").strip()
                names.append(re.sub(r"\[.*\]", "", item))
        assert names == [f[0] for f in mirror._fields_], struct


def test_library_argument_errors_need_no_device():
    from blademaster_hip import _lib

    lib = _lib.load()
    assert lib.bm_refresh_stress(None, None, None) < 0
    assert lib.bm_substep(None, None, 0, 0, 1, 0, 0, None, None, 0, None) < 0
    assert lib.bm_mlp_forward(None, None, 1, None) < 0
    m = _lib.Mlp(inputs=96, hidden=128, outputs=6)
    assert lib.bm_mlp_infer(m, None, None, 1, None) < 0
    params, packed, work = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    assert lib.bm_mlp_sizes(96, 256, 6, ctypes.byref(params), ctypes.byref(packed), ctypes.byref(work)) == 0
    assert work.value == 16 * 257 * 256
    assert params.value == 97 * 256 + 257 * 256 + 257 * 6
    assert packed.value == 256 * 128 + 256 * 288 + 16 * 288


def test_static_labels_scratch_bound_is_checked_on_the_host():
    """bm_piece_labels_static rejects missing buffers and sizes beyond its static scratch
    before it touches a device."""
    from blademaster_hip import _lib

    lib = _lib.load()
    cap = 64 << 20
    need = lib.bm_piece_labels_static_bytes(1_000_000, 133_926, 16)
    assert 0 < need <= cap  # the 1M-particle block fits with 16 codes per node
    assert lib.bm_piece_labels_static_bytes(1_000_000, 133_926, 64) > cap
    assert lib.bm_piece_labels_static_bytes(10, 10, 0) < 0
    fake = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks first
    p = _lib.Params()
    s = _lib.State()
    s.codes = ctypes.cast(fake, _lib.c_i64p)
    s.base_node = ctypes.cast(fake, _lib.c_i32p)
    assert lib.bm_piece_labels_static(None, None, 16, fake, fake, fake, fake, None) == -1
    p.n, p.node_count = 1_000_000, 133_926
    assert lib.bm_piece_labels_static(ctypes.byref(p), ctypes.byref(s), 64, fake, fake, fake, fake, None) == -3
    assert lib.bm_piece_labels_static(ctypes.byref(p), ctypes.byref(s), 0, fake, fake, fake, fake, None) == -1
    assert lib.bm_piece_labels_static(ctypes.byref(p), ctypes.byref(s), 16, None, fake, fake, fake, None) == -1


@pytest.mark.parametrize("shape,size,spacing,fields", [("block", (0.3, 0.15, 0.075), 0.003, 4),
                                                       ("banana", BANANA_DEMO_SIZE_M, 0.003, 4),
                                                       ("block", (0.048, 0.024, 0.018), 0.006, 32)])
def test_brick_bins_cover_every_particle_window(shape, size, spacing, fields):
    from blademaster_hip.geometry import bin_bricks

    ref, origin, cells = points(size, spacing, shape)
    base, local, base_cell = compact_stencil(ref, origin, spacing, cells)
    brick, bricks, slots = bin_bricks(cells, base_cell, fields)
    used = slots[slots >= 0]
    np.testing.assert_array_equal(np.sort(used), np.arange(len(ref)))
    assert np.all(bricks * brick >= cells)
    window, corners = fields * np.prod(brick + 1), 8 * np.prod(brick)
    assert max(7 * window + 7 * corners, 3 * window + 3 * corners) * 4 <= 65536
    per_brick = int(np.prod(brick))
    assert len(slots) == 8 * int(np.prod(bricks)) * per_brick
    cell_slots = slots.reshape(-1, 8)
    # used slots are a prefix of each cell, in reference order
    filled = cell_slots >= 0
    assert np.all(filled[:, 1:] <= filled[:, :-1])
    both = filled[:, 1:] & filled[:, :-1]
    assert np.all(cell_slots[:, 1:][both] > cell_slots[:, :-1][both])
    # the particles of slot group c are exactly those of cell c of the brick order
    c = np.repeat(np.arange(len(cell_slots)), filled.sum(1))
    nb, cl = c // per_brick, c % per_brick
    bz = nb % bricks[2]
    by = (nb // bricks[2]) % bricks[1]
    bx = nb // (bricks[1] * bricks[2])
    lz = cl % brick[2]
    ly = (cl // brick[2]) % brick[1]
    lx = cl // (brick[1] * brick[2])
    cell = np.stack([bx * brick[0] + lx, by * brick[1] + ly, bz * brick[2] + lz], 1)
    np.testing.assert_array_equal(base_cell[cell_slots[filled]], cell)


def test_multiply_high_division_constants_are_exact():
    """The brick kernels divide window, cell and node indices by launch constants as
    floor(n / d) = (n * m) >> 32, m = ceil(2^32 / d), for n < 2^16; and (n * m) >> 64,
    m = ceil(2^64 / d), for n < 2^32 (bm_kernels.hip, Divisors)."""
    n = np.arange(65536, dtype=np.uint64)
    for d in list(range(2, 4100)) + [65535, 65536]:
        m = (0xFFFFFFFF // d) + 1
        np.testing.assert_array_equal((n * np.uint64(m)) >> np.uint64(32), n // np.uint64(d))
    rng = np.random.default_rng(0)
    for d in [int(v) for v in rng.integers(2, 2**32 - 1, 500)] + [2, 3, 133926, 2**31, 2**32 - 1]:
        m = (2**64 - 1) // d + 1
        for x in [0, d - 1, d, 2**32 - 1] + [int(v) for v in rng.integers(0, 2**32, 20)]:
            assert (x * m) >> 64 == x // d


@pytest.mark.parametrize("shape, size, spacing", [("block", (0.06, 0.03, 0.024), 0.003),
                                                  ("banana", BANANA_DEMO_SIZE_M, 0.006)])
def test_mls_p2g_groups_cover_every_particle_once(shape, size, spacing):
    """The MLS P2G workgroups hold every particle exactly once, 256 slots per group, and a
    group's particles lie in one brick of 3 x 3 x 3 stencil base cells."""
    from blademaster_hip.mlsmpm import p2g_groups

    ref, _, _ = points(size, spacing, shape, 0)
    origin = ref.min(0) - 0.01
    groups = p2g_groups(ref, origin, spacing)
    assert len(groups) % 256 == 0
    used = groups[groups >= 0]
    assert np.array_equal(np.sort(used), np.arange(len(ref)))
    base = np.floor((ref - origin) / spacing - 0.5).astype(int)
    for g in groups.reshape(-1, 256):
        members = g[g >= 0]
        assert len(members) > 0
        bricks = np.unique((base[members] - base.min(0)) // 3, axis=0)
        assert len(bricks) == 1
