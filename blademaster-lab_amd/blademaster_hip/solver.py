"""TLMPM material solver with a rigid blade; every per-substep stage runs in libbm_hip.

Keeps the `Solver` surface of `blademaster_warp/solver.py:157-1051` (constructor, buffers
read by callers, `step`, `advance`, `prepare`, `state`, stage methods) so experiment
drivers call it unchanged.  Device buffers are PyTorch-ROCm tensors wrapped in `DArray`;
each stage is one call into the C-ABI on torch's current stream, and `prepare` records a
batch of substeps into a HIP graph (torch.cuda.CUDAGraph) keyed like the reference's graph
cache (solver.py:995).  The reference's conditional graph nodes are replaced by kernels
that read their device flag (topology_dirty, pair_dirty) and return early.
"""

import ctypes
import time
import numpy as np
import torch

from . import _lib
from .arrays import DArray
from .config import Config, derived, max_cut_index
from .geometry import points, stencil, compact_stencil, bin_bricks, lattice_slots

# Listed pair partners per particle before the bucket fallback (solver.py:152-154).
PAIR_LIST_LANES = 16
# Pair search (solver.py:260-263): Verlet partner lists ("list", "auto"), partner lists rebuilt
# from the cell buckets every step ("hash"), or every pair tested ("all", O(n^2)).
PAIR_MODES = {"auto": 0, "list": 0, "hash": 1, "all": 2}
# "all" is a check of the other searches on small solvers: above this many particles its n^2
# pair tests per step are refused.
ALL_PAIRS_LIMIT = 100_000
CUT_ON, CUT_ADVANCE, CUT_HOLD = 1, 2, 4


def _pow2_at_least(v):
    return 1 << (max(int(v), 16) - 1).bit_length()


class _Device:
    def __init__(self, device):
        self.torch = torch.device(device)
        self.name = str(self.torch)
        self.is_cuda = self.torch.type == "cuda"


class BladeConstraint:
    """Held-blade grid constraint state (blade_constraint.py:213-324); the solve itself runs
    inside the material step of the C-ABI."""

    def __init__(self, solver):
        self.solver = solver
        T = solver._t
        self.side = DArray(T["bc_side"])
        self.impulse = DArray(T["bc_impulse"])
        self.max_violation = DArray(T["bc_sums"][6:7])
        self.residual = DArray(T["bc_sums"][7:8])
        self.batches = DArray(T["bc_counts"][3:4])

    def solve(self):
        self.solver._call("bm_blade_constraint")

    def store_residual(self, result):
        self.solver._call("bm_store_residual", ctypes.c_void_p(result.data_ptr()), int(result.numel()))


class Solver:
    """brick_z: height in cells of the particle bricks of the transfer kernels (a tuning knob;
    results do not depend on it beyond float summation order)."""

    # Relabelled particles listed per substep for the incremental re-registration; more in one
    # substep re-register every particle instead.
    CHANGED_CAPACITY = 65536
    # Movers per step, and displaced particles per full build, of the pair-list refresh (more:
    # every list is rebuilt).  The searches scan the displaced list, so it stays short.
    MOVER_CAPACITY = 256
    PAIR_REFRESH = True

    def __init__(self, config=Config(), device="cuda:0", brick_z=None, _layout=None):
        self.config = cfg = config
        self._brick_z = brick_z
        # _layout (batched instances): reference, origin, cells of the packed domain, the
        # instance of each particle and per-instance Lame constants.
        self.instances = 1 if _layout is None else len(_layout["mu"])
        self.device = _Device(device)
        if not self.device.is_cuda or not torch.cuda.is_available():
            raise ValueError("CUDA device required")
        self.tools = int(cfg.tool_count)
        if _layout is not None and self.tools != 1:
            raise ValueError("Batched instances carry one knife each")
        self._lib = _lib.load()
        dev = self.device.torch
        if _layout is None:
            self.reference, self.origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
        else:
            self.reference, self.origin, cells = _layout["reference"], _layout["origin"], _layout["cells"]
        self.cells = np.asarray(cells)
        self.n = n = len(self.reference)
        if cfg.pair_search == "all" and n > ALL_PAIRS_LIMIT:
            raise ValueError(f"pair_search='all' tests every pair each step: at most {ALL_PAIRS_LIMIT} particles")
        self.node_numpy, self.weight_numpy, self._gradient_numpy = stencil(
            self.reference, self.origin, cfg.grid_spacing, cells
        )
        self.dims = self.cells + 1
        self.node_count = int(np.prod(self.dims))
        if n and (self.node_numpy.min() < 0 or self.node_numpy.max() >= self.node_count):
            raise ValueError("Reference stencil leaves the grid")
        self.max_cut_index = max_cut_index(self.node_count)
        # Sizes kept from the reference for callers that read them (solver.py:177-180).
        self.capacity = _pow2_at_least(16 * n)
        self.pair_capacity = _pow2_at_least(64 * n)
        self.fields = int(cfg.fields_per_node)
        d = derived(cfg)
        self.volume, self.mass, self.radius = d["volume"], d["mass"], d["radius"]
        self.blade_offset = d["blade_offset"]
        self.mu, self.lam = d["mu"], d["lam"]
        self.threshold = d["threshold"]
        self.normal_scale, self.tangent_stiffness = d["normal_scale"], d["tangent_stiffness"]
        self.board_scale = d["board_scale"]
        self.inertia_inv = d["inertia_inv"]
        self.maxwell = d["maxwell"]
        self.mu_maxwell, self.maxwell_relax = d["mu_maxwell"], d["maxwell_relax"]
        if _layout is not None and (self.maxwell or cfg.dynamic_viscosity):
            raise ValueError("Batched instances take neither the Maxwell branch nor dynamic_viscosity")
        self.compact = cfg.grid_storage == "compact"

        f32, i64, i32 = torch.float32, torch.int64, torch.int32

        def t(shape, dtype=f32, value=0):
            return torch.full(shape, value, dtype=dtype, device=dev)

        def arr(tensor, hook=None):
            return DArray(tensor, hook)

        base_node, local, base_cell = compact_stencil(self.reference, self.origin, cfg.grid_spacing, cells)
        # Regular lattice: the transfer kernels use the separable stencil.
        self._lattice = lattice_slots(local) if n else None
        self.regular = self._lattice is not None
        self._bin_bricks(base_cell)
        eye = torch.eye(3, dtype=f32, device=dev).expand(n, 3, 3).contiguous()
        self._t = {}
        T = self._t
        T["x"] = torch.tensor(self.reference, dtype=f32, device=dev).reshape(n, 3)
        T["ref"] = T["x"].clone()
        T["v"] = t((n, 3))
        T["F"] = eye.clone()
        T["Fp"] = eye.clone()
        T["stress"] = t((n, 3, 3))
        T["damage"] = t((n,))
        T["codes"] = t((n,), i64)
        T["single_codes"] = t((n,), i64)
        T["local"] = torch.tensor(local, device=dev).reshape(n, 3)
        T["base_node"] = torch.tensor(base_node, device=dev)
        T["field_idx"] = t((n, 8), torch.uint8)
        T["impulses"] = t((n, 3))
        T["histories"] = t((2, n, 3))
        # Rows of 64-bit words, one bit per particle: its friction history may be nonzero.
        T["history_mask"] = t((2, max(1, -(-n // 64))), i64)
        F_, N_ = self.fields, self.node_count
        T["field_code"] = t((F_, N_), i64, -1)
        T["node_nf"] = t((N_,), i32)
        # Multi-field node list of the node kernels (appended by P2G, emptied every substep).
        T["multi_nodes"] = t((N_,), i32)
        T["multi_stamp"] = t((N_,), i32)
        T["multi_count"] = torch.tensor([0, 1], dtype=i32, device=dev)
        T["grid_mass"] = t((F_, N_))
        for name in ("momentum", "grid_impulse", "old_v", "new_v", "remapped", "grid_vel"):
            T[name] = t((F_, N_, 3))
        hz = cfg.blade_half_size[2]
        I = self.instances
        # One row per batched instance, or per knife of a single solver (knife k > 0 is an
        # extra free knife, contact.py:500-566).
        K = self.tools if _layout is None else I
        T["tool_pos"] = torch.tensor([[0.0, 0.0, cfg.size[2] + 0.012 + hz]] * K, dtype=f32, device=dev)
        T["tool_vel"] = t((K, 3))
        T["tool_rot"] = torch.tensor([[0.0, 0.0, 0.0, 1.0]] * K, dtype=f32, device=dev)
        T["tool_omega"] = t((K, 3))
        T["tool_mobility"] = t((6, 6)) if I == 1 else t((I, 6, 6))
        T["reaction"] = t((K, 3))
        T["torque"] = t((K, 3))
        T["board"] = t((I, 3))
        T["work"] = t((I,))
        T["speed"] = t((1,))
        T["table"] = t((K, 3))
        T["interval_impulse"] = t((I, 3))
        T["interval_angular"] = t((I, 3))
        if _layout is not None:
            T["instance"] = torch.tensor(np.asarray(_layout["instance"], dtype=np.int32), device=dev)
            T["inst_mu"] = torch.tensor(np.asarray(_layout["mu"], dtype=np.float32), device=dev)
            T["inst_lam"] = torch.tensor(np.asarray(_layout["lam"], dtype=np.float32), device=dev)
            T["inst_sums"] = t((I, 16, 10))
        if self.maxwell:
            # Fv - I of the Maxwell branch in float64 (material.py:485-535).
            T["viscous"] = t((n, 3, 3), torch.float64)
        if cfg.dynamic_viscosity:
            T["viscous_impulse"] = t((F_, N_, 3))
        T["errors"] = t((3,), i32)
        T["topology_dirty"] = t((1,), i32, 1)
        T["fp_identity"] = t((1,), i32, 1)
        T["update_pending"] = t((1,), i32)
        T["update_done"] = t((max(n, 1),), i32)
        T["field_idx_prev"] = t((max(n, 1), 8), torch.uint8)
        # pair contacts
        P_ = self.pair_capacity
        self.list_buckets = _pow2_at_least(2 * n)
        if _layout is not None:
            # One bucket region per instance (a power of two of at least twice its particles),
            # so each instance's pair list is binned and rebuilt on its own.
            counts = np.bincount(np.asarray(_layout["instance"]), minlength=I)
            sizes = np.array([_pow2_at_least(2 * max(int(c), 1)) for c in counts], dtype=np.int64)
            bases = np.concatenate([[0], np.cumsum(sizes)[:-1]])
            self.list_buckets = int(sizes.sum())
            T["inst_buckets"] = torch.tensor(np.stack([bases, sizes], 1).astype(np.int32), device=dev)
            T["inst_pair_builds"] = t((I,), i32)
            T["inst_pair_dirty"] = t((I,), i32, 1)
        T["pair_keys"] = t((2, P_), i64, -1)
        T["pair_history"] = t((2, P_, 3))
        T["pair_used"] = t((2, P_), i32)
        T["pair_count"] = t((2,), i32)
        T["pair_dirty"] = t((1,), i32, 1)
        T["pair_builds"] = t((1,), i32)
        T["list_reach"] = t((1,))
        T["built_x"] = t((n, 3))
        T["built_codes"] = t((n,), i64)
        T["partners"] = t((max(n, 1) * PAIR_LIST_LANES,), i32)
        T["partner_count"] = t((n,), i32)
        # Bucket heads carry the build count in their upper 32 bits; a head of an older build
        # reads as an empty bucket, so the lists need no clearing pass.
        T["list_heads"] = t((self.list_buckets,), i64)
        T["list_links"] = t((max(n, 1),), i32, -1)
        self.changed_capacity = int(self.CHANGED_CAPACITY)
        T["changed"] = t((self.changed_capacity,), i32)
        # Pair-list refresh of the particles that moved too far (one solver, list mode, at least
        # 2^17 buckets so that a search cube's cells have distinct buckets): the other lists keep
        # their build.  Otherwise every list is rebuilt.
        refresh = (self.PAIR_REFRESH and not _layout and cfg.pair_search in ("list", "auto")
                   and self.list_buckets >= 1 << 17 and PAIR_LIST_LANES <= 16)
        self.mover_capacity = int(self.MOVER_CAPACITY) if refresh else 0
        T["movers"] = t((max(self.mover_capacity, 1),), i32)
        T["displaced"] = t((max(self.mover_capacity, 1),), i32)
        T["refresh_mark"] = t((max(n, 1),), i32)
        T["refresh_state"] = t((4,), i32)
        T["changed_count"] = t((1,), i32)
        T["pair_mark"] = t((max(n, 1),), i32)
        T["pair_epoch"] = t((1,), i32)
        T["perm"] = torch.tensor(self.brick_slots, dtype=i32, device=dev)
        # Bricks with particles and nodes in some particle's stencil never change in TLMPM; the
        # kernels walk only these when they are a strict subset (sparse shapes, packed batches).
        nb_all = int(np.prod(self.bricks))
        occupied = np.flatnonzero((self.brick_slots.reshape(nb_all, -1) >= 0).any(axis=1)).astype(np.int32)
        self.active_bricks = len(occupied) if len(occupied) < nb_all else 0
        if self.active_bricks:
            T["brick_list"] = torch.tensor(occupied, device=dev)
        used_nodes = np.unique(self.node_numpy).astype(np.int32) if n else np.zeros(0, np.int32)
        self.active_nodes = len(used_nodes) if 0 < len(used_nodes) < self.node_count else 0
        if self.active_nodes:
            T["node_list"] = torch.tensor(used_nodes, device=dev)
        window = int(np.prod(np.asarray(self.brick) + 1))
        nb = int(np.prod(self.bricks))
        T["p2g_partial"] = t((max(nb * F_ * window * 7, 1),))
        T["remap_partial"] = t((max(nb * F_ * window * 3, 1),))
        # Rows of contact sums: one per contact-kernel workgroup, then 16 the pair kernel reduces
        # them to.
        self.rigid_blocks = max(1, -(-n // 256)) + 16
        T["rigid_partial"] = t((self.rigid_blocks * 12,))
        if self.tools > 1:
            # Knives 1 .. K-1: friction history row, damage and per-block contact sums of each.
            T["extra_histories"] = t((self.tools - 1, n, 3))
            T["extra_damage"] = t((self.tools - 1, n))
            T["extra_partial"] = t(((self.tools - 1) * self.rigid_blocks * 8,))
        T["counters"] = t((4,), i32)
        # Held-blade grid constraint (blade_constraint.py): allocated for an externally held tool.
        self.blade_hold = bool(cfg.external_tool and cfg.tool_substep == "hold")
        bn = n if self.blade_hold else 1
        bN = N_ if self.blade_hold else 1
        T["bc_side"] = t((bn,))
        T["bc_normal"] = t((bn, 3))
        T["bc_lever"] = t((bn, 3))
        for name in ("bc_rhs", "bc_bound", "bc_impulse", "bc_extrapolated"):
            T[name] = t((bn,))
        T["bc_active"] = t((bn,), i32)
        T["bc_nodes"] = t((bN,), i32)
        T["bc_node_mark"] = t((bN,), i32)
        T["bc_grid"] = t((F_ if self.blade_hold else 1, bN, 3))
        T["bc_coupled"] = t((F_ if self.blade_hold else 1, bN, 3))
        T["bc_counts"] = t((4,), i32)
        T["bc_sums"] = t((8,))
        T["single_target"] = t((1, 3))
        T["single_rotation"] = torch.tensor([[0.0, 0.0, 0.0, 1.0]], dtype=f32, device=dev)

        mark_topology = self._mark_topology
        mark_pairs = self._mark_pairs
        self.x, self.v = arr(T["x"], mark_pairs), arr(T["v"], mark_pairs)
        self.ref = arr(T["ref"])
        self.F, self.stress = arr(T["F"]), arr(T["stress"])
        self.viscous = arr(T["viscous"]) if self.maxwell else None
        # Host writes to Fp end the device's identity shortcut for Fp.
        self.Fp = arr(T["Fp"], lambda: T["fp_identity"].fill_(0))
        self.codes = arr(T["codes"], mark_topology)
        self.single_field_codes = arr(T["single_codes"], mark_topology)
        self.damage = arr(T["damage"])
        # Damage of each particle against knife k (knife 0: `damage`).
        self.tool_damage = [self.damage] + [arr(T["extra_damage"][k]) for k in range(self.tools - 1)]
        self.extra_histories = arr(T["extra_histories"]) if self.tools > 1 else None
        self.impulses = arr(T["impulses"])
        # Host writes to the histories: every particle's history may now be nonzero.
        self.histories = arr(T["histories"], lambda: T["history_mask"].fill_(-1))
        self.tool_pos, self.tool_vel = arr(T["tool_pos"]), arr(T["tool_vel"])
        self.tool_rot, self.tool_omega = arr(T["tool_rot"]), arr(T["tool_omega"])
        self.tool_mobility = arr(T["tool_mobility"])
        self.reaction, self.torque, self.board = arr(T["reaction"]), arr(T["torque"]), arr(T["board"])
        self.work, self.speed = arr(T["work"]), arr(T["speed"])
        self.tool_table_impulse = arr(T["table"])
        self.interval_impulse = arr(T["interval_impulse"])
        self.interval_angular_impulse = arr(T["interval_angular"])
        self.errors = arr(T["errors"])
        self.topology_dirty = arr(T["topology_dirty"])
        self.pair_count = [arr(T["pair_count"][0:1]), arr(T["pair_count"][1:2])]
        self.pair_list_builds = arr(T["pair_builds"])
        self.pair_list_dirty = arr(T["pair_dirty"])
        self.grid_mass = arr(T["grid_mass"])
        self.new_v = arr(T["new_v"])
        self.old_v = arr(T["old_v"])
        self.single_target = arr(T["single_target"])
        self.single_rotation = arr(T["single_rotation"])
        self.pair_skin = cfg.pair_skin or self.radius
        self.pair_reach = 2 * self.radius + self.pair_skin
        self._single_field = False
        self.step_count = 0
        self.cut_index = 0
        self._batches = {}
        self.graph_build_s = 0.0
        # Knife 0 uses the full arrays; knife k > 0 row views (solver.py:288-300).
        self.tool_views = [
            dict(pos=self.tool_pos, vel=self.tool_vel, rot=self.tool_rot, omega=self.tool_omega,
                 reaction=self.reaction, torque=self.torque, table=self.tool_table_impulse)
        ] + [
            dict(pos=self.tool_pos[k:k + 1], vel=self.tool_vel[k:k + 1], rot=self.tool_rot[k:k + 1],
                 omega=self.tool_omega[k:k + 1], reaction=self.reaction[k:k + 1], torque=self.torque[k:k + 1],
                 table=self.tool_table_impulse[k:k + 1])
            for k in range(1, self.tools if _layout is None else 1)
        ]
        self._params = self._make_params()
        self._state = self._make_state()
        self.blade_constraint = BladeConstraint(self) if self.blade_hold else None
        self.refresh_stress()

    # ------------------------------------------------------------------ C-ABI plumbing
    def _make_params(self):
        cfg = self.config
        p = _lib.Params()
        p.n = self.n
        p.node_count = self.node_count
        p.fields = self.fields
        for i in range(3):
            p.dims[i] = int(self.dims[i])
            p.half[i] = cfg.blade_half_size[i]
            p.inertia_inv[i] = float(self.inertia_inv[i])
        p.pair_capacity = self.pair_capacity
        p.list_buckets = self.list_buckets
        p.list_lanes = PAIR_LIST_LANES
        p.tool_model = self._tool_model()
        p.free_blade = int(cfg.free_blade)
        p.lock_rotation = int(not (cfg.free_blade or cfg.pose_servo))
        p.predict = int(cfg.tool_substep == "predict")
        p.plastic = int(cfg.yield_stress > 0)
        p.dt, p.spacing, p.mass, p.volume = cfg.dt, cfg.grid_spacing, self.mass, self.volume
        p.radius, p.blade_offset, p.mu, p.lam = self.radius, self.blade_offset, self.mu, self.lam
        p.yield_stress, p.flip = cfg.yield_stress, cfg.flip_fraction
        p.sn, p.gk = self.normal_scale, self.tangent_stiffness
        p.mus, p.mud, p.gc = cfg.friction_static, cfg.friction_dynamic, cfg.tangent_damping
        p.board_sn, p.threshold = self.board_scale, self.threshold
        p.bevel, p.blade_mass = cfg.blade_bevel_height, cfg.blade_mass
        p.board_half[0], p.board_half[1] = cfg.board_half_extent
        p.drive_k, p.drive_c, p.drive_limit = cfg.drive_stiffness, cfg.drive_damping, cfg.drive_force_limit
        p.rot_k, p.rot_c = cfg.rotational_stiffness, cfg.rotational_damping
        p.pair_skin = self.pair_skin
        for i in range(3):
            p.brick[i] = int(self.brick[i])
            p.bricks[i] = int(self.bricks[i])
        p.rigid_blocks = self.rigid_blocks
        p.blade_hold = int(self.blade_hold)
        p.changed_capacity = self.changed_capacity
        p.instances = getattr(self, "instances", 1)
        p.active_bricks = self.active_bricks
        p.active_nodes = self.active_nodes
        p.regular = int(self.regular)
        p.pair_mode = PAIR_MODES[cfg.pair_search]
        p.maxwell = int(self.maxwell)
        p.tools = self.tools
        p.mu_maxwell, p.maxwell_relax = self.mu_maxwell, self.maxwell_relax
        p.viscosity = cfg.dynamic_viscosity
        p.mover_capacity = self.mover_capacity
        return p

    def _make_state(self):
        s = _lib.State()
        T = self._t
        for field, ctype in _lib.STATE_FIELDS:
            key = field
            if field == "mobility":
                key = "tool_mobility"
            elif field == "transfer_codes":
                key = "single_codes" if self._single_field_flag() else "codes"
            tensor = T.get(key)
            if tensor is None:
                continue  # absent buffers stay NULL (e.g. the batched-instance arrays)
            setattr(s, field, ctypes.cast(ctypes.c_void_p(tensor.data_ptr()), ctype))
        return s

    def _single_field_flag(self):
        return getattr(self, "_single_field", False)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device.torch).cuda_stream)

    def _call(self, name, *args):
        code = getattr(self._lib, name)(ctypes.byref(self._params), ctypes.byref(self._state), *args, self._stream())
        _lib.check(code, name)

    def _mark_topology(self):
        self._t["topology_dirty"].fill_(1)
        self._t["pair_dirty"].fill_(1)

    def _mark_pairs(self):
        # Host writes to x or v: rebuild the pair list before the next contact pass.
        self._t["pair_dirty"].fill_(1)

    def _bin_bricks(self, base_cell):
        kw = {} if self._brick_z is None else dict(brick_z=int(self._brick_z))
        self.brick, self.bricks, self.brick_slots = bin_bricks(self.cells, base_cell, int(self.config.fields_per_node),
                                                               lattice=self._lattice, **kw)

    def launch(self, name, *args):
        """Launch one C-ABI stage by name with the solver's params/state (cf. solver.py:331)."""
        self._call(name, *args)

    # ------------------------------------------------------------------ stages
    def refresh_stress(self):
        """Recompute the stored stress after host writes to F, Fp or viscous (solver.py:318-329)."""
        self._call("bm_refresh_stress")

    def material_step(self):
        """Registration when dirty, P2G, exchange, G2P + MUSL, F update (solver.py:334-376)."""
        self._call("bm_material_step")

    @property
    def single_field(self):
        return self._single_field

    def set_single_field(self, enabled):
        """Paper Fig. 4 baseline: grid transfers ignore the cut codes (solver.py:382-389)."""
        if bool(enabled) != self._single_field:
            self._single_field = bool(enabled)
            self._state = self._make_state()
            self._mark_topology()

    def _register_topology(self):
        self._mark_topology()
        self._call("bm_register")

    def _tool_model(self):
        """-1 articulated (external), 0 guided translation, 1 free rigid blade (solver.py:571-576)."""
        cfg = self.config
        if cfg.external_tool:
            return -1
        return int(cfg.free_blade or cfg.pose_servo)

    def _cut_bits(self, cutting):
        if not cutting:
            return 0
        cfg = self.config
        return CUT_ON | (CUT_ADVANCE if cfg.cut_work == "advance" else 0) | (CUT_HOLD if cfg.cut_hold else 0)

    def rigid_contacts(self, cutting=True):
        """Gravity plus board and blade contact impulses on every particle (solver.py:588-630)."""
        self._call("bm_rigid_contacts", self._cut_bits(cutting))

    def particle_contacts(self):
        """Contacts between particles on incompatible sides of a cut (solver.py:670-751)."""
        self._call("bm_pair_contacts", self.step_count % 2)

    def assign_sides(self, cut_index, cutting=True):
        """Side digit 3^(cut_index + k) for particles crossed by knife k's edge (solver.py:866-890)."""
        self._call("bm_assign", 3**cut_index, int(cutting))

    def step(self, target_edge=(0.0, 0.0, 0.042), cutting=True, cut_index=0, orientation=None):
        if self.config.external_tool:
            raise ValueError("Use ExternalToolPort.advance for an externally owned tool")
        cut_index = self._set_cut(cut_index)
        self._step_kernels(target_edge, cutting, cut_index, orientation=orientation)

    def _set_cut(self, cut_index):
        if isinstance(cut_index, (bool, np.bool_)) or not isinstance(cut_index, (int, np.integer)):
            raise ValueError("Cut index must be an integer")
        cut_index = int(cut_index)
        last = self.max_cut_index - (self.config.tool_count - 1)
        if cut_index < self.cut_index or cut_index > last:
            raise ValueError(f"Cut indices must be monotonic and within [0, {last}]")
        if cut_index != self.cut_index:
            for damage in self.tool_damage:
                damage.zero_()
            self.cut_index = cut_index
        return cut_index

    def _drive_tool(self, target_edge, orientation, buffer, rotations, index):
        """Servo, free or guided knife step with the shared board contact (solver.py:824-864)."""
        if buffer is None:
            self.single_target.assign([target_edge])
            self.single_rotation.assign([orientation if orientation is not None else [0.0, 0.0, 0.0, 1.0]])
            buffer, rotations, index = self._t["single_target"], self._t["single_rotation"], 0
        self._call("bm_tool_step", ctypes.c_void_p(buffer.data_ptr()), ctypes.c_void_p(rotations.data_ptr()), index)

    def _step_kernels(self, target_edge, cutting, cut_index, buffer=None, index=0, rotations=None, orientation=None,
                      defer=False):
        """One substep: contacts, material update, tool update, side labels (solver.py:920-940).

        defer: leave this substep's F update to the next substep's P2G when no labels changed
        (`bm_substep_deferred`); the next call must follow with nothing reading F in between.
        """
        cfg = self.config
        if buffer is None and not cfg.external_tool:
            self.single_target.assign([target_edge])
            self.single_rotation.assign([orientation if orientation is not None else [0.0, 0.0, 0.0, 1.0]])
            buffer, rotations, index = self._t["single_target"], self._t["single_rotation"], 0
        tptr = ctypes.c_void_p(buffer.data_ptr()) if buffer is not None else ctypes.c_void_p(
            self._t["single_target"].data_ptr())
        rptr = ctypes.c_void_p(rotations.data_ptr()) if rotations is not None else ctypes.c_void_p(
            self._t["single_rotation"].data_ptr())
        self._call(
            "bm_substep_deferred" if defer else "bm_substep",
            self._cut_bits(cutting),
            int(bool(cutting)),
            3**cut_index,
            self.step_count % 2,
            int(cfg.accumulate_tool_impulses),
            tptr,
            rptr,
            int(index),
        )
        self.step_count += 1

    def advance(self, targets, cutting=True, cut_index=0, graph=True, orientations=None):
        """Submit exact per-substep edge targets (solver.py:942-973)."""
        if self.config.external_tool:
            raise ValueError("Use ExternalToolPort.advance for an externally owned tool")
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        if targets.ndim != 2 or targets.shape[1] != 3 or not np.isfinite(targets).all() or not len(targets):
            raise ValueError("targets must be finite (substeps, 3) edge positions")
        orientations = (
            np.tile([0.0, 0.0, 0.0, 1.0], (len(targets), 1)) if orientations is None else np.asarray(orientations)
        )
        if orientations.shape != (len(targets), 4) or not np.isfinite(orientations).all():
            raise ValueError("orientations must be finite (substeps, 4) xyzw quaternions")
        lengths = np.linalg.norm(orientations, axis=1)
        if np.any(lengths < 1.0e-6):
            raise ValueError("Zero quaternion")
        orientations = np.ascontiguousarray(orientations / lengths[:, None], dtype=np.float32)
        cut_index = self._set_cut(cut_index)
        if not graph:
            for target, orientation in zip(targets, orientations):
                self._step_kernels(target, cutting, cut_index, orientation=orientation)
            return
        buffer, rotations, captured = self.prepare(targets, cutting, cut_index, orientations)
        buffer.copy_(torch.from_numpy(targets).to(buffer.device))
        rotations.copy_(torch.from_numpy(orientations).to(rotations.device))
        captured.replay()
        self.step_count += len(targets)

    def prepare(self, targets, cutting=True, cut_index=0, orientations=None, before=(), after=(), tag=None,
                after_step=None):
        """Record a batch of substeps into a HIP graph without advancing state (solver.py:975-1019).

        `before` and `after` are callables run around the substeps during capture (each
        issues C-ABI launches on the current stream); `tag` names that wrapping in the cache
        key.
        """
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        if orientations is None:
            orientations = np.tile([0.0, 0.0, 0.0, 1.0], (len(targets), 1))
        key = (len(targets), bool(cutting), cut_index, self.step_count % 2, self._single_field, tag)
        if key not in self._batches:
            start = time.perf_counter()
            dev = self.device.torch
            buffer = torch.tensor(targets, dtype=torch.float32, device=dev)
            rotations = torch.tensor(np.asarray(orientations, dtype=np.float32), device=dev)
            count = self.step_count
            graph = torch.cuda.CUDAGraph()
            torch.cuda.synchronize(dev)
            try:
                with torch.cuda.graph(graph):
                    for fn in before:
                        fn()
                    # Inside the recorded batch each substep's F update may wait for the next
                    # substep's P2G; the last substep completes its own.
                    for i in range(len(targets)):
                        self._step_kernels(targets[i], cutting, cut_index, buffer, i, rotations,
                                           defer=i + 1 < len(targets))
                        if after_step is not None:
                            after_step(i)
                    for fn in after:
                        fn()
            finally:
                self.step_count = count
            self._batches[key] = buffer, rotations, graph
            self.graph_build_s += time.perf_counter() - start
        return self._batches[key]

    def release_graphs(self, owner):
        """Drop the cached graphs recorded for `owner` (a port's tag prefix, e.g.
        ("external", serial)): their launches hold raw pointers into that port's buffers."""
        owner = tuple(owner)
        for key in [k for k in self._batches if isinstance(k[-1], tuple) and k[-1][: len(owner)] == owner]:
            del self._batches[key]

    # ------------------------------------------------------------------ readback
    def state(self):
        """Host copies of the particle and blade state; raises on device errors (solver.py:1021-1051)."""
        errors = self.errors.numpy()
        if errors.any():
            raise RuntimeError(
                f"GPU errors [storage overflow, inverted material, coincident proxies]: {errors.tolist()}"
            )
        state = {
            "position": self.x.numpy(),
            "velocity": self.v.numpy(),
            "deformation": self.F.numpy(),
            "plastic_deformation": self.Fp.numpy(),
            "sides": self.codes.numpy(),
            "damage": self.damage.numpy(),
            "blade_position": self.tool_pos.numpy(),
            "blade_velocity": self.tool_vel.numpy(),
            "blade_orientation": self.tool_rot.numpy(),
            "blade_omega": self.tool_omega.numpy(),
        }
        if self.maxwell:
            state["viscous_deformation"] = np.eye(3) + self.viscous.numpy()
        if self.tools > 1:
            state["tool_damage"] = np.stack([damage.numpy() for damage in self.tool_damage])
        state["time"] = self.step_count * self.config.dt
        if any(not np.isfinite(v).all() for v in state.values()):
            raise RuntimeError("Nonfinite state; run is invalid")
        if np.linalg.det(state["deformation"]).min() <= 0:
            raise RuntimeError("Inverted material; run is invalid")
        return state

    def embedded_particles(self):
        """Particles beside the blade body (solver.py:892-900)."""
        count = torch.zeros(1, dtype=torch.int32, device=self.device.torch)
        self._call("bm_count_embedded", ctypes.c_void_p(count.data_ptr()))
        return int(count.item())

    def embedded_particles_previous(self):
        """`embedded_particles` of the previous call; this call's count is copied back
        asynchronously into pinned memory (solver.py:902-918).  The next synchronising call
        (an `advance`, a readback) completes the copy before it is read."""
        if not hasattr(self, "_embedded"):
            self._embedded = torch.zeros(1, dtype=torch.int32, device=self.device.torch)
            self._embedded_host = torch.zeros(1, dtype=torch.int32, pin_memory=True)
        previous = int(self._embedded_host[0])
        self._embedded.zero_()
        self._call("bm_count_embedded", ctypes.c_void_p(self._embedded.data_ptr()))
        self._embedded_host.copy_(self._embedded, non_blocking=True)
        return previous

    def blade_patch(self, out=None):
        """Particle count and summed damage in 8 x 8 x 8 cells around each knife edge (the
        surrogate's patch): a [instances][2][8][8][8] device tensor, filled asynchronously."""
        if out is None:
            out = torch.empty(self.instances, 2, 8, 8, 8, dtype=torch.float32, device=self.device.torch)
        if out.numel() != self.instances * 1024 or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError("Patch buffer must be a contiguous float32 tensor of instances x 1024 values")
        self._call("bm_blade_patch", ctypes.c_void_p(out.data_ptr()))
        return out

    def snapshot(self):
        """Device copy of the whole simulation state (every buffer and the host counters);
        `restore` rewinds to it.  Buffers are copied in place, so captured graphs stay valid."""
        return dict(tensors={k: t.clone() for k, t in self._t.items()}, step_count=self.step_count,
                    cut_index=self.cut_index)

    def restore(self, snap):
        for k, t in self._t.items():
            t.copy_(snap["tensors"][k])
        self.step_count = snap["step_count"]
        self.cut_index = snap["cut_index"]

    def piece_labels(self):
        """Piece labels of the current codes on the device, equal bit for bit to
        `components(node_numpy, weight_numpy, codes)` (topology.py:17-57): labels (n,)
        ranked by the lowest particle index of each piece, and the piece sizes."""
        dev = self.device.torch
        n, N = self.n, self.node_count
        if n == 0:
            return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
        # The reference has no cap on codes per node: the scratch table grows until it fits.
        F = getattr(self, "_label_fields", 16)
        while True:
            B = self._label_scratch(F)
            B["errors"].zero_()
            L = _lib.Labels()
            L.fields = F
            for field, ctype in _lib.LABELS_FIELDS[1:]:
                setattr(L, field, ctypes.cast(ctypes.c_void_p(B[field].data_ptr()), ctype))
            self._call("bm_piece_labels", ctypes.byref(L))
            if int(B["errors"].cpu()[0]) == 0:
                break
            if F >= 255:
                raise RuntimeError("Piece labels: a grid node holds more than 255 cut codes")
            F = min(255, 2 * F)
        self._label_fields = F
        labels = B["labels"].cpu().numpy().astype(np.int64)
        return labels, np.bincount(labels, minlength=int(B["count"].cpu()[0]))

    def _label_scratch(self, F):
        B = getattr(self, "_label_buffers", None)
        if B is not None and B["field_code"].shape[0] == F:
            return B
        dev = self.device.torch
        n, N = self.n, self.node_count
        bits = (np.asarray(self.weight_numpy) > 0).astype(np.uint8) << np.arange(8, dtype=np.uint8)
        z = lambda *shape, dtype=torch.int32: torch.zeros(shape, dtype=dtype, device=dev)  # noqa: E731
        self._label_buffers = dict(
            corner_mask=torch.tensor(bits.sum(axis=1).astype(np.uint8), device=dev),
            field_code=z(F, N, dtype=torch.int64), node_nf=z(N), field_idx=z(n, 8, dtype=torch.uint8),
            parent=z(n + F * N), root=z(n), block_sum=z(-(-n // 1024)), labels=z(n), count=z(1), errors=z(1),
        )
        return self._label_buffers

    # ------------------------------------------------------------------ field-table views
    @property
    def weights(self):
        return DArray(torch.tensor(self.weight_numpy, dtype=torch.float32, device=self.device.torch))

    @property
    def nodes(self):
        return DArray(torch.tensor(self.node_numpy, dtype=torch.int32, device=self.device.torch))

    @property
    def gradients(self):
        return DArray(torch.tensor(self._gradient_numpy, dtype=torch.float32, device=self.device.torch))

    def field_table(self):
        """(codes [F, N] int64 with -1 for empty slots, used slots per node [N])."""
        return self._t["field_code"].cpu().numpy(), self._t["node_nf"].cpu().numpy()

    @property
    def active_count(self):
        _, nf = self.field_table()
        return DArray(torch.tensor([int(nf.sum())], dtype=torch.int32))

    @property
    def grid_keys(self):
        """Registered field keys code * node_count + node, ordered by slot then node."""
        table, _ = self.field_table()
        f, node = np.nonzero(table >= 0)
        keys = table[f, node] * self.node_count + node
        return DArray(torch.tensor(keys, dtype=torch.int64))

    @property
    def keys(self):
        return self.grid_keys

    @property
    def slots(self):
        """Global field slot (f * node_count + node) of every (particle, corner)."""
        idx = self._t["field_idx"].cpu().numpy().astype(np.int64)
        return DArray(torch.tensor(idx * self.node_count + self.node_numpy, dtype=torch.int64))

    def corner_keys(self):
        """Key code * node_count + node of the field each (particle, corner) transfers into."""
        table, _ = self.field_table()
        idx = self._t["field_idx"].cpu().numpy().astype(np.int64)
        return table[idx, self.node_numpy] * self.node_count + self.node_numpy
