"""Single-field MLS-MPM cutting baseline after TopoCut (`blademaster_warp/mlsmpm.py`).

Updated-Lagrangian MLS-MPM: quadratic B-splines, APIC transfers, one velocity field, J-based
damage (TopoCut Eqs. 6, 7).  Particles, material, time step and the knife/board contact
impulses (the TLMPM contact kernel with cutting off, an articulated tool of zero mobility)
are the BladeMaster ones, so the two solvers differ only in the field representation.
The surface (`MlsSolver.advance/start_damage/diagnostics/state`) follows mlsmpm.py:232-446.
"""

from dataclasses import dataclass
import ctypes
import numpy as np
import torch

from . import _lib
from .arrays import DArray
from .config import derived
from .coupling import Reaction
from .geometry import points


@dataclass(frozen=True)
class TopoCutParams:
    compression: float = 0.025  # epsilon_c, Eq. 6
    stretch: float = 0.01  # epsilon_s
    sensitivity: float = 1.0  # m
    softening: float = 0.0  # gamma in Eq. 7, Pa per unit Hencky plastic strain
    damaged_shear: float = 0.0  # shear modulus factor of damaged particles
    margin: float = 0.06  # domain margin around the material, m
    headroom: float = 0.06  # height above the material, m


def p2g_groups(positions, origin, dx, brick=3, slots=256):
    """P2G workgroups of the particles: grouped by bricks of brick^3 stencil base cells (the
    base cell of cell_base, mlsmpm.py:54-61, at these positions), each brick's particles in
    base-cell order, at most `slots` per group (a brick with more takes several).  Returns
    the particle of each slot, -1 for an empty one, `slots` per group."""
    n = len(positions)
    if n == 0:
        return np.full(slots, -1, dtype=np.int32)
    base = np.floor((np.asarray(positions, np.float64) - origin) / dx - 0.5).astype(np.int64)
    base -= base.min(0)
    span = base.max(0) + 1
    brick_of = base // brick
    bspan = brick_of.max(0) + 1
    brick_id = (brick_of[:, 0] * bspan[1] + brick_of[:, 1]) * bspan[2] + brick_of[:, 2]
    cell_id = (base[:, 0] * span[1] + base[:, 1]) * span[2] + base[:, 2]
    order = np.lexsort((np.arange(n), cell_id, brick_id))
    b_sorted = brick_id[order]
    starts = np.flatnonzero(np.r_[True, b_sorted[1:] != b_sorted[:-1]])
    ends = np.r_[starts[1:], n]
    out = []
    for s0, s1 in zip(starts, ends):
        for g in range(s0, s1, slots):
            chunk = order[g:min(g + slots, s1)]
            out.append(np.pad(chunk.astype(np.int32), (0, slots - len(chunk)), constant_values=-1))
    return np.concatenate(out)


class MlsSolver:
    """damage_from_start=True applies Eq. 6 to J of the reference configuration from the
    first step; False leaves damage off until `start_damage` (mlsmpm.py:232-235)."""

    def __init__(self, cfg, params=TopoCutParams(), device="cuda:0", damage_from_start=True):
        self.config, self.params = cfg, params
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise ValueError("CUDA device required")
        self._lib = _lib.load()
        dx = cfg.grid_spacing
        reference, _, _ = points(cfg.size, dx, cfg.shape, cfg.sampling_seed)
        self.reference = reference
        self.n = n = len(reference)
        d = derived(cfg)
        self.volume, self.mass, self.radius = d["volume"], d["mass"], d["radius"]
        self.blade_offset = d["blade_offset"]
        self.normal_scale, self.tangent_stiffness = d["normal_scale"], d["tangent_stiffness"]
        self.mu, self.lam = d["mu"], d["lam"]
        low = reference.min(0) - params.margin
        high = reference.max(0) + params.margin
        low[2] = -2 * dx
        high[2] = reference[:, 2].max() + params.headroom
        self.origin = low
        self.dims = np.ceil((high - low) / dx).astype(int) + 1
        self.compressed = (1 - params.compression) ** params.sensitivity
        self.stretched = (1 + params.stretch) ** params.sensitivity
        dev = self.device
        f32 = torch.float32

        def z(*shape, dtype=f32):
            return torch.zeros(shape, dtype=dtype, device=dev)

        nodes = int(np.prod(self.dims))
        T = self._t = dict(
            x=torch.tensor(reference, dtype=f32, device=dev),
            v=z(n, 3),
            C=z(n, 3, 3),
            F=torch.eye(3, dtype=f32, device=dev).expand(n, 3, 3).contiguous(),
            yield_stress=torch.full((n,), float(cfg.yield_stress), dtype=f32, device=dev),
            damage=z(n),
            damage_reference=torch.full((n,), 1.0 if damage_from_start else 0.0, dtype=f32, device=dev),
            grid_m=z(nodes),
            grid_p=z(3, nodes),
            grid_v=z(3, nodes),
            tool_pos=z(1, 3),
            tool_vel=z(1, 3),
            tool_rot=torch.tensor([[0.0, 0.0, 0.0, 1.0]], dtype=f32, device=dev),
            tool_omega=z(1, 3),
            tool_mobility=z(6, 6),
            histories=z(2, n, 3),
            contact_impulse=z(n, 3),
            unused_damage=z(n),
            reaction=z(1, 3),
            torque=z(1, 3),
            board=z(1, 3),
            speed=z(1),
            work=z(1),
            impulse=z(1, 3),
            angular_impulse=z(1, 3),
            inversions=z(1, dtype=torch.int32),
            rigid_partial=z(max(1, -(-n // 256)) * 12),
            counters=z(4, dtype=torch.int32),
        )
        groups = p2g_groups(reference, self.origin, dx)
        self.p2g_groups = len(groups) // 256
        T["perm"] = torch.tensor(groups, dtype=torch.int32, device=dev)
        for name in ("x", "v", "C", "F", "damage", "tool_pos", "tool_vel", "tool_rot", "tool_omega", "tool_mobility",
                     "impulse", "angular_impulse", "inversions", "damage_reference", "yield_stress"):
            setattr(self, name, DArray(T[name]))
        self.step_count = 0
        self.graphs = {}
        self._params, self._state = self._contact_structs()
        self._mls_params, self._mls_state = self._mls_structs()

    def _ptr(self, name, ctype):
        return ctypes.cast(ctypes.c_void_p(self._t[name].data_ptr()), ctype)

    def _contact_structs(self):
        """Parameters of the contact kernel: articulated tool of zero mobility, no cutting
        (mlsmpm.py:296-336)."""
        cfg = self.config
        p = _lib.Params()
        p.n = self.n
        p.tool_model = -1
        p.dt, p.spacing, p.mass, p.volume = cfg.dt, cfg.grid_spacing, self.mass, self.volume
        p.radius, p.blade_offset = self.radius, self.blade_offset
        p.sn, p.gk = self.normal_scale, self.tangent_stiffness
        p.mus, p.mud, p.gc = cfg.friction_static, cfg.friction_dynamic, cfg.tangent_damping
        p.board_sn = cfg.board_normal_scale or self.normal_scale
        p.threshold = 1.0
        for i in range(3):
            p.half[i] = cfg.blade_half_size[i]
            p.inertia_inv[i] = 0.0
        p.bevel, p.blade_mass = cfg.blade_bevel_height, cfg.blade_mass
        p.board_half[0], p.board_half[1] = cfg.board_half_extent
        p.rigid_blocks = max(1, -(-self.n // 256))
        s = _lib.State()
        f = _lib.c_f32p
        for field, name in [("x", "x"), ("v", "v"), ("tool_pos", "tool_pos"), ("tool_vel", "tool_vel"),
                            ("tool_rot", "tool_rot"), ("tool_omega", "tool_omega"), ("mobility", "tool_mobility"),
                            ("histories", "histories"), ("damage", "unused_damage"), ("impulses", "contact_impulse"),
                            ("reaction", "reaction"), ("torque", "torque"), ("board", "board"), ("work", "work"),
                            ("speed", "speed"), ("rigid_partial", "rigid_partial")]:
            setattr(s, field, self._ptr(name, f))
        s.counters = self._ptr("counters", ctypes.POINTER(ctypes.c_uint32))
        return p, s

    def _mls_structs(self):
        cfg, prm = self.config, self.params
        m = _lib.MlsParams()
        m.n = self.n
        for i in range(3):
            m.dims[i] = int(self.dims[i])
            m.origin[i] = float(self.origin[i])
        m.plastic = int(cfg.yield_stress > 0)
        m.dx, m.dt, m.mass, m.volume = cfg.grid_spacing, cfg.dt, self.mass, self.volume
        m.mu, m.lam, m.damaged_shear, m.softening = self.mu, self.lam, prm.damaged_shear, prm.softening
        m.compressed, m.stretched = self.compressed, self.stretched
        s = _lib.MlsState()
        f = _lib.c_f32p
        for field, name in [("x", "x"), ("v", "v"), ("C", "C"), ("F", "F"), ("damage", "damage"),
                            ("damage_reference", "damage_reference"), ("yield_stress", "yield_stress"),
                            ("impulse", "contact_impulse"), ("grid_m", "grid_m"), ("grid_p", "grid_p"),
                            ("grid_v", "grid_v"), ("interval_impulse", "impulse"),
                            ("interval_angular", "angular_impulse")]:
            setattr(s, field, self._ptr(name, f))
        s.inversions = self._ptr("inversions", _lib.c_i32p)
        m.groups = self.p2g_groups
        s.perm = self._ptr("perm", _lib.c_i32p)
        return m, s

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _substep(self):
        code = self._lib.bm_mls_substep(ctypes.byref(self._params), ctypes.byref(self._state),
                                        ctypes.byref(self._mls_params), ctypes.byref(self._mls_state), self._stream())
        _lib.check(code, "bm_mls_substep")

    def advance(self, state, duration_s, graph=True, check=True):
        """Hold the knife pose over duration_s (mlsmpm.py:395-426).  check=False skips the
        host-side domain check after the interval (it reads every position back)."""
        state = state.checked()
        dt = self.config.dt
        steps = int(round(duration_s / dt))
        if steps < 1 or not np.isclose(steps * dt, duration_s, rtol=1.0e-10, atol=1.0e-12):
            raise ValueError("Duration must be an integer number of substeps")
        self.tool_pos.assign([state.position])
        self.tool_rot.assign([state.orientation_xyzw])
        self.tool_vel.assign([state.linear_velocity])
        self.tool_omega.assign([state.angular_velocity])
        self.impulse.zero_()
        self.angular_impulse.zero_()
        if graph:
            if steps not in self.graphs:
                g = torch.cuda.CUDAGraph()
                torch.cuda.synchronize(self.device)
                with torch.cuda.graph(g):
                    for _ in range(steps):
                        self._substep()
                self.graphs[steps] = g
            self.graphs[steps].replay()
        else:
            for _ in range(steps):
                self._substep()
        self.step_count += steps
        lost = self.diagnostics()["left_domain"] if check else 0
        if lost:
            raise RuntimeError(
                f"MLS-MPM diverged by t = {self.step_count * dt:.3f} s: {lost} particles non-finite or outside the grid"
            )
        return Reaction(
            duration_s=steps * dt,
            impulse_ns=self.impulse.numpy()[0].astype(float),
            angular_impulse_nms=self.angular_impulse.numpy()[0].astype(float),
            origin_world_m=np.asarray(state.position, dtype=float),
        )

    def start_damage(self):
        """Enable Eq. 6 with each particle's current J as its reference volume (mlsmpm.py:428-430)."""
        self.damage_reference.assign(np.linalg.det(self.F.numpy()))

    def diagnostics(self):
        """left_domain: particles with a non-finite position or a quadratic stencil (nodes
        floor(u - 1/2) .. floor(u - 1/2) + 2 per axis, u the position in cells) not inside the
        grid; clamped_inversions: singular values clamped so far (the counts of
        mlsmpm.py:432-437)."""
        u = (self.x.numpy().astype(np.float64) - self.origin) / self.config.grid_spacing
        with np.errstate(invalid="ignore"):
            first = np.floor(u - 0.5)
            inside = (first >= 0) & (first + 2 <= np.asarray(self.dims) - 1)
        valid = np.isfinite(u).all(axis=1) & inside.all(axis=1)
        return {"left_domain": int(np.count_nonzero(~valid)),
                "clamped_inversions": int(self.inversions.numpy()[0])}

    def state(self):
        return dict(
            position=self.x.numpy(),
            velocity=self.v.numpy(),
            damage=self.damage.numpy(),
            deformation=self.F.numpy(),
            time=self.step_count * self.config.dt,
        )
