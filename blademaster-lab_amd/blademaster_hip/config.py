"""Solver configuration and the per-run constants derived from it.

`Config` keeps the field names, defaults and `ValueError` messages of
`blademaster_warp/solver.py:12-149` so that callers construct it unchanged.  One field is
new:

* ``fields_per_node``: field slots per grid node in the node-major field table that replaces
  the reference's open-addressing key table (`material.py:37-75`).  Exceeding it is reported
  through the same storage-overflow counter (``errors[0]``) the reference uses for a full table.
"""

from dataclasses import dataclass, asdict
import numpy as np


@dataclass(frozen=True)
class Config:
    size: tuple = (0.06, 0.03, 0.03)
    grid_spacing: float = 0.0075
    dt: float = 0.0005
    density: float = 1100.0
    young_modulus: float = 6000.0
    poisson_ratio: float = 0.35
    dynamic_viscosity: float = 0.0
    maxwell_tau: float = 0.0
    maxwell_fraction: float = 0.0
    yield_stress: float = 0.0
    flip_fraction: float = 0.95
    normal_scale: float = 0.1
    friction_static: float = 0.3
    friction_dynamic: float = 0.2
    tangent_stiffness: float = 0.05
    tangent_damping: float = 0.1
    contact_reference_dt: float = 0.0
    cutting_energy: float = 0.5
    sharpness: float = 0.9
    cut_work: str = "relative"
    cut_hold: bool = False
    blade_contact_radius: float = 0.0
    blade_mass: float = 0.2
    blade_half_size: tuple = (0.0005, 0.025, 0.02)
    drive_stiffness: float = 250.0
    drive_damping: float = 14.0
    drive_force_limit: float = 10.0
    shape: str = "block"
    sampling_seed: int = 0
    blade_bevel_height: float = 0.0
    free_blade: bool = False
    pose_servo: bool = False
    external_tool: bool = False
    cache_topology: bool = False
    grid_storage: str = "hash"
    pair_search: str = "auto"
    pair_skin: float = 0.0
    tool_substep: str = "hold"
    accumulate_tool_impulses: bool = False
    rotational_stiffness: float = 0.05
    rotational_damping: float = 0.003
    board_half_extent: tuple = (0.0, 0.0)
    board_normal_scale: float = 0.0
    tool_count: int = 1
    fields_per_node: int = 4

    def __post_init__(self):
        # Same checks and messages as `solver.py:87-144`.
        if self.grid_storage not in ("hash", "compact"):
            raise ValueError("grid_storage must be hash or compact")
        if self.pair_search not in ("auto", "hash", "all", "list"):
            raise ValueError("pair_search must be auto, hash, all or list")
        if self.pair_skin < 0:
            raise ValueError("pair_skin must be non-negative")
        if self.blade_contact_radius < 0:
            raise ValueError("blade_contact_radius must be non-negative")
        if self.cut_work not in ("relative", "advance"):
            raise ValueError("cut_work must be relative or advance")
        if self.tool_substep not in ("hold", "predict"):
            raise ValueError("tool_substep must be hold or predict")
        if self.tool_substep == "predict" and not self.external_tool:
            raise ValueError("tool_substep=predict requires external_tool")
        numeric = [float(v) for v in asdict(self).values() if isinstance(v, (int, float))]
        if not np.isfinite(numeric).all():
            raise ValueError("Nonfinite configuration")
        positive = (
            "dt",
            "grid_spacing",
            "density",
            "young_modulus",
            "normal_scale",
            "tangent_stiffness",
            "blade_mass",
            "cutting_energy",
            "drive_force_limit",
        )
        if any(getattr(self, k) <= 0 for k in positive):
            raise ValueError("Invalid positive parameter")
        if not (0 < self.poisson_ratio < 0.5 and 0 < self.sharpness < 1):
            raise ValueError("Invalid material/sharpness")
        if not (0 <= self.flip_fraction <= 1 and 0 <= self.friction_dynamic <= self.friction_static):
            raise ValueError("Invalid transfer/friction")
        if min(self.dynamic_viscosity, self.yield_stress, self.tangent_damping, self.contact_reference_dt) < 0:
            raise ValueError("Negative damping/plasticity/reference step")
        if self.maxwell_tau < 0 or not 0 <= self.maxwell_fraction <= 1:
            raise ValueError("maxwell_tau must be non-negative and maxwell_fraction in [0, 1]")
        if max(self.contact_scales()) > 1:
            raise ValueError("Rescaled contact factor exceeds 1 (one-step overshoot); reduce dt")
        if not 0 <= self.board_normal_scale <= 1:
            raise ValueError("Board normal scale must be in [0, 1]")
        if min(self.board_half_extent) < 0:
            raise ValueError("Board half extent must be non-negative")
        if not 0 <= self.blade_bevel_height < 2 * self.blade_half_size[2]:
            raise ValueError("Invalid blade geometry")
        if sum((self.free_blade, self.pose_servo, self.external_tool)) > 1:
            raise ValueError("Choose free, pose-servo, or externally prescribed tool")
        if self.tool_count < 1 or (
            self.tool_count > 1 and not (self.free_blade and not self.accumulate_tool_impulses)
        ):
            raise ValueError("Several tools are supported only as free knives")
        if self.grid_spacing <= 0 or min(np.asarray(self.size) / self.grid_spacing) < 2:
            raise ValueError("Grid spacing must be positive with at least two cells per axis")
        if not 1 <= int(self.fields_per_node) <= 255:
            raise ValueError("fields_per_node must be in [1, 255]")

    def contact_scales(self):
        """Effective (s_n, gamma_k) at this dt (`solver.py:146-149`)."""
        ratio = 1.0 if self.contact_reference_dt == 0 else (self.dt / self.contact_reference_dt) ** 2
        return self.normal_scale * ratio, self.tangent_stiffness * ratio


def derived(cfg):
    """Per-particle constants of `solver.py:181-200` (volume, mass, radius, Lame moduli,
    damage threshold, blade inertia)."""
    h = cfg.grid_spacing
    volume = (h / 2) ** 3
    radius = h / 4
    hx, hy, hz = cfg.blade_half_size
    mu = cfg.young_modulus / (2 * (1 + cfg.poisson_ratio))
    lam = cfg.young_modulus * cfg.poisson_ratio / ((1 + cfg.poisson_ratio) * (1 - 2 * cfg.poisson_ratio))
    normal_scale, tangent_stiffness = cfg.contact_scales()
    maxwell = cfg.maxwell_tau > 0 and cfg.maxwell_fraction > 0
    return dict(
        volume=volume,
        mass=cfg.density * volume,
        radius=radius,
        blade_offset=max(cfg.blade_contact_radius or radius, radius - hx),
        mu=mu,
        lam=lam,
        maxwell=maxwell,
        mu_maxwell=cfg.maxwell_fraction * mu,
        maxwell_relax=float(-np.expm1(-cfg.dt / cfg.maxwell_tau)) if maxwell else 0.0,
        threshold=cfg.cutting_energy * volume ** (2 / 3) * (1 - cfg.sharpness) / cfg.sharpness,
        normal_scale=normal_scale,
        tangent_stiffness=tangent_stiffness,
        board_scale=cfg.board_normal_scale or normal_scale,
        inertia_inv=3.0 / (cfg.blade_mass * np.array([hy * hy + hz * hz, hx * hx + hz * hz, hx * hx + hy * hy])),
    )


def max_cut_index(node_count):
    """Largest cut index whose int64 keys code * node_count + node cannot overflow
    (`solver.py:168-171`)."""
    k = 0
    while 3 ** (k + 2) * node_count <= 2**63 - 1:
        k += 1
    return k
