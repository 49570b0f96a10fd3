"""Many independent cutting instances in one solver: the batched driver of force_convergence.

`scripts/force_convergence.cut_trace` (`:92-129`) runs one prescribed cut per Solver, one
after the other.  Here the instances are packed side by side on one grid, with a gap wider than
a knife between neighbours, so no particle, grid node or knife of one instance reaches
another.  Every substep kernel then runs once for all instances. The particle arrays carry an
instance id; the knife state, the Lame constants, the step sums and the interval sums are kept
per instance.  Instances share the grid spacing, time step, density, knife geometry and contact
parameters.  Size, shape, sampling seed, Young's modulus and Poisson's ratio vary per instance,
and so does every knife's pose and twist (submitted per interval, in the instance's own frame).

The brick height of the transfer kernels is chosen by timing (`autotune_brick_z`).
"""

from dataclasses import asdict, replace
import ctypes
import time

import numpy as np
import torch

from . import _lib
from .config import Config, derived
from .coupling import Reaction, ToolState, register_port
from .geometry import grid_cells, points
from .solver import Solver

# Fields that may differ between packed instances.
PER_INSTANCE = {"size", "young_modulus", "poisson_ratio", "sampling_seed", "shape"}


def _reach(cfg):
    """Radius of a sphere around the knife origin that holds the knife."""
    return float(np.linalg.norm(cfg.blade_half_size)) + cfg.grid_spacing


def pack(configs, gap=None):
    """Side-by-side placement of the instances' reference particles on one grid.

    Shelf packing in x and y with `gap` metres between neighbours: a number, or (x, y) gaps.
    The default is a knife diameter plus two cells, safe for any knife orientation.  Each
    instance keeps its own lattice: its grid corner lands on a node of the packed grid, at an
    even cell index in x and y, so that an instance starts on a brick boundary (bricks are
    two cells wide in x and y) and a block with even cell counts fills its bricks.
    Returns the layout for `Solver` and, per instance, the offset added to its own coordinates
    and its particle range.
    """
    base = configs[0]
    h = base.grid_spacing
    if gap is None:
        gap = 2.0 * max(_reach(c) for c in configs) + 2 * h
    gx, gy = (gap, gap) if np.isscalar(gap) else gap
    if min(gx, gy) < 2 * h:
        raise ValueError("Instances need at least two empty cells between them")
    gap_cells = np.array([int(np.ceil(gx / h)), int(np.ceil(gy / h))])
    parts = []
    for cfg in configs:
        ref, origin, cells = points(cfg.size, h, cfg.shape, cfg.sampling_seed)
        parts.append((ref, origin, np.asarray(cells)))
    widths = np.array([p[2][0] for p in parts]) + gap_cells[0]
    depths = np.array([p[2][1] for p in parts]) + gap_cells[1]
    row_width = max(int(np.ceil(np.sqrt(float(np.sum(widths * depths))))), int(widths.max()))

    def even(v):
        return v + (v & 1)

    corner = []
    x = y = row_depth = 0
    for w, d in zip(widths, depths):
        if x and x + w > row_width:
            x, y, row_depth = 0, even(y + row_depth), 0
        corner.append((x, y))
        x = even(x + w)
        row_depth = max(row_depth, d)
    corner = np.array(corner, dtype=np.int64)
    cells = np.array([max(c[0] + p[2][0] for c, p in zip(corner, parts)),
                      max(c[1] + p[2][1] for c, p in zip(corner, parts)),
                      max(p[2][2] for p in parts)], dtype=np.int64)
    origin = np.zeros(3)
    reference, instance, offsets, slices = [], [], [], []
    start = 0
    mu, lam = [], []
    for i, ((ref, own_origin, own_cells), (cx, cy)) in enumerate(zip(parts, corner)):
        offset = origin + np.array([cx * h, cy * h, 0.0]) - own_origin
        reference.append(ref + offset)
        instance.append(np.full(len(ref), i, dtype=np.int32))
        offsets.append(offset)
        slices.append(slice(start, start + len(ref)))
        start += len(ref)
        d = derived(configs[i])
        mu.append(d["mu"])
        lam.append(d["lam"])
    layout = dict(reference=np.concatenate(reference), origin=origin, cells=cells,
                  instance=np.concatenate(instance), mu=np.array(mu), lam=np.array(lam))
    return layout, np.array(offsets), slices


class BatchSolver(Solver):
    """Independent cutting instances advanced together.

    configs: one Config per instance; they must agree on every field but size, shape,
    sampling seed, Young's modulus and Poisson's ratio, and use an external knife with the
    mobility predictor (prescribed motion: a zero mobility matrix).
    """

    def __init__(self, configs, device="cuda:0", brick_z=None, gap=None):
        """gap: spacing between packed instances, metres (a number or (x, y); see `pack`)."""
        configs = list(configs)
        if not configs:
            raise ValueError("No instances")
        base = configs[0]
        shared = {k: v for k, v in asdict(base).items() if k not in PER_INSTANCE}
        for cfg in configs[1:]:
            other = {k: v for k, v in asdict(cfg).items() if k not in PER_INSTANCE}
            if other != shared:
                diff = sorted(k for k in shared if shared[k] != other[k])
                raise ValueError(f"Batched instances must share {diff}")
        if not base.external_tool or base.tool_substep != "predict":
            raise ValueError("Batched instances need external_tool=True with tool_substep='predict'")
        layout, self.offsets, self.slices = pack(configs, gap)
        self.instance_configs = configs
        size = tuple(float(c) * base.grid_spacing for c in layout["cells"])
        super().__init__(replace(base, size=size), device, brick_z, _layout=layout)

    # ------------------------------------------------------------------ per-instance views
    def instance_state(self, i, name):
        """Host copy of a per-particle array of instance i, positions in its own frame."""
        values = getattr(self, name).numpy()[self.slices[i]]
        if name in ("x", "ref"):
            values = values - self.offsets[i].astype(values.dtype)
        return values

    def instance_labels(self, i):
        """Piece labels of instance i alone (`topology.components` of its particles)."""
        labels, _ = self.piece_labels()
        _, local = np.unique(labels[self.slices[i]], return_inverse=True)
        local = local.ravel().astype(np.int64)
        return local, np.bincount(local)


class BatchPort:
    """Per-instance coupling intervals for a BatchSolver (the batched ExternalToolPort).

    submit(states, duration) enqueues one interval for all instances: each ToolState is in
    its instance's own frame; collect() returns one Reaction per instance, with the angular
    impulse about the tool origin, as `ExternalToolPort` returns it (coupling.py:181-182).

    `submit_records` takes the 49-float records of all instances as one array and writes the
    interval's result into one of `slots` result buffers, so several intervals can be in
    flight with one readback for all of them (`collect_wrenches`).  With `patch=True` every
    interval also leaves the occupancy/damage patch of its end state in its slot
    (`Solver.blade_patch`).
    """

    def __init__(self, solver, slots=1, patch=False):
        self.solver = s = solver
        dev = s.device.torch
        I = s.instances
        self.slots = int(slots)
        self.packed = torch.zeros(self.slots, I, 49, dtype=torch.float32, device=dev)
        self.results = torch.zeros(self.slots, I, 6, dtype=torch.float32, device=dev)
        self.result = self.results[0]
        self.patches = torch.zeros(self.slots, I, 2, 8, 8, 8, dtype=torch.float32, device=dev) if patch else None
        self._staging = [torch.zeros(I, 49, dtype=torch.float32, pin_memory=True) for _ in range(4 * self.slots)]
        self._staged = [None] * len(self._staging)
        self._stage_next = 0
        register_port(self, "batch")

    def _ptr(self, tensor):
        return ctypes.c_void_p(tensor.data_ptr())

    def _steps(self, duration_s):
        s = self.solver
        steps = int(round(duration_s / s.config.dt))
        if steps < 1 or abs(steps * s.config.dt - duration_s) > 1.0e-12 + 1.0e-10 * abs(duration_s):
            raise ValueError("Coupling duration must be an integer number of material substeps")
        return steps

    def submit(self, states, duration_s, *, cutting=True, cut_index=0):
        s = self.solver
        if len(states) != s.instances:
            raise ValueError(f"{len(states)} tool states for {s.instances} instances")
        states = [st.checked() for st in states]
        records = np.stack([np.concatenate([st.position, st.orientation_xyzw, st.linear_velocity,
                                            st.angular_velocity, np.asarray(st.mobility).ravel()]) for st in states])
        steps = self.submit_records(records, duration_s, cutting=cutting, cut_index=cut_index)
        return states, steps

    def submit_records(self, records, duration_s, *, cutting=True, cut_index=0, slot=0):
        """Enqueue one interval; records [instances][49] in the instances' own frames."""
        s = self.solver
        records = np.asarray(records, dtype=np.float64)
        if records.shape != (s.instances, 49) or not np.isfinite(records).all():
            raise ValueError(f"Tool records must be a finite ({s.instances}, 49) array")
        if not 0 <= slot < self.slots:
            raise ValueError(f"Slot {slot} outside [0, {self.slots})")
        steps = self._steps(duration_s)
        cut_index = s._set_cut(cut_index)
        host = records.astype(np.float32)
        host[:, :3] += s.offsets.astype(np.float32)
        k = self._stage_next
        self._stage_next = (k + 1) % len(self._staging)
        if self._staged[k] is not None:
            self._staged[k].synchronize()
        self._staging[k].copy_(torch.from_numpy(host))
        packed, result = self.packed[slot], self.results[slot]
        packed.copy_(self._staging[k], non_blocking=True)
        done = torch.cuda.Event()
        done.record(torch.cuda.current_stream(s.device.torch))
        self._staged[k] = done

        def load():
            s.launch("bm_load_external_batch", self._ptr(packed))

        def finish():
            s.launch("bm_validate_interval")
            s.launch("bm_store_interval_batch", self._ptr(result))
            if self.patches is not None:
                s.blade_patch(self.patches[slot])

        unused_targets = np.zeros((steps, 3), dtype=np.float32)
        _, _, graph = s.prepare(unused_targets, cutting, cut_index, before=[load], after=[finish],
                                tag=self._owner + (slot,))
        graph.replay()
        s.step_count += steps
        return steps

    def load_records(self, records):
        """Set every knife's state (records as for `submit_records`) without running an interval."""
        s = self.solver
        host = np.asarray(records, dtype=np.float32).reshape(s.instances, 49).copy()
        host[:, :3] += s.offsets.astype(np.float32)
        self.packed[0].copy_(torch.from_numpy(host))
        s.launch("bm_load_external_batch", self._ptr(self.packed[0]))

    def _check_errors(self):
        """The checks of `ExternalToolPort.collect` (coupling.py:183-188): any device error
        count (storage overflow, inverted material, coincident proxies) invalidates the run."""
        errors = self.solver.errors.numpy()
        if errors.any():
            raise RuntimeError(f"Invalid material coupling interval: GPU errors {errors.astype(int).tolist()}")

    def _about_tools(self, out, positions):
        """Per slot and instance (impulse, angular impulse about the tool origin) from the raw
        sums about the packed domain's origin."""
        impulse = out[..., :3]
        tool = np.asarray(positions, dtype=np.float64) + self.solver.offsets
        return impulse, out[..., 3:] - np.cross(tool, impulse)

    def collect(self, states, duration_s):
        self._check_errors()
        out = self.results[0].cpu().numpy().astype(np.float64)
        impulse, angular = self._about_tools(out, np.stack([np.asarray(st.position, dtype=float) for st in states]))
        return [Reaction(duration_s=duration_s, impulse_ns=impulse[i], angular_impulse_nms=angular[i],
                         origin_world_m=np.asarray(st.position, dtype=float)) for i, st in enumerate(states)]

    def collect_wrenches(self, positions, duration_s, slots=None, extra=None):
        """Wait for the submitted intervals and return their wrenches on the knives,
        [len(slots)][instances][6] (force 3, torque 3 about the tool origin, averaged over the
        interval); positions [len(slots)][instances][3] are the tool origins each interval
        started from.  `extra`, a device tensor, is read back in the same copy and returned
        second."""
        slots = list(range(self.slots)) if slots is None else list(slots)
        self._check_errors()
        raw = self.results[slots].reshape(len(slots), -1)
        if extra is not None:
            raw = torch.cat([raw.reshape(-1), extra.reshape(-1).to(torch.float32)])
        host = raw.cpu().numpy().astype(np.float64)
        n = len(slots) * self.solver.instances * 6
        out = host[:n].reshape(len(slots), self.solver.instances, 6)
        impulse, angular = self._about_tools(out, np.asarray(positions, dtype=np.float64).reshape(len(slots), -1, 3))
        wrench = np.concatenate([impulse, angular], axis=-1) / duration_s
        if extra is not None:
            return wrench, host[n:].reshape(tuple(extra.shape))
        return wrench

    def advance(self, states, duration_s, *, cutting=True, cut_index=0):
        states, _ = self.submit(states, duration_s, cutting=cutting, cut_index=cut_index)
        return self.collect(states, duration_s)


def autotune_brick_z(configs, candidates=(4, 8, 12, 16, 24), substeps=20, device="cuda:0"):
    """Brick height with the shortest substep for this batch, by timing a prescribed
    plunge of every knife over `substeps` substeps per candidate."""
    best, timings = None, {}
    for bz in candidates:
        solver = BatchSolver(configs, device=device, brick_z=bz)
        port = BatchPort(solver)
        states = [ToolState([0.0, 0.0, c.size[2] + 0.01 + c.blade_half_size[2]], [0, 0, 0, 1], [0, 0, -0.05],
                            [0, 0, 0]) for c in configs]
        dt = configs[0].dt
        port.submit(states, substeps * dt)  # capture
        torch.cuda.synchronize()
        start = time.perf_counter()
        port.submit(states, substeps * dt)
        torch.cuda.synchronize()
        timings[bz] = (time.perf_counter() - start) / substeps
        if best is None or timings[bz] < timings[best]:
            best = bz
        del port, solver
        torch.cuda.empty_cache()
    return best, timings


def sweep_instances(count=256, seed=0, spacing=0.003, dt=25.0e-6, blade_half_size=(0.0005, 0.035, 0.02)):
    """A batched sweep of uneven cutting instances: per instance a block of nx x ny x nz
    particles (nx in [10, 60], ny in [10, 40], nz in [8, 40]) on the h/2 lattice, Young's
    modulus log-uniform in [1e4, 1e6] Pa, Poisson's ratio uniform in [0.25, 0.45], and a knife
    tilted about x by up to +-15 degrees, offset laterally by [0, h/2] and plunging at
    [0.02, 0.2] m/s from 2 mm above the block.

    Returns (configs, knife(t) per instance, gap for `pack`): the knives stay in their
    instances' x range, so x gaps are three cells; y gaps hold the knives' overhang."""
    rng = np.random.default_rng(seed)
    lattice = spacing / 2
    configs, knives, overhang = [], [], 0.0
    for _ in range(count):
        nx, ny, nz = int(rng.integers(10, 61)), int(rng.integers(10, 41)), int(rng.integers(8, 41))
        size = (nx * lattice, ny * lattice, nz * lattice)
        cfg = Config(size=size, grid_spacing=spacing, dt=dt, young_modulus=float(10 ** rng.uniform(4, 6)),
                     poisson_ratio=float(rng.uniform(0.25, 0.45)), external_tool=True, tool_substep="predict",
                     cache_topology=True, grid_storage="compact", blade_half_size=blade_half_size)
        tilt = np.deg2rad(rng.uniform(-15.0, 15.0))
        offset = rng.uniform(0.0, spacing / 2)
        speed = rng.uniform(0.02, 0.2)
        q = np.array([np.sin(tilt / 2), 0.0, 0.0, np.cos(tilt / 2)])
        hy, hz = blade_half_size[1], blade_half_size[2]
        reach_z = hz * np.cos(tilt) + hy * abs(np.sin(tilt))
        z0 = size[2] + 0.002 + reach_z

        def knife(t, q=q, z0=z0, offset=offset, speed=speed):
            return ToolState([offset, 0.0, z0 - speed * t], q, [0.0, 0.0, -speed], [0.0, 0.0, 0.0])

        configs.append(cfg)
        knives.append(knife)
        overhang = max(overhang, hy * abs(np.cos(tilt)) + hz * abs(np.sin(tilt)) - size[1] / 2)
    gap = (3 * spacing, 2 * max(overhang, 0.0) + 3 * spacing)
    return configs, knives, gap
