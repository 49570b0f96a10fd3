"""Blade force under grid refinement and particle resampling: the counterpart of
`scripts/force_convergence.py`, with the per-seed runs of one spacing batched.

The protocol is the reference's: the edge moves from 36 mm to the board at 20 mm/s, then holds
for 0.2 s, in 1 ms coupling intervals.  `cut_trace` runs one cut in a Solver, with the knife
held over each interval as in the reference.  `batch_cut_traces` runs every (seed, lateral
offset) cut of one spacing as the instances of one BatchSolver.  There the knife follows the
same path continuously (prescribed twist, zero mobility) and no hold constraint applies.
`summarize` is the reference's statistics (force_convergence.py:132-164) on either.
"""

from dataclasses import asdict, replace
import time

import numpy as np

from .batch import BatchPort, BatchSolver
from .coupling import ExternalToolPort, ToolState
from .presets import banana_config, banana_dt
from .runio import git_revision
from .topology import components

INTERVAL = 0.001
DESCENT_S, HOLD_S, SPEED = 1.8, 0.2, 0.02
START_EDGE = DESCENT_S * SPEED


def effective_config(spacing, seed, blade_contact_radius=None, tool_substep="hold"):
    base = banana_config()
    cfg = replace(base, grid_spacing=spacing, dt=banana_dt(base, spacing, INTERVAL), sampling_seed=seed,
                  external_tool=True, grid_storage="compact", cache_topology=True, tool_substep=tool_substep)
    if blade_contact_radius is not None:
        cfg = replace(cfg, blade_contact_radius=blade_contact_radius)
    return cfg


def knife(t, half, x_offset):
    moving = t < DESCENT_S
    edge = START_EDGE - SPEED * t if moving else 0.0
    return ToolState([x_offset, 0.0, edge + half], [0, 0, 0, 1], [0.0, 0.0, -SPEED if moving else 0.0], [0, 0, 0])


def _run_record(cfg, spacing, seed, x_offset, force, pieces, wall, errors, min_det, work, damaged, particles):
    return dict(spacing=spacing, seed=seed, x_offset_m=x_offset, dt=cfg.dt,
                blade_contact_radius=cfg.blade_contact_radius, particles=particles, force_n=force.tolist(),
                pieces=pieces.tolist(), wall_s=wall, gpu_errors=errors, min_det_F=min_det, directional_work_j=work,
                damaged_particles=damaged, config=asdict(cfg), revision=git_revision())


def cut_trace(spacing, seed, blade_contact_radius=None, x_offset_fraction=0.0, duration=DESCENT_S + HOLD_S):
    """One cut in its own Solver (force_convergence.py:92-129)."""
    from .solver import Solver

    cfg = effective_config(spacing, seed, blade_contact_radius)
    s = Solver(cfg)
    port = ExternalToolPort(s)
    half = cfg.blade_half_size[2]
    steps = round(duration / INTERVAL)
    force = np.empty((steps, 3))
    x_offset = x_offset_fraction * spacing / 2
    begin = time.perf_counter()
    for i in range(steps):
        force[i] = port.advance(knife(i * INTERVAL, half, x_offset), INTERVAL, cutting=True).force_n
    wall = time.perf_counter() - begin
    _, pieces = components(s.node_numpy, s.weight_numpy, s.codes.numpy())
    return _run_record(cfg, spacing, seed, x_offset, force, pieces, wall, s.errors.numpy().tolist(),
                       float(np.linalg.det(s.F.numpy()).min()), float(s.work.numpy()[0]),
                       int((s.damage.numpy() >= 1).sum()), s.n)


def batch_cut_traces(spacing, seeds, x_offset_fractions=(0.0,), blade_contact_radius=None,
                     duration=DESCENT_S + HOLD_S, brick_z=None):
    """Every (seed, offset) cut of one spacing in one BatchSolver; one run record per cut."""
    cases = [(seed, f) for seed in seeds for f in x_offset_fractions]
    cfgs = [effective_config(spacing, seed, blade_contact_radius, tool_substep="predict") for seed, _ in cases]
    s = BatchSolver(cfgs, brick_z=brick_z)
    port = BatchPort(s)
    half = cfgs[0].blade_half_size[2]
    steps = round(duration / INTERVAL)
    force = np.empty((len(cases), steps, 3))
    begin = time.perf_counter()
    for i in range(steps):
        states = [knife(i * INTERVAL, half, f * spacing / 2) for _, f in cases]
        for j, r in enumerate(port.advance(states, INTERVAL, cutting=True)):
            force[j, i] = r.force_n
    wall = time.perf_counter() - begin
    F = s.F.numpy()
    damage = s.damage.numpy()
    work = s.work.numpy()
    runs = []
    for j, ((seed, f), cfg) in enumerate(zip(cases, cfgs)):
        sl = s.slices[j]
        _, pieces = s.instance_labels(j)
        runs.append(_run_record(cfg, spacing, seed, f * spacing / 2, force[j], pieces, wall / len(cases),
                                s.errors.numpy().tolist(), float(np.linalg.det(F[sl]).min()), float(work[j]),
                                int((damage[sl] >= 1).sum()), sl.stop - sl.start))
    return runs


def _mean_std(values):
    """Mean and sample standard deviation (0 for a single value)."""
    values = np.asarray(values, dtype=float)
    return dict(mean=float(values.mean()), std=float(values.std(ddof=1)) if values.size > 1 else 0.0)


class _ForceTable:
    """Vertical knife force of the runs of one spacing, one row per run and one column per
    coupling interval (force at the end of interval k + 1, t = (k + 1) INTERVAL)."""

    def __init__(self, runs):
        self.runs = runs
        self.fz = np.stack([np.asarray(r["force_n"], dtype=float)[:, 2] for r in runs])
        self.t = INTERVAL * np.arange(1, self.fz.shape[1] + 1)

    def window_mean(self, start, stop):
        """Per run: the mean force over start < t <= stop."""
        return self.fz[:, (self.t > start) & (self.t <= stop)].mean(axis=1)

    def impulse_until(self, stop):
        """Per run: the impulse of the intervals ending at t <= stop."""
        return INTERVAL * self.fz[:, self.t <= stop].sum(axis=1)

    def bins(self, width_s):
        """Per run: the force averaged over whole bins of width_s (a partial last bin dropped)."""
        per = int(round(width_s / INTERVAL))
        whole = self.fz.shape[1] // per
        return self.fz[:, : whole * per].reshape(len(self.runs), whole, per).mean(axis=2)


def summarize(runs, window=(0.4, 1.4), bin_s=0.02):
    """Statistics over sampling seeds and lateral offsets of each spacing's cut traces: the
    fields of scripts/force_convergence.py:132-164, with the same meaning."""
    groups = {}
    for r in runs:
        groups.setdefault(r["spacing"], []).append(r)
    by_spacing = {}
    for spacing in sorted(groups, reverse=True):
        rows = groups[spacing]
        table = _ForceTable(rows)
        binned = table.bins(bin_s)
        trace_std = binned.std(axis=0, ddof=1) if len(rows) > 1 else np.zeros(binned.shape[1])
        by_spacing[f"{spacing * 1000:g} mm"] = {
            "particles": rows[0]["particles"],
            "dt_us": rows[0]["dt"] * 1.0e6,
            "seeds": len({r["seed"] for r in rows}),
            "samples": len(rows),
            "x_offsets_m": sorted({r.get("x_offset_m", 0.0) for r in rows}),
            "steady_force_n": _mean_std(table.window_mean(*window)),
            "descent_impulse_ns": _mean_std(table.impulse_until(DESCENT_S)),
            "peak_single_run_n": {"mean": float(table.fz.max(axis=1).mean()), "max": float(table.fz.max())},
            "peak_of_mean_trace_n": float(binned.mean(axis=0).max()),
            "severed_runs": sum(1 for r in rows if len(r["pieces"]) > 1),
            "wall_s_per_run": float(np.mean([r["wall_s"] for r in rows])),
            "gpu_errors": int(sum(np.sum(r["gpu_errors"]) for r in rows)),
            "mean_trace_n": binned.mean(axis=0).tolist(),
            "std_trace_n": trace_std.tolist(),
        }
    return {"window_s": list(window), "bin_s": bin_s, "by_spacing": by_spacing}
