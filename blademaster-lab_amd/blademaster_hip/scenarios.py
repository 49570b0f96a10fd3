"""Prescribed-knife cut scenes: the counterpart of `blademaster_warp/scenarios.py`.

A scene is a list of keyframes (time s, tool-origin position m, Z-Y-X Euler angles deg, cut
event; -1 = passive contact).  Between keyframes the pose moves linearly and the orientation
by spherical interpolation; the cut index of a time is the largest event reached so far
(`command_at`, scenarios.py:293-307).  `run` drives a Solver through `ExternalToolPort` one
interval at a time and writes the reference's outputs (scenarios.py:310-430): `manifest.json`,
`states.npz` (every `Solver.state()` key each 20 intervals, `reference`, `forces` rows
[t, F, tau], `energy` rows [t, Metrics columns]) and the piece counts every 500 intervals.

The reference's scenes (parallel, cross-shallow, slice-lift, fig4, fig7-shallow, fig7-deep,
fig8) keep their keyframes and materials (scenarios.py:14-245).  `three-cuts` is new: three
sequential prescribed cuts through a horizontal log (18 mm diameter, 160 mm long, 3 mm grid,
11,984 particles), at 40, 80 and 120 mm from its -x end, each plunging at 50 mm/s to the board
and then pushing the cut piece 10 mm away; E = 2e5 Pa, nu = 0.35, 40 us steps, 4 field slots.
"""

import argparse
from dataclasses import asdict, replace
from pathlib import Path
import time

import numpy as np
from scipy.spatial.transform import Rotation, Slerp

from .config import Config
from .coupling import ExternalToolPort, ToolState
from .diagnostics import Metrics
from .geometry import DEMO_KNIFE_BEVEL_HEIGHT_M, DEMO_KNIFE_HALF_SIZE_M
from .runio import StateLog, git_revision, write
from .topology import components

# Scene dimensions of the reference (scenarios.py:14-62).
FIG4_RADIUS, FIG4_HEIGHT = 0.08, 0.056
FIG4_KNIFE_HALF_SIZE = (0.0005, 0.05, 0.04)
FIG4_CUT_ANGLES_DEG = (60.0, -60.0)
FIG4_CUT_CENTRE_RADIUS = 0.04
FIG7_SIZE = (0.08, 0.08, 0.07)
FIG7_TOP = 0.0665
FIG7_ROLL_DEG = -30.0
FIG7_DEPTHS = {"fig7-shallow": 0.006, "fig7-deep": 0.020}
FIG8_SIZE = (0.10, 0.05, 0.042)
FIG8_KNIFE_HALF_SIZE = (0.0005, 0.04, 0.03)
FIG8_SLICE = 0.012
FIG8_CUTS = 4

# The three-cut log.
LOG_SIZE = (0.16, 0.018, 0.018)
LOG_KNIFE_HALF_SIZE = (0.0005, 0.02, 0.02)
LOG_CUTS_FROM_END = (0.04, 0.08, 0.12)
LOG_PLUNGE_SPEED = 0.05
LOG_PUSH = 0.01


def _euler(angles):
    return Rotation.from_euler("ZYX", angles, degrees=True)


def _score(t0, start, end, index, depth, speed=0.02, clearance=0.01):
    """Rolled knife: its lowest corner plunges at `start`, runs to `end` at `depth`, withdraws
    (scenarios.py:34-57)."""
    start, end = np.asarray(start, float), np.asarray(end, float)
    heading = (end - start) / np.linalg.norm(end - start)
    angles = [np.degrees(np.arctan2(-heading[0], heading[1])), 0.0, FIG7_ROLL_DEG]
    hx, hy, hz = FIG4_KNIFE_HALF_SIZE
    corners = _euler(angles).apply([[0.0, -hy, -hz], [0.0, hy, -hz]])
    tip = corners[np.argmin(corners[:, 2])]

    def origin(xy, z):
        return (np.array([xy[0], xy[1], z]) - tip).tolist()

    t_plunge = t0 + 0.2 + (clearance + depth) / speed
    t_run = t_plunge + np.linalg.norm(end - start) / speed
    return [(t0, origin(start, FIG7_TOP + clearance), angles, index),
            (t0 + 0.2, origin(start, FIG7_TOP + clearance), angles, index),
            (t_plunge, origin(start, FIG7_TOP - depth), angles, index),
            (t_run, origin(end, FIG7_TOP - depth), angles, -1),
            (t_run + 1.0, origin(end, FIG7_TOP + clearance + 0.02), angles, -1)]


def _slice(t0, x, index, speed=0.03, push=0.02):
    """Cut through at x, rise to 12 mm below the top, tip the slice towards +x (scenarios.py:67-91)."""
    hz = FIG8_KNIFE_HALF_SIZE[2]
    top, bottom, tip_over = FIG8_SIZE[2] + 0.01 + hz, 0.0003 + hz, FIG8_SIZE[2] - 0.012 + hz
    t_down = t0 + 0.2 + (top - bottom) / speed
    t_rise = t_down + 0.2 + 0.6
    t_push = t_rise + push / 0.025
    t_up = t_push + 0.6
    flat = [0, 0, 0]
    return [(t0, [x, 0, top], flat, index), (t0 + 0.2, [x, 0, top], flat, index),
            (t_down, [x, 0, bottom], flat, index), (t_down + 0.2, [x, 0, bottom], flat, -1),
            (t_rise, [x, 0, tip_over], flat, -1), (t_push, [x + push, 0, tip_over], flat, -1),
            (t_up, [x + push, 0, top + 0.01], flat, -1), (t_up + 0.8, [x + push, 0, top + 0.01], flat, -1)]


def _radial(t0, angle_deg, index, top, bottom, descend_s, hold_s=0.2, withdraw_s=1.0):
    """Vertical stroke with the knife along the radius at angle_deg (scenarios.py:94-107)."""
    a = np.deg2rad(angle_deg)
    xy = [FIG4_CUT_CENTRE_RADIUS * np.cos(a), FIG4_CUT_CENTRE_RADIUS * np.sin(a)]
    angles = [angle_deg - 90.0, 0, 0]
    t_down = t0 + 0.2 + descend_s
    return [(t0, [*xy, top], angles, index), (t0 + 0.2, [*xy, top], angles, index),
            (t_down, [*xy, bottom], angles, index), (t_down + hold_s, [*xy, bottom], angles, -1),
            (t_down + hold_s + withdraw_s, [*xy, top], angles, -1)]


def _log_cut(t0, x, index):
    """Plunge to the board at x, dwell, push the cut piece (on the -x side) 10 mm, lift."""
    hz = LOG_KNIFE_HALF_SIZE[2]
    top, bottom = LOG_SIZE[2] + 0.01 + hz, 0.0003 + hz
    t_down = t0 + 0.2 + (top - bottom) / LOG_PLUNGE_SPEED
    t_push = t_down + 0.2 + LOG_PUSH / 0.025
    flat = [0, 0, 0]
    return [(t0, [x, 0, top], flat, index), (t0 + 0.2, [x, 0, top], flat, index),
            (t_down, [x, 0, bottom], flat, index), (t_down + 0.2, [x, 0, bottom], flat, -1),
            (t_push, [x - LOG_PUSH, 0, bottom], flat, -1), (t_push + 0.6, [x - LOG_PUSH, 0, top], flat, -1)]


def keyframes(name):
    """(time, position, Z-Y-X Euler degrees, cut event) of a scene (scenarios.py:110-194)."""
    still = [0, 0, 0]
    if name == "parallel":
        return [(0, [-0.012, 0, 0.06], still, 0), (0.3, [-0.012, 0, 0.06], still, 0),
                (2.3, [-0.012, 0, 0.0202], still, 0), (2.5, [-0.012, 0, 0.0202], still, -1),
                (3.5, [-0.012, 0, 0.07], still, -1), (4.2, [0.012, 0, 0.07], still, 1),
                (6.5, [0.012, 0, 0.0202], still, 1), (6.7, [0.012, 0, 0.0202], still, -1),
                (8, [0.012, 0, 0.07], still, -1)]
    if name == "cross-shallow":
        turned = [90, 0, 0]
        return [(0, [0, 0, 0.06], still, 0), (2, [0, 0, 0.035], still, 0), (2.2, [0, 0, 0.035], still, -1),
                (3.2, [0, 0, 0.07], still, -1), (4.2, [0, 0, 0.07], turned, 1), (6.2, [0, 0, 0.035], turned, 1),
                (6.4, [0, 0, 0.035], turned, -1), (7.5, [0, 0, 0.07], turned, -1)]
    if name == "slice-lift":
        pitched = [0, 90, 0]
        return keyframes("parallel") + [
            (9, [0.065, 0, 0.07], still, -1), (10, [0.065, 0, 0.05], pitched, -1),
            (11, [0.065, 0, 0.003], pitched, -1), (13, [0, 0, 0.003], pitched, -1),
            (15, [0, 0, 0.035], pitched, -1), (16, [0, 0, 0.035], pitched, -1)]
    if name == "fig4":
        hz = FIG4_KNIFE_HALF_SIZE[2]
        top, bottom = FIG4_HEIGHT + 0.01 + hz, 0.0003 + hz
        descend = (top - bottom) / 0.04
        first = _radial(0.0, FIG4_CUT_ANGLES_DEG[0], 0, top, bottom, descend)
        second = _radial(first[-1][0] + 0.8, FIG4_CUT_ANGLES_DEG[1], 1, top, bottom, descend)
        # Flat knife slid under the slice at board level and lifted 40 mm (scenarios.py:145-167).
        flat, t, low = [0, 90, 0], second[-1][0], -FIG4_KNIFE_HALF_SIZE[0]
        outside = [0.15, 0, top]
        return first + second + [
            (t + 0.6, outside, second[-1][2], -1), (t + 1.2, outside, flat, -1), (t + 2.0, [0.15, 0, low], flat, -1),
            (t + 4.0, [0.07, 0, low], flat, -1), (t + 4.3, [0.07, 0, low], flat, -1),
            (t + 5.8, [0.07, 0, low + 0.04], flat, -1), (t + 6.8, [0.07, 0, low + 0.04], flat, -1)]
    if name in FIG7_DEPTHS:
        depth = FIG7_DEPTHS[name]
        first = _score(0.0, [-0.03, 0.0], [0.03, 0.0], 0, depth)
        second = _score(first[-1][0] + 1.0, [0.0, -0.03], [0.0, 0.03], 1, depth)
        return first + second + [(second[-1][0] + 0.5, second[-1][1], second[-1][2], -1)]
    if name == "fig8":
        keys = []
        for k in range(FIG8_CUTS):
            keys += _slice(keys[-1][0] + 0.6 if keys else 0.0, FIG8_SIZE[0] / 2 - (k + 1) * FIG8_SLICE, k)
        return keys + [(keys[-1][0] + 1.5, keys[-1][1], keys[-1][2], -1)]
    if name == "three-cuts":
        keys = []
        for k, from_end in enumerate(LOG_CUTS_FROM_END):
            keys += _log_cut(keys[-1][0] + 0.2 if keys else 0.0, from_end - LOG_SIZE[0] / 2, k)
        return keys + [(keys[-1][0] + 0.5, keys[-1][1], keys[-1][2], -1)]
    raise ValueError(name)


def _scene(name):
    """Material, geometry and knife of a scene (scenarios.py:197-284)."""
    # The reference's field table has no cap (one hash slot per code and node); crossing cuts
    # put up to 9 codes (3 sides of each of two cuts) on a node near the crossing, which 4 field
    # slots cannot hold: these scenes take 16.
    common = dict(external_tool=True, grid_storage="compact", blade_contact_radius=0.0,
                  blade_bevel_height=DEMO_KNIFE_BEVEL_HEIGHT_M, fields_per_node=16)
    if name == "fig4":
        return Config(size=(2 * FIG4_RADIUS, 2 * FIG4_RADIUS, FIG4_HEIGHT), shape="cylinder", grid_spacing=0.01,
                      dt=0.002, normal_scale=1.0, board_normal_scale=1.0, young_modulus=10000.0, poisson_ratio=0.4,
                      density=1100.0, cutting_energy=10.0, blade_half_size=FIG4_KNIFE_HALF_SIZE, **common)
    if name in FIG7_DEPTHS:
        return Config(size=FIG7_SIZE, shape="block", grid_spacing=0.01, dt=0.002, normal_scale=1.0,
                      board_normal_scale=1.0, young_modulus=6000.0, poisson_ratio=0.35, density=1100.0,
                      cutting_energy=10.0, blade_half_size=FIG4_KNIFE_HALF_SIZE, **common)
    if name == "fig8":
        return Config(size=FIG8_SIZE, shape="block", grid_spacing=0.006, dt=0.001, normal_scale=1.0,
                      board_normal_scale=1.0, young_modulus=20000.0, poisson_ratio=0.4, density=1200.0,
                      yield_stress=2000.0, cutting_energy=10.0, blade_half_size=FIG8_KNIFE_HALF_SIZE, **common)
    if name == "three-cuts":
        return Config(size=LOG_SIZE, shape="horizontal_cylinder", grid_spacing=0.003, dt=40.0e-6,
                      young_modulus=2.0e5, poisson_ratio=0.35, cache_topology=True,
                      blade_half_size=LOG_KNIFE_HALF_SIZE, **dict(common, fields_per_node=4))
    return Config(size=(0.06, 0.03, 0.03), grid_spacing=0.006, dt=0.00025, young_modulus=6000, poisson_ratio=0.35,
                  external_tool=True, blade_half_size=DEMO_KNIFE_HALF_SIZE_M,
                  blade_bevel_height=DEMO_KNIFE_BEVEL_HEIGHT_M, fields_per_node=16)


def scene_config(name, dt=None, normal_scale=None):
    """A scene's Config; another dt keeps the contact stiffness of the scene's step, and an
    explicit normal_scale is the per-step s_n at the actual dt (scenarios.py:197-212)."""
    cfg = _scene(name)
    if dt is not None and dt != cfg.dt:
        cfg = replace(cfg, dt=dt, contact_reference_dt=cfg.dt)
    if normal_scale is not None:
        cfg = replace(cfg, normal_scale=normal_scale * cfg.normal_scale / cfg.contact_scales()[0])
    return cfg


PROVENANCE = {
    "fig4": "Paper Fig. 4 E/nu/rho and Table I step; local cylinder size, knife, path, Gc=10 and contact numerics",
    "fig7-shallow": "Paper Fig. 7 E/nu/rho; local block size, knife, roll, path, depth, Gc=10 and step",
    "fig7-deep": "Paper Fig. 7 E/nu/rho; local block size, knife, roll, path, depth, Gc=10 and step",
    "fig8": "Paper Fig. 6/8 E/nu/rho/yield; local block size, knife, slice width, tip-over push, Gc=10, step and contact",
    "three-cuts": "Three sequential prescribed cuts through an 18 mm log at 40/80/120 mm, 50 mm/s plunge, 10 mm push",
}
SCENARIOS = ["parallel", "cross-shallow", "slice-lift", "fig4", *FIG7_DEPTHS, "fig8", "three-cuts"]


def command_at(t, keys):
    """Tool state, cutting flag and cut index at time t (scenarios.py:287-307)."""
    times = np.array([k[0] for k in keys], dtype=float)
    j = int(np.clip(np.searchsorted(times, t, side="right") - 1, 0, len(keys) - 2))
    (t0, p0, e0, c0), (t1, p1, e1, _) = keys[j], keys[j + 1]
    span = t1 - t0
    u = float(np.clip((t - t0) / span, 0.0, 1.0))
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    ends = _euler([e0, e1])
    orientation = Slerp([0.0, 1.0], ends)(u)
    # World angular velocity: the relative rotation's vector, expressed in world axes.
    omega = ends[0].apply((ends[0].inv() * ends[1]).as_rotvec()) / span
    index = max(k[3] for k in keys[: j + 1])
    return ToolState((1 - u) * p0 + u * p1, orientation.as_quat(), (p1 - p0) / span, omega), c0 >= 0, index


def run(output, scenario="parallel", dt=None, single_field=False, normal_scale=None, board_normal_scale=None,
        duration=None, log=print):
    """Drive one scene and write its outputs into `output` (which must not exist)."""
    from .solver import Solver

    out = Path(output)
    out.mkdir(parents=True, exist_ok=False)
    cfg = scene_config(scenario, dt, normal_scale)
    if board_normal_scale is not None:
        cfg = replace(cfg, board_normal_scale=board_normal_scale)
    normal, tangent = cfg.contact_scales()
    log(f"{scenario}: dt {cfg.dt:g} s, per-step s_n {normal:.4g}, gamma_k {tangent:.4g}")
    keys = keyframes(scenario)
    interval = max(0.002, cfg.dt)
    end = keys[-1][0] if duration is None else min(duration, keys[-1][0])
    s = Solver(cfg)
    port = ExternalToolPort(s)
    meter = Metrics(s)
    meta = {
        "scenario": scenario, "config": asdict(cfg), "contact_scales_at_dt": dict(normal=normal, tangent=tangent),
        "particles": s.n, "keyframes": keys, "driver": "prescribed tool (infinite effective mass)",
        "parameter_provenance": PROVENANCE.get(
            scenario, "Fig.7 elastic E/nu/rho; local geometry, path, Gc=.5 and contact numerics"),
        "single_field_after_withdrawal": bool(single_field), "revision": git_revision(), "no_presplit": True,
        "surface_interval_s": None, "status": "running",
    }
    write(out / "manifest.json", meta)
    states = StateLog()
    forces, energy, events = [], [], []
    minimum, status, error, physics = 1.0, "completed", None, 0.0
    count = int(round(end / interval))
    begin = time.perf_counter()
    try:
        for i in range(count):
            t = i * interval
            tool, active, index = command_at(t, keys)
            s.set_single_field(single_field and not active)
            tick = time.perf_counter()
            reaction = port.advance(tool, interval, cutting=active, cut_index=index)
            physics += time.perf_counter() - tick
            forces.append([t, *reaction.force_n, *reaction.torque_nm])
            if i % 20 == 0 or i == count - 1:
                states.add(s.state())
                energy.append([t, *meter.read()])
                minimum = min(minimum, float(energy[-1][-1]))
            if (i + 1) % 250 == 0:
                log(f"{scenario} t={t + interval:.2f}s wall={time.perf_counter() - begin:.1f}s")
            if i % 500 == 0:
                _, sizes = components(s.node_numpy, s.weight_numpy, s.codes.numpy())
                events.append(dict(time_s=t, components=sizes.tolist()))
    except Exception as exc:  # the reference records the failure in the manifest, then raises
        status, error = "failed", repr(exc)
    wall = time.perf_counter() - begin
    np.savez_compressed(out / "states.npz", **{k: np.asarray(v) for k, v in states.saved.items()},
                        reference=s.reference, forces=np.array(forces), energy=np.array(energy))
    _, sizes = components(s.node_numpy, s.weight_numpy, s.codes.numpy())
    simulated = s.step_count * cfg.dt
    meta.update(status=status, error=error, wall_s=wall, completed_sim_s=simulated,
                realtime_ratio=simulated / wall if wall else None, physics_wall_s=physics,
                physics_wall_per_sim=physics / simulated if simulated else None, graph_build_s=s.graph_build_s,
                components=sizes.tolist(), events=events, min_det_F=minimum, surface_wall_s=0.0)
    write(out / "manifest.json", meta)
    if error:
        raise RuntimeError(error)
    return meta


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--output", required=True)
    p.add_argument("--scenario", choices=SCENARIOS, default="parallel")
    p.add_argument("--dt", type=float, default=None, help="Material step; default per scene")
    p.add_argument("--single-field", action="store_true",
                   help="Grid transfers ignore cut-side codes while the knife is not cutting")
    p.add_argument("--normal-scale", type=float, default=None, help="Per-step contact penalty s_n at the actual --dt")
    p.add_argument("--board-normal-scale", type=float, default=None, help="Per-step board penalty in [0, 1]")
    p.add_argument("--duration", type=float, default=None, help="Stop early (seconds of the scene)")
    a = p.parse_args(argv)
    run(a.output, a.scenario, a.dt, a.single_field, a.normal_scale, a.board_normal_scale, a.duration)


if __name__ == "__main__":
    main()
