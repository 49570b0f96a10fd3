"""Gravity-knife drop (paper Fig. 3), the counterpart of `blademaster_warp/paper_freefall.py`.

A 0.1 kg free knife released at 25 degrees above a soft block falls under gravity and the
material/board reaction only.  Same arguments, and the same outputs in `output/`:
`states.npz` (every `Solver.state()` key per 10 ms snapshot, plus `reference`),
`manifest.json` and `result.json` with the reference's fields (`paper_freefall.py:19-155`).
"""

import argparse
from dataclasses import asdict
from pathlib import Path
import sys
import time

import numpy as np
import torch

from .config import Config
from .runio import StateLog, edge_height, git_revision, write
from .topology import components

FIG3_CUTTING_ENERGY = 100.0


def run(output, sharpness=0.9, dt=0.002, duration=1.0, sampling_seed=1, cutting_energy=FIG3_CUTTING_ENERGY,
        normal_scale=1.0, board_normal_scale=1.0, cut_hold=False, tilt_axis="thickness", cut_work="relative",
        grid_spacing=0.01, contact_reference_dt=0.0, size=(0.14, 0.10, 0.08)):
    from .solver import Solver

    if tilt_axis not in ("length", "thickness"):
        raise ValueError(tilt_axis)
    cfg = Config(
        size=tuple(size), grid_spacing=grid_spacing, dt=dt, young_modulus=6000.0, poisson_ratio=0.4, density=1100.0,
        blade_mass=0.1, blade_half_size=(0.0005, 0.06, 0.02), free_blade=True, sharpness=sharpness,
        cutting_energy=cutting_energy, sampling_seed=sampling_seed, normal_scale=normal_scale,
        board_normal_scale=board_normal_scale, cut_hold=cut_hold, cut_work=cut_work,
        contact_reference_dt=contact_reference_dt,
    )
    steps = round(duration / dt)
    if steps < 1 or not np.isclose(steps * dt, duration):
        raise ValueError("Invalid duration")
    out = Path(output)
    out.mkdir(parents=True, exist_ok=False)
    s = Solver(cfg)
    # Tilt about the blade thickness (inclined edge, one corner enters first) or its length.
    angle = np.deg2rad(25.0)
    hx, hy, hz = cfg.blade_half_size
    if tilt_axis == "length":
        q = np.array([0.0, np.sin(angle / 2), 0.0, np.cos(angle / 2)])
        reach = hz * np.cos(angle) + hx * np.sin(angle)
    else:
        q = np.array([np.sin(angle / 2), 0.0, 0.0, np.cos(angle / 2)])
        reach = hz * np.cos(angle) + hy * np.sin(angle)
    height = cfg.size[2] + 0.015 + reach
    s.tool_rot.assign([q])
    s.tool_pos.assign([[0.0, 0.0, height]])
    meta = {
        "status": "running",
        "experiment": "fig3_freefall",
        "config": asdict(cfg),
        "device": s.device.name,
        "published_values": ["E=6kPa", "nu=0.4", "rho=1100kg/m3", "knife_mass=0.1kg", "tilt=25deg",
                             f"tilt axis (not given): blade {tilt_axis}", "Table I: 8960 particles, dt=2ms"],
        "independent_choices": [f"block size {cfg.size[0] * 1e3:g}x{cfg.size[1] * 1e3:g}x{cfg.size[2] * 1e3:g}mm",
                                f"grid spacing {grid_spacing:g} m", "blade geometry/inertia",
                                "initial clearance 15mm", "sharpness/Gc", "contact/transfer coefficients"],
        "dt_is_table_value": dt == 0.002,
        "actuation": "none; gravity plus material/board reaction",
        "surface": "particles only; reconstruction pending",
        "revision": git_revision(),
        "command": sys.argv,
    }
    log = StateLog(s.state())
    start = time.perf_counter()
    minimum = 1.0
    try:
        stride = max(1, round(0.01 / dt))
        for i in range(0, steps, stride):
            count = min(stride, steps - i)
            s.advance(np.zeros((count, 3)), cutting=True)
            state = s.state()
            minimum = min(minimum, float(np.linalg.det(state["deformation"]).min()))
            log.add(state)
        torch.cuda.synchronize()
        meta.update(status="complete", wall_s_excluding_jit=time.perf_counter() - start)
    except Exception as exc:
        meta.update(status="failed", error=repr(exc), wall_s_excluding_jit=time.perf_counter() - start)
        raise
    finally:
        meta.update(completed_steps=s.step_count, minimum_det_f_sampled=minimum)
        log.save(out / "states.npz", s.reference)
        write(out / "manifest.json", meta)
    _, counts = components(s.node_numpy, s.weight_numpy, log["sides"][-1])
    qfinal = log["blade_orientation"][-1][0]
    turn = 2 * np.arccos(np.clip(abs(np.dot(qfinal, q)), 0.0, 1.0))
    lowest, _ = edge_height([log["blade_position"][-1]], [log["blade_orientation"][-1]], cfg.blade_half_size)
    assigned = log["sides"][-1] != 0
    damage = log["damage"][-1]
    result = {
        "status": "complete",
        "component_sizes": counts.tolist(),
        "wall_s": meta["wall_s_excluding_jit"],
        "orientation_change_deg": float(np.rad2deg(turn)),
        "horizontal_displacement_m": log["blade_position"][-1][0, :2].tolist(),
        "final_blade_height_m": float(log["blade_position"][-1][0, 2]),
        "final_lowest_edge_m": float(lowest[0, 0]),
        "particle_count": s.n,
        "assigned_particles": int(assigned.sum()),
        "assigned_below_threshold_final": int(np.count_nonzero(assigned & (damage < 1.0))),
        "released_particles": int(np.count_nonzero(damage >= 1.0)),
        "directional_contact_work_j": float(s.work.numpy()[0]),
        "gpu_errors": s.errors.numpy().tolist(),
        "minimum_det_f_sampled": minimum,
    }
    write(out / "result.json", result)
    print(result, flush=True)
    return result


if __name__ == "__main__":
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--output", required=True)
    p.add_argument("--sharpness", type=float, default=0.9)
    p.add_argument("--dt", type=float, default=0.002)
    p.add_argument("--grid-spacing", type=float, default=0.01)
    p.add_argument("--contact-reference-dt", type=float, default=0.0)
    p.add_argument("--duration", type=float, default=1.0)
    p.add_argument("--sampling-seed", type=int, default=1)
    p.add_argument("--cutting-energy", type=float, default=FIG3_CUTTING_ENERGY)
    p.add_argument("--normal-scale", type=float, default=1.0)
    p.add_argument("--board-normal-scale", type=float, default=1.0)
    p.add_argument("--cut-hold", action="store_true")
    p.add_argument("--tilt-axis", choices=["length", "thickness"], default="thickness")
    p.add_argument("--cut-work", choices=["relative", "advance"], default="relative")
    a = p.parse_args()
    run(a.output, sharpness=a.sharpness, dt=a.dt, duration=a.duration, sampling_seed=a.sampling_seed,
        cutting_energy=a.cutting_energy, normal_scale=a.normal_scale, board_normal_scale=a.board_normal_scale,
        cut_hold=a.cut_hold, tilt_axis=a.tilt_axis, cut_work=a.cut_work, grid_spacing=a.grid_spacing,
        contact_reference_dt=a.contact_reference_dt)
