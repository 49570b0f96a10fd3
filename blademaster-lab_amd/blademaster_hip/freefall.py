"""Knife drops: a free rigid knife released above a block falls under gravity and the reaction
of the material and the board, with nothing driving it.

A `DropScene` holds everything a drop needs: the block, its material, the knife, how the knife
is tilted and how far above the block its lowest point starts, the time step and the duration.
Two scenes are defined:

- ``fig3``: the paper's Fig. 3 drop, the scene of the reference's `paper_freefall.py`: a 0.1 kg
  knife tilted 25 degrees, 15 mm above a soft 140 x 100 x 80 mm block on a 10 mm grid;
- ``gravity-knife``: a 56 x 20 x 8 = 8960-particle block on a 3 mm lattice (6 mm grid) under a
  20 g, 100 x 30 mm plate released level and at rest 2 mm above it, 62.5 us substeps for
  0.5 s.

`run` writes the reference's outputs into a new directory: ``states.npz`` (every
`Solver.state()` key once per snapshot interval, plus ``reference``), ``manifest.json`` and
``result.json``; their keys are those of `paper_freefall.py:82-155`, so tools written for the
reference's runs read these.

Command line: ``python -m blademaster_hip.freefall OUTPUT [--scene NAME] [--set FIELD=VALUE ...]``.
"""

from dataclasses import asdict, dataclass, field, fields, replace
from pathlib import Path
import argparse
import sys
import time

import numpy as np
import torch
from scipy.spatial.transform import Rotation

from .config import Config
from .runio import StateLog, edge_height, git_revision, write
from .topology import components


@dataclass(frozen=True)
class DropScene:
    name: str
    size: tuple  # block, m
    grid_spacing: float
    young_modulus: float
    poisson_ratio: float
    density: float
    knife_mass: float
    knife_half_size: tuple  # thickness, length, height halves, m
    tilt_deg: float = 0.0
    tilt_axis: str = "thickness"  # the knife axis the tilt turns about: "thickness" or "length"
    clearance: float = 0.0  # lowest knife point above the block's top face, m
    dt: float = 0.002
    duration: float = 1.0
    snapshot_every: float = 0.01  # s of simulated time between saved states
    sampling_seed: int = 0
    cutting_energy: float = 0.5
    sharpness: float = 0.9
    # Config fields passed through unchanged (contact coefficients, cut options).
    extra: dict = field(default_factory=dict)
    # What the scene takes from a publication, and what this scene chooses itself (manifest).
    sources: tuple = ()
    choices: tuple = ()

    def config(self):
        return Config(size=tuple(self.size), grid_spacing=self.grid_spacing, dt=self.dt,
                      young_modulus=self.young_modulus, poisson_ratio=self.poisson_ratio, density=self.density,
                      blade_mass=self.knife_mass, blade_half_size=tuple(self.knife_half_size), free_blade=True,
                      sampling_seed=self.sampling_seed, cutting_energy=self.cutting_energy,
                      sharpness=self.sharpness, **self.extra)

    def steps(self):
        steps = int(round(self.duration / self.dt))
        if steps < 1 or not np.isclose(steps * self.dt, self.duration):
            raise ValueError("The duration must be a whole number of time steps")
        return steps


SCENES = {
    "fig3": DropScene(
        name="fig3_freefall", size=(0.14, 0.10, 0.08), grid_spacing=0.01, young_modulus=6000.0,
        poisson_ratio=0.4, density=1100.0, knife_mass=0.1, knife_half_size=(0.0005, 0.06, 0.02), tilt_deg=25.0,
        clearance=0.015, dt=0.002, duration=1.0, sampling_seed=1, cutting_energy=100.0,
        sources=("paper Fig. 3 / Table I: E 6 kPa, nu 0.4, rho 1100 kg/m3, knife 0.1 kg, tilt 25 deg, "
                 "8960 particles, dt 2 ms",),
        choices=("tilt axis: the paper does not name it", "block 140 x 100 x 80 mm", "knife 1 x 120 x 40 mm",
                 "15 mm initial clearance", "work-based damage with a cutting energy of 100 J/m^2"),
    ),
    "gravity-knife": DropScene(
        name="gravity_knife_8960", size=(0.168, 0.06, 0.024), grid_spacing=0.006, young_modulus=1.0e5,
        poisson_ratio=0.3, density=1000.0, knife_mass=0.02, knife_half_size=(0.0005, 0.05, 0.015),
        clearance=0.002, dt=62.5e-6, duration=0.5,
        sources=("56 x 20 x 8 particles on a 3 mm lattice, 6 mm grid, E 1e5 Pa, nu 0.3, rho 1000 kg/m3, "
                 "20 g knife 100 x 30 mm released at rest 2 mm above the block, 62.5 us substeps, 0.5 s",),
        choices=("fixed-corotated elasticity (the solver's material model)",
                 "work-based damage with the default cutting energy 0.5 J/m^2 and sharpness 0.9",
                 "knife 1 mm thick"),
    ),
}


def release_pose(scene, top):
    """Orientation (xyzw) and origin height of the knife: tilted by tilt_deg about its
    thickness (x) or length (y) axis, lowest corner `clearance` above the plane z = top."""
    axis = {"thickness": [1.0, 0.0, 0.0], "length": [0.0, 1.0, 0.0]}.get(scene.tilt_axis)
    if axis is None:
        raise ValueError(f"tilt_axis must be 'thickness' or 'length', not {scene.tilt_axis!r}")
    rotation = Rotation.from_rotvec(np.deg2rad(scene.tilt_deg) * np.asarray(axis))
    # The lowest corner of the rotated box sits sum_k |R_zk| h_k below its centre.
    below = float(np.abs(rotation.as_matrix()[2]) @ np.asarray(scene.knife_half_size))
    return rotation.as_quat(), top + scene.clearance + below


def outcome(solver, log, start_orientation, wall_s, minimum_det):
    """The result.json fields of a finished drop."""
    sides, damage = log["sides"][-1], log["damage"][-1]
    position, orientation = log["blade_position"][-1], log["blade_orientation"][-1]
    _, sizes = components(solver.node_numpy, solver.weight_numpy, sides)
    # Rotation angle between the start and end orientations (unit quaternions, either sign).
    cos_half = min(1.0, abs(float(np.dot(orientation[0], start_orientation))))
    lowest, _ = edge_height([position], [orientation], solver.config.blade_half_size)
    assigned = sides != 0
    released = damage >= 1.0
    return {
        "status": "complete",
        "component_sizes": sizes.tolist(),
        "wall_s": wall_s,
        "orientation_change_deg": float(np.degrees(2.0 * np.arccos(cos_half))),
        "horizontal_displacement_m": [float(position[0, 0]), float(position[0, 1])],
        "final_blade_height_m": float(position[0, 2]),
        "final_lowest_edge_m": float(lowest[0, 0]),
        "particle_count": int(solver.n),
        "assigned_particles": int(np.count_nonzero(assigned)),
        "assigned_below_threshold_final": int(np.count_nonzero(assigned & ~released)),
        "released_particles": int(np.count_nonzero(released)),
        "directional_contact_work_j": float(solver.work.numpy()[0]),
        "gpu_errors": solver.errors.numpy().tolist(),
        "minimum_det_f_sampled": minimum_det,
    }


def run(output, scene="fig3", **overrides):
    """Run one drop into the new directory `output` and return its result.

    scene: a name in SCENES or a DropScene; overrides: DropScene fields to change (for example
    duration=0.1, size=(0.06, 0.04, 0.03)) or Config fields (collected into `extra`).
    """
    from .solver import Solver

    scene = SCENES[scene] if isinstance(scene, str) else scene
    own = {f.name for f in fields(DropScene)}
    extra = dict(scene.extra, **{k: v for k, v in overrides.items() if k not in own})
    scene = replace(scene, extra=extra, **{k: v for k, v in overrides.items() if k in own})
    cfg = scene.config()
    steps = scene.steps()
    out = Path(output)
    out.mkdir(parents=True, exist_ok=False)
    solver = Solver(cfg)
    q0, height = release_pose(scene, cfg.size[2])
    solver.tool_rot.assign([q0])
    solver.tool_pos.assign([[0.0, 0.0, height]])
    manifest = {
        "status": "running",
        "experiment": scene.name,
        "config": asdict(cfg),
        "device": solver.device.name,
        "published_values": list(scene.sources),
        "independent_choices": list(scene.choices) + [f"tilt {scene.tilt_deg:g} deg about the knife "
                                                      f"{scene.tilt_axis}"],
        "dt_is_table_value": scene.name == "fig3_freefall" and scene.dt == 0.002,
        "actuation": "none: gravity and the reaction of the material and the board",
        "surface": "particle positions only (no surface mesh)",
        "revision": git_revision(),
        "command": sys.argv,
    }
    log = StateLog(solver.state())
    chunk = max(1, int(round(scene.snapshot_every / scene.dt)))
    minimum_det = 1.0
    t0 = time.perf_counter()
    try:
        done = 0
        while done < steps:
            count = min(chunk, steps - done)
            # A free knife follows no target: the targets are placeholders.
            solver.advance(np.zeros((count, 3)), cutting=True)
            done += count
            state = solver.state()
            minimum_det = min(minimum_det, float(np.linalg.det(state["deformation"]).min()))
            log.add(state)
        torch.cuda.synchronize()
        manifest.update(status="complete", wall_s_excluding_jit=time.perf_counter() - t0)
    except Exception as exc:
        manifest.update(status="failed", error=repr(exc), wall_s_excluding_jit=time.perf_counter() - t0)
        raise
    finally:
        manifest.update(completed_steps=solver.step_count, minimum_det_f_sampled=minimum_det)
        log.save(out / "states.npz", solver.reference)
        write(out / "manifest.json", manifest)
    result = outcome(solver, log, q0, manifest["wall_s_excluding_jit"], minimum_det)
    write(out / "result.json", result)
    return result


def _value(text):
    """A --set value: a number, a comma-separated tuple of numbers, true/false, or a string."""
    if text.lower() in ("true", "false"):
        return text.lower() == "true"
    try:
        parts = [float(v) for v in text.split(",")]
        parts = [int(v) if v.is_integer() and "." not in text and "e" not in text.lower() else v for v in parts]
        return tuple(parts) if "," in text else parts[0]
    except ValueError:
        return text


def main(argv=None):
    parser = argparse.ArgumentParser(description="Free knife drop onto a block (writes states.npz, manifest.json "
                                     "and result.json).")
    parser.add_argument("output", help="new directory for the run's files")
    parser.add_argument("--scene", choices=sorted(SCENES), default="fig3")
    parser.add_argument("--set", action="append", default=[], metavar="FIELD=VALUE",
                        help="change a DropScene or Config field, e.g. --set duration=0.2 --set tilt_axis=length")
    args = parser.parse_args(argv)
    overrides = {}
    for item in args.set:
        key, sep, text = item.partition("=")
        if not sep:
            parser.error(f"--set expects FIELD=VALUE, got {item!r}")
        overrides[key] = _value(text)
    print(run(args.output, args.scene, **overrides), flush=True)


if __name__ == "__main__":
    main()
