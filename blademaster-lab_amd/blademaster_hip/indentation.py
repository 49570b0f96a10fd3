"""Knife indentation without cutting: the banana's stiffness on a grid and the relaxation of the
held force (the reference's scripts/indentation.py).

The level knife is pressed into the top of the paper banana at a constant speed, one 1 ms
coupling interval at a time through `ExternalToolPort`, with cutting off; `press` returns the
vertical force at 2, 4 and 6 mm depth (linear interpolation over the intervals) and, with
`hold` > 0, the force while the knife then stays at its last depth.  `fit_relaxation` fits
F(t) = F_inf + (F_0 - F_inf) exp(-t / tau) to a held force.

Command line: ``python -m blademaster_hip.indentation --spacings 0.006 --seeds 1 --hold 0.35
--maxwell-tau 0.05 --maxwell-fraction 0.5``.
"""

import argparse
import json
from dataclasses import replace
from pathlib import Path

import numpy as np
from scipy.optimize import curve_fit

from .presets import banana_config, banana_dt

INTERVAL = 0.001  # s, one robot step
SPEED = 0.02  # m/s
START = 0.03  # m, first edge height
DEPTHS = (0.002, 0.004, 0.006)  # m


def indentation_config(spacing, seed, **fields):
    """The paper banana on a `spacing` grid, externally held knife, cached compact storage;
    `fields` change further Config fields (maxwell_tau, maxwell_fraction, ...)."""
    base = banana_config()
    return replace(base, grid_spacing=spacing, dt=banana_dt(base, spacing, INTERVAL), sampling_seed=seed,
                   external_tool=True, grid_storage="compact", cache_topology=True, **fields)


def press(spacing, seed, hold=0.0, speed=SPEED, config=None, **fields):
    """Press the knife to 6 mm below the banana's contact height and optionally hold it.

    Returns (forces at DEPTHS in N, contact height in m, final depth in m, held forces in N
    per interval).  The contact height is that of the highest particle within one spacing of
    the knife plane plus the knife's contact offset: contact starts when the edge reaches it.
    `config` replaces the banana configuration (the knife still moves as described).
    """
    from .coupling import ExternalToolPort, ToolState
    from .solver import Solver

    cfg = config if config is not None else indentation_config(spacing, seed, **fields)
    solver = Solver(cfg)
    port = ExternalToolPort(solver)
    half = cfg.blade_half_size[2]
    near = np.abs(solver.reference[:, 0]) < cfg.grid_spacing
    top = float(solver.reference[near, 2].max() + solver.blade_offset)
    bottom = top - max(DEPTHS)
    heights, forces = [], []
    edge = START
    while edge > bottom:
        edge = START - speed * INTERVAL * len(heights)
        state = ToolState([0.0, 0.0, edge + half], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, -speed], [0.0, 0.0, 0.0])
        forces.append(port.advance(state, INTERVAL, cutting=False).force_n[2])
        heights.append(edge)
    depth = top - np.asarray(heights)
    still = ToolState([0.0, 0.0, edge + half], [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    held = [port.advance(still, INTERVAL, cutting=False).force_n[2] for _ in range(int(round(hold / INTERVAL)))]
    solver.state()  # raises on device errors or an invalid state
    at = [float(np.interp(d, depth, forces)) for d in DEPTHS]
    return at, top, float(depth[-1]), np.asarray(held, dtype=np.float64)


def fit_relaxation(force, skip=0.02, interval=INTERVAL):
    """Least-squares fit of F_inf + (F_0 - F_inf) exp(-t / tau) to a held force sampled every
    `interval`, leaving out its first `skip` seconds (t restarts at the first kept sample)."""
    force = np.asarray(force, dtype=np.float64)
    t = interval * np.arange(1, len(force) + 1)
    keep = t > skip
    t, f = t[keep] - skip, force[keep]

    def curve(t, f_inf, f_0, tau):
        return f_inf + (f_0 - f_inf) * np.exp(-t / tau)

    (f_inf, f_0, tau), _ = curve_fit(curve, t, f, p0=(f[-1], f[0], t[-1] / 5), maxfev=20000)
    residual = curve(t, f_inf, f_0, tau) - f
    return dict(f_0=float(f_0), f_inf=float(f_inf), tau_s=float(tau), rms_n=float(np.sqrt(np.mean(residual**2))),
                skip_s=skip)


def main(argv=None):
    parser = argparse.ArgumentParser(description="Knife indentation of the paper banana without cutting.")
    parser.add_argument("--spacings", type=float, nargs="+", default=[0.006, 0.004, 0.003, 0.002])
    parser.add_argument("--seeds", type=int, default=3)
    parser.add_argument("--speed", type=float, default=SPEED, help="indentation speed, m/s")
    parser.add_argument("--hold", type=float, default=0.0, help="hold at the final depth, s")
    parser.add_argument("--maxwell-tau", type=float, default=0.0, help="Config.maxwell_tau, s")
    parser.add_argument("--maxwell-fraction", type=float, default=0.0, help="Config.maxwell_fraction")
    parser.add_argument("--output", help="JSON file for the held force traces and their fits")
    args = parser.parse_args(argv)
    traces = []
    for spacing in args.spacings:
        runs = [press(spacing, seed, args.hold, args.speed, maxwell_tau=args.maxwell_tau,
                      maxwell_fraction=args.maxwell_fraction) for seed in range(1, args.seeds + 1)]
        forces = np.array([r[0] for r in runs])
        text = ", ".join(f"{1000 * d:g} mm {m:.2f}+-{s:.2f} N" for d, m, s in zip(DEPTHS, forces.mean(0), forces.std(0)))
        print(f"{1000 * spacing:g} mm grid: top {1000 * np.mean([r[1] for r in runs]):.1f} mm, force at {text}",
              flush=True)
        for seed, (_, _, depth, held) in enumerate(runs, start=1):
            if len(held) == 0:
                continue
            fit = fit_relaxation(held)
            traces.append(dict(spacing=spacing, seed=seed, depth_m=depth, hold_force_n=held.tolist(), **fit))
            print(f"  seed {seed}: held at {1000 * depth:.2f} mm, {held[0]:.3f} -> {held[-1]:.3f} N; "
                  f"F0 {fit['f_0']:.3f} N, F_inf {fit['f_inf']:.3f} N, tau {fit['tau_s']:.3f} s, "
                  f"rms {fit['rms_n']:.3f} N", flush=True)
    if args.output:
        Path(args.output).write_text(json.dumps(traces))


if __name__ == "__main__":
    main()
