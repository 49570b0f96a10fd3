"""Throughput of the TLMPM cutting substep on one MI355X.

Workload (N=1): a 200 x 100 x 50 = 1,000,000-particle block on a 1.5 mm lattice, 3 mm grid
(133,926 nodes), fixed-corotated elasticity with E = 10**U(4, 6) drawn from seed 0 and
nu = 0.3, a blade tilted 10 degrees about its thickness axis plunging at 0.1 m/s
(prescribed kinematics: external tool re-posed every 1 ms coupling interval, integrated at
its twist inside the interval), 25 us substeps, cutting on.  One bench step is one substep
of all particles; intervals of 40 substeps are replayed as HIP graphs.  The blade's lowest
edge point starts 2 mm above the block and reaches it after about 800 substeps; the first
labelled (cut) particles appear after about 1,560.  The state the timed substeps start from
is made by running `--lead` substeps (1,600 by default) untimed, so warm-up and timing fall
in the cut, where contact, relabelling, multi-field nodes and partner-list rebuilds run.

Prints one JSON line with particle-substeps/s, the HBM roofline of the dominant kernel
(live HIP-event timing of that kernel on the bench stream) and the CPU oracle timed on a
bounded sample of the same workload on this host.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "blademaster-lab_amd"))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle of each kernel: the per-particle arrays of the substep state
# (x, v, F, Fp, stress, codes, friction histories) that the kernel must read or write, each
# array counted once per kernel.  On this regular lattice the stencil is implied by a
# particle's cell slot, so no kernel reads stencil arrays; Fp is not read while every Fp is the
# identity (no plastic flow; a device flag).  Intermediates between kernels (contact
# impulses, grid partials, the particle bins and field slots) are not charged.
#   k_rigid       x, v read 24; friction histories read 24 + written 24            = 72
#   k_pairs       codes read 8                                                      = 8
#   k_p2g         v read 12; stress read 36                                         = 48
#   k_g2p         x, v read 24 + written 24; codes read 8                           = 56
#   k_update_F    F read 36 + written 36; stress written 36                         = 108
#   k_p2g_update  k_p2g with the previous substep's F update inside it: v read 12, F read 36 +
#                 written 36; the stress it computes is used at once, neither read nor
#                 written                                                           = 84
ALG_BYTES = {"k_rigid": 72, "k_pairs": 8, "k_p2g": 48, "k_g2p": 56, "k_update_F": 108, "k_p2g_update": 84}
# Per particle-substep in the reference's layout (x, v, F, Fp, stress, codes, histories and the
# stored stencil): the substep-level figure, hbm_peak_frac.
SUBSTEP_BYTES = 444
INTERVAL_SUBSTEPS = 40


def interval_schedule(steps, warmup, per=INTERVAL_SUBSTEPS):
    """Interval lengths, in substeps, of the untimed warm-up and of the timed region.

    The timed region is exactly `steps` substeps: intervals of `per` substeps and one shorter
    interval for the remainder.  Each interval length is one captured HIP graph, cached under
    (length, step parity), so the warm-up runs every timed length at the parity it is timed
    at, and runs at least `warmup` substeps: the remainder first (twice when it is odd, which
    restores the parity), then intervals of `per`.  No capture falls in the timed region.
    """
    if steps < 1:
        raise ValueError("steps must be at least 1")
    if warmup < 0:
        raise ValueError("warmup must be non-negative")
    full, rest = divmod(steps, per)
    timed = [per] * full + ([rest] if rest else [])
    step = 2 if rest % 2 else 1  # remainder intervals in pairs when odd
    warm = [rest] * step if rest else []
    if full:
        warm += [per] * max(1, -(-(warmup - sum(warm)) // per))
    else:
        while sum(warm) < warmup:
            warm += [rest] * step
    return warm, timed


_SLEEP_CYCLES_PER_MS = []


def hold_stream(stream, ms=3.0):
    """Keep `stream` busy for about `ms` milliseconds with a bounded spin kernel.

    Enqueued just before a timed region's start event, it lets the host enqueue the first
    timed interval (tool-state staging, graph launch) while the device still spins, so the
    events time the device's work and not the host's first enqueue.  Later intervals are
    enqueued while earlier ones run (the host is many times faster than an interval).
    """
    import torch

    if not _SLEEP_CYCLES_PER_MS:
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        probe = 1_000_000
        torch.cuda._sleep(probe)
        a.record(stream)
        torch.cuda._sleep(probe)
        b.record(stream)
        torch.cuda.synchronize()
        _SLEEP_CYCLES_PER_MS.append(probe / max(a.elapsed_time(b), 1e-3))
    torch.cuda._sleep(int(ms * _SLEEP_CYCLES_PER_MS[0]))


def workload(seed=0, n_side=(0.3, 0.15, 0.075)):
    from blademaster_hip import Config

    rng = np.random.default_rng(seed)
    E = float(10 ** rng.uniform(4, 6))
    cfg = Config(
        size=n_side,
        grid_spacing=0.003,
        dt=25e-6,
        young_modulus=E,
        poisson_ratio=0.3,
        external_tool=True,
        tool_substep="predict",
        cache_topology=True,
        grid_storage="compact",
        blade_half_size=(0.0005, 0.08, 0.03),
    )
    tilt = np.deg2rad(10.0)
    q = np.array([np.sin(tilt / 2), 0.0, 0.0, np.cos(tilt / 2)])
    hz = cfg.blade_half_size[2]
    # lowest edge point 2 mm above the top face
    edge_z0 = cfg.size[2] + 0.002 + cfg.blade_half_size[1] * np.sin(tilt)
    return cfg, q, edge_z0, hz, E


def pose_at(t, q, edge_z0, hz, speed=0.1):
    from scipy.spatial.transform import Rotation

    edge = np.array([0.0, 0.0, edge_z0 - speed * t])
    origin = edge + Rotation.from_quat(q).apply([0.0, 0.0, hz])
    return origin


def run_gpu(args):
    import torch
    from blademaster_hip import Solver, ExternalToolPort, ToolState
    from blademaster_hip import _lib

    cfg, q, edge_z0, hz, E = workload()
    s = Solver(cfg, brick_z=args.brick_z)
    port = ExternalToolPort(s)
    vel = np.array([0.0, 0.0, -0.1])
    warm, timed = interval_schedule(args.steps, args.warmup)
    clock = [0]  # substeps submitted so far: the start time of the next interval

    def submit(length):
        state = ToolState(pose_at(clock[0] * cfg.dt, q, edge_z0, hz), q, vel, [0.0, 0.0, 0.0])
        port.submit(state, length * cfg.dt, cutting=True, cut_index=0)
        clock[0] += length
        return state

    # The plunge up to the timed window: untimed, checked for device errors at its end.
    lead = [INTERVAL_SUBSTEPS] * (args.lead // INTERVAL_SUBSTEPS)
    lead += [args.lead % INTERVAL_SUBSTEPS] if args.lead % INTERVAL_SUBSTEPS else []
    for length in lead:
        last = submit(length)
    if lead:
        port.collect(last, lead[-1] * cfg.dt, 2)
    for length in warm:
        submit(length)
    torch.cuda.synchronize()
    start_state = snapshot(s) if not args.no_cpu else None
    start_codes = int(torch.count_nonzero(s._t["codes"]).item())
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    hold_stream(stream)
    start.record(stream)
    for length in timed:
        last = submit(length)
    stop.record(stream)
    torch.cuda.synchronize()
    gpu_ms = start.elapsed_time(stop)
    steps_timed = sum(timed)
    reaction = port.collect(last, timed[-1] * cfg.dt, 2)
    st = s.state()
    ms_per_step = gpu_ms / steps_timed
    rate = s.n * steps_timed / (gpu_ms * 1e-3)
    # Live per-kernel timing of the dominant kernels on the bench stream.
    # k_p2g_update is k_p2g with the previous substep's F update pending, as it runs on all
    # but the last substep of a recorded interval.
    kernels = {
        "k_rigid": ("bm_rigid_contacts", (1,), 0),
        "k_p2g": ("bm_p2g", (), 0),
        "k_p2g_update": ("bm_p2g", (), 1),
        "k_g2p": ("bm_g2p_remap", (), 0),
        "k_update_F": ("bm_update_F", (), 0),
    }
    per_kernel_ms = {}
    reps = args.kernel_reps
    saved = {name: t.clone() for name, t in s._t.items() if name in ("x", "v", "F", "Fp", "stress", "impulses",
                                                                        "histories", "damage", "codes", "work")}
    for kname, (entry, extra, pending) in kernels.items():
        s._t["update_pending"].fill_(pending)
        # The launches are captured in a HIP graph and the replay is timed, so the figure is
        # the kernels' execution time without the host's launch gaps.
        s.launch(entry, *extra)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(reps):
                s.launch(entry, *extra)
        g.replay()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        g.replay()
        b.record(stream)
        torch.cuda.synchronize()
        per_kernel_ms[kname] = a.elapsed_time(b) / reps
        del g
    s._t["update_pending"].fill_(0)
    for name, t in saved.items():
        s._t[name].copy_(t)
    dominant = max(per_kernel_ms, key=per_kernel_ms.get)
    t_dom = per_kernel_ms[dominant] * 1e-3
    achieved = ALG_BYTES[dominant] * s.n / t_dom / 1e9
    # HBM bytes per launch are hardware counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): they
    # cannot be read inside this process, so they come from the committed counter profile of the
    # same workload, named in traffic_source.
    traffic = traffic_source = None
    if args.traffic_file and os.path.exists(args.traffic_file):
        profile = json.load(open(args.traffic_file))
        entry = profile.get("kernels", {}).get(dominant)
        if entry:
            traffic = entry["total"]
            traffic_source = os.path.relpath(args.traffic_file, ROOT)
    return dict(
        solver=s,
        cfg=cfg,
        E=E,
        rate=rate,
        ms_per_step=ms_per_step,
        gpu_ms=gpu_ms,
        steps=steps_timed,
        warmup=sum(warm),
        per_kernel_ms=per_kernel_ms,
        dominant=dominant,
        achieved=achieved,
        traffic=traffic,
        traffic_source=traffic_source,
        reaction=reaction,
        codes_nonzero=int(np.count_nonzero(st["sides"])),
        codes_at_start=start_codes,
        lead=sum(lead),
        start_state=start_state,
        start_time=(sum(lead) + sum(warm)) * cfg.dt,
    )


# Per-particle arrays and knife state copied from the solver into the oracle for the CPU
# baseline, so that it runs the same part of the plunge as the timed substeps.
SNAPSHOT = ("x", "v", "F", "Fp", "stress", "damage", "histories", "codes", "tool_pos", "tool_vel", "tool_rot",
            "tool_omega")


def snapshot(s):
    return {name: getattr(s, name).numpy().copy() for name in SNAPSHOT}


def run_cpu_baseline(cfg, q, edge_z0, hz, substeps, threads=1, state=None):
    sys.path.insert(0, ROOT)
    from oracle import OracleSim
    from blademaster_hip.geometry import points, stencil
    from blademaster_hip.config import max_cut_index

    ref, origin, cells = points(cfg.size, cfg.grid_spacing, cfg.shape, cfg.sampling_seed)
    nodes, w, g = stencil(ref, origin, cfg.grid_spacing, cells)
    node_count = int(np.prod(cells + 1))
    o = OracleSim(cfg, ref, nodes, w, g, node_count, 32, max_cut_index(node_count))
    threads = o.set_threads(threads)
    o.set("tool_pos", pose_at(0.0, q, edge_z0, hz))
    o.set("tool_rot", q)
    o.set("tool_vel", [0.0, 0.0, -0.1])
    for name, value in (state or {}).items():
        o.set(name, value)
    o.substep(cutting=True)  # first substep registers the topology (setup)
    t0 = time.perf_counter()
    for _ in range(substeps):
        o.substep(cutting=True)
    dt = time.perf_counter() - t0
    return len(ref) * substeps / dt, dt, threads


def run_batch(args):
    """Batched sweep of uneven instances (256 by default) in one solver: particle-substeps/s
    over all instances, brick height chosen by the autotuner on a 16-instance subset."""
    import torch
    from blademaster_hip.batch import BatchPort, BatchSolver, autotune_brick_z, sweep_instances

    configs, knives, gap = sweep_instances(args.instances, seed=0)
    brick_z, timings = autotune_brick_z(configs[:16], substeps=20) if args.brick_z is None else (args.brick_z, {})
    s = BatchSolver(configs, brick_z=brick_z, gap=gap)
    port = BatchPort(s)
    dt = configs[0].dt
    warm, timed = interval_schedule(args.steps, args.warmup)
    clock = [0]

    def submit(length):
        states = [kn(clock[0] * dt) for kn in knives]
        port.submit(states, length * dt)
        clock[0] += length
        return states

    for length in warm:
        submit(length)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    hold_stream(stream)
    a.record(stream)
    for length in timed:
        last = submit(length)
    b.record(stream)
    torch.cuda.synchronize()
    steps = sum(timed)
    gpu_ms = a.elapsed_time(b)
    reactions = port.collect(last, timed[-1] * dt)
    rate = s.n * steps / (gpu_ms * 1e-3)
    cut = int(np.count_nonzero(s.codes.numpy()))
    return {
        "metric": "particle-substeps/s",
        "value": rate,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": sum(warm),
        "ms_per_step": gpu_ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched sweep: {args.instances} uneven blocks (10-60 x 10-40 x 8-40 particles on a "
            "1.5 mm lattice, 3 mm grid), E log-uniform 1e4-1e6 Pa, nu 0.25-0.45, knives tilted +-15 deg, "
            "offset 0-h/2, plunging 0.02-0.2 m/s, dt 25 us, one solver",
            "instances": args.instances,
            "particles": s.n,
            "grid_nodes": s.node_count,
            "active_nodes": s.active_nodes or s.node_count,
            "brick": [int(v) for v in s.brick],
            "autotune_us_per_substep": {str(k): round(v * 1e6, 1) for k, v in timings.items()},
        },
        "cut_particles": cut,
        "mean_abs_force_n": float(np.mean([np.abs(r.force_n).max() for r in reactions])),
    }


# MLS-MPM baseline (mlsmpm.py), per kernel the per-particle arrays it must read or write, as
# above: contact reads x, v and reads/writes the friction histories (72); P2G reads x, v, C, F
# and the damage flag (100); G2P reads x, F and the damage reference, and writes x, v, C and F
# (148).  Per particle-substep: x, v 24 r + 24 w; C, F 72 r + 72 w; damage and its reference
# 8 r; friction histories 24 r + 24 w = 248 B.
MLS_ALG_BYTES = {"k_rigid": 72, "k_mls_p2g": 100, "k_mls_g2p": 148}
MLS_SUBSTEP_BYTES = 248
# Node updates per particle of the MLS P2G: 27 nodes x (mass + 3 momentum components), summed
# per workgroup in LDS before one global atomic per node and component.
MLS_P2G_UPDATES = 108


def run_mls(args):
    """The same 1M-particle block and knife through the MLS-MPM baseline: the knife pose of
    each 1 ms interval held over its 40 substeps (MlsSolver.advance, mlsmpm.py:395-426)."""
    import ctypes
    import torch
    from blademaster_hip import ToolState
    from blademaster_hip.mlsmpm import MlsSolver

    cfg, q, edge_z0, hz, E = workload()
    s = MlsSolver(cfg)
    vel = np.array([0.0, 0.0, -0.1])
    warm, timed = interval_schedule(args.steps, args.warmup)
    clock = [0]

    def advance(length, check=True):
        state = ToolState(pose_at(clock[0] * cfg.dt, q, edge_z0, hz), q, vel, [0.0, 0.0, 0.0])
        clock[0] += length
        return s.advance(state, length * cfg.dt, check=check)

    for length in warm:
        advance(length)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    gpu_ms = 0.0
    for length in timed:
        # Each interval ends with the reaction's readback: time the device work of each
        # interval and add them (the domain check runs once, after the timed intervals).
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        hold_stream(stream)
        a.record(stream)
        advance(length, check=False)
        b.record(stream)
        torch.cuda.synchronize()
        gpu_ms += a.elapsed_time(b)
    steps = sum(timed)
    rate = s.n * steps / (gpu_ms * 1e-3)
    lost = s.diagnostics()["left_domain"]
    mp, ms = ctypes.byref(s._mls_params), ctypes.byref(s._mls_state)
    lib = s._lib
    entries = {
        "k_rigid": lambda st: lib.bm_rigid_contacts(ctypes.byref(s._params), ctypes.byref(s._state), 0, st),
        "k_mls_p2g": lambda st: lib.bm_mls_p2g(mp, ms, st),
        "k_mls_grid": lambda st: lib.bm_mls_grid_update(mp, ms, st),
        "k_mls_g2p": lambda st: lib.bm_mls_g2p(mp, ms, st),
    }
    saved = {k: t.clone() for k, t in s._t.items()}
    per_kernel_ms = {}
    for name, fn in entries.items():
        st = ctypes.c_void_p(stream.cuda_stream)
        fn(st)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            cap = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
            for _ in range(args.kernel_reps):
                # P2G adds into the grid: clear it between launches as the grid update does.
                fn(cap)
                if name == "k_mls_p2g":
                    lib.bm_mls_grid_update(mp, ms, cap)
        g.replay()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        g.replay()
        b.record(stream)
        torch.cuda.synchronize()
        per_kernel_ms[name] = a.elapsed_time(b) / args.kernel_reps
        del g
        for k, t in saved.items():
            s._t[k].copy_(t)
    per_kernel_ms["k_mls_p2g"] -= per_kernel_ms["k_mls_grid"]
    dominant = max(MLS_ALG_BYTES, key=lambda k: per_kernel_ms[k])
    achieved = MLS_ALG_BYTES[dominant] * s.n / (per_kernel_ms[dominant] * 1e-3) / 1e9
    atomics = MLS_P2G_UPDATES * s.n / (per_kernel_ms["k_mls_p2g"] * 1e-3)
    return {
        "metric": "particle-substeps/s",
        "value": rate,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": sum(warm),
        "ms_per_step": gpu_ms / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "MLS-MPM baseline: block 200x100x50 = 1M particles, 1.5 mm lattice, 3 mm dense grid, "
            "blade 10 deg tilt re-posed each 1 ms (0.1 m/s plunge), dt 25 us",
            "particles": s.n,
            "grid_nodes": int(np.prod(s.dims)),
            "young_modulus": E,
        },
        "hbm_peak_frac": MLS_SUBSTEP_BYTES * rate / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_particle": MLS_ALG_BYTES[dominant]},
        "p2g_node_updates_per_s": atomics,
        "kernel_ms": {k: round(v, 5) for k, v in per_kernel_ms.items()},
        "left_domain": lost,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--lead", type=int, default=1600,
                    help="untimed substeps of the plunge before warm-up and timing (the cut starts near 1,560)")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--cpu-substeps", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_1M_block.json"))
    ap.add_argument("--brick-z", type=int, default=None, help="brick height of the transfer kernels (cells)")
    ap.add_argument("--batch", action="store_true",
                    help="measure the batched sweep of uneven instances instead of the 1M-particle block")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--mls", action="store_true", help="measure the MLS-MPM baseline on the 1M-particle block")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB three times (the calibration kernel of "
                    "HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the substep on a single device; "
              "--gpus N with N > 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.batch:
        print(json.dumps(run_batch(args)), flush=True)
        return
    if args.mls:
        print(json.dumps(run_mls(args)), flush=True)
        return
    r = run_gpu(args)
    s, cfg = r["solver"], r["cfg"]
    if args.copy_mib > 0:
        import torch

        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cpu = None
    if not args.no_cpu:
        _, q, edge_z0, hz, _ = workload()
        # The host's share of cores (OMP_NUM_THREADS where the environment sets it), and one.
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        start = r["start_state"]
        many_rate, many_s, threads = run_cpu_baseline(cfg, q, edge_z0, hz, args.cpu_substeps, threads, start)
        one_rate, one_s, _ = run_cpu_baseline(cfg, q, edge_z0, hz, args.cpu_substeps, 1, start)
        cpu = dict(
            value=many_rate,
            unit="particle-substeps/s",
            cores=threads,
            kind="port",
            sample=f"{args.cpu_substeps} substeps of the same 1M-particle plunge from the state the timed GPU "
            f"substeps start from (t = {1000 * r['start_time']:.1f} ms), after one setup substep; float32 oracle "
            f"on {threads} OpenMP threads ({many_s:.1f} s) and on one thread ({one_s:.1f} s)",
            one_thread=dict(value=one_rate, cores=1),
        )
    peak = HBM_PEAK_GBS
    line = {
        "metric": "particle-substeps/s",
        "value": r["rate"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": r["warmup"],
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "block 200x100x50 = 1M particles, 1.5 mm lattice, 3 mm grid, blade 10 deg tilt plunging "
            "0.1 m/s, dt 25 us, cutting; timed substeps start after the lead-in, in the cut",
            "lead_substeps": r["lead"],
            "timed_from_s": r["start_time"],
            "particles": s.n,
            "grid_nodes": s.node_count,
            "young_modulus": r["E"],
            "poisson_ratio": cfg.poisson_ratio,
            "interval_substeps": INTERVAL_SUBSTEPS,
            "brick": [int(v) for v in s.brick],
        },
        "hbm_peak_frac": SUBSTEP_BYTES * r["rate"] / (peak * 1e9),
        "sim_per_device_s": r["steps"] * cfg.dt / (r["gpu_ms"] * 1e-3),
        "roofline": {
            "bound": "hbm",
            "kernel": r["dominant"],
            "achieved": r["achieved"],
            "peak": peak,
            "unit": "GB/s",
            "frac": r["achieved"] / peak,
            "traffic": r["traffic"],
            "traffic_source": r["traffic_source"],
            "alg_bytes_per_particle": ALG_BYTES[r["dominant"]],
        },
        "kernel_ms": {k: round(v, 5) for k, v in r["per_kernel_ms"].items()},
        "cut_particles": r["codes_nonzero"],
        "cut_particles_at_start": r["codes_at_start"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
