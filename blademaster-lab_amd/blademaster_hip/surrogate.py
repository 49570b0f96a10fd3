"""Learned interval-wrench predictor for robot couplings wider than one robot step.

The reference exchanges the wrench with the arm once per robot step (`franka_run.py:394-422`):
the material runs one interval from the arm's state and the arm then integrates with that
interval's wrench.  Inside an interval its only predictor is the mobility matrix of
`tool_substep="predict"` (`tool.py:97-111`), which moves the knife by its response to the
contact impulses.  Widening the exchange to a window of W robot steps needs the wrench at every
step of the window before the material has run it.  Without a learned model the last measured
wrench is held over the window ("hold"); here a small MLP predicts it ("mlp").

One window (`CutRun.window_step`):
1. The material interval 0 is enqueued from the arm state at the window start.  While it runs,
   the predictor gives the wrenches w^_k of all W steps in one inference, from the arm state
   k steps into the window (along the path the held wrench would give), the last measured
   wrench and the occupancy/damage patch (both from the window start, k steps old) and the
   material.  Then for k = 0..W-1 the arm steps with w^_k and interval k + 1 is enqueued from
   the arm state at its start.
2. One readback returns the W measured wrenches w_k: the full two-way TLMPM coupling along the
   same knife path judges the prediction.  The residual of an instance is
   |mean_k (F^_k - F_k)| / max(|mean_k F_k|, floor) on the force.
3. Above the threshold (5 %) the instance falls back: its arm is re-integrated over the window
   with the measured wrenches, as a 1-step coupling would have applied them.

Network input (96 floats):
    [0:3] tool displacement since the wrench was measured (since the window start),
    [3:6] rotation vector, [6:9] velocity, [9:12] angular velocity,
    [12:18] last measured wrench (force, torque about the tool origin),
    [18:24] the wrench measured one robot step before it,
    [24:88] particle counts of the 8x8x8 edge patch summed over 2x2x2 blocks (4x4x4),
    [88:94] mean damage in 2 (along the edge) x 3 (below / at / above the edge) bands,
    [94] log10 Young's modulus, [95] age of the wrench and patch in robot steps (0..W-1).
Output (6): the change of the knife's force and torque since the last measured wrench; the
prediction is that wrench plus the output, so a zero output is the hold predictor.

The 8x8x8 patch of the task statement would be 512 values per channel; 96 inputs hold it only
pooled, as above.
"""

from dataclasses import replace
import time

import numpy as np
import torch

from .arm import ImpedanceArm, rotation_matrix, rotvec_of
from .batch import BatchPort, BatchSolver
from .mlp import Mlp
from .presets import banana_config

INTERVAL = 0.001
INPUTS, HIDDEN, OUTPUTS = 96, 256, 6
THRESHOLD = 0.05
# Predicted force changes below this (N) are taken as zero (chosen on the validation tasks).
SHRINK = 0.15
PREV = slice(12, 18)
PATCH_FEATURES = 70
# Smallest standard deviation used to standardise each input: the physical resolution of the
# feature (displacement 10 um, angle 0.01 rad, velocity 1 mm/s, angular velocity 0.01 rad/s, force
# 0.05 N, torque 0.001 N m, one particle, damage 0.05, 0.05 decades of modulus, one step).
INPUT_FLOOR = np.concatenate([[1e-5] * 3, [1e-2] * 3, [1e-3] * 3, [1e-2] * 3, ([0.05] * 3 + [1e-3] * 3) * 2,
                              [1.0] * 64, [0.05] * 6, [0.05, 1.0]]).astype(np.float32)
MAX_AGE = 4
OUTPUT_FLOOR = np.array([0.01] * 3 + [1e-4] * 3, dtype=np.float32)


# ---------------------------------------------------------------------------- features
def patch_features(patches):
    """[..., 2, 8, 8, 8] patch (counts, damage sums) -> [..., 70] pooled features (device)."""
    lead = patches.shape[:-4]
    count, damage = patches[..., 0, :, :, :], patches[..., 1, :, :, :]
    occupancy = count.reshape(*lead, 4, 2, 4, 2, 4, 2).sum((-5, -3, -1)).reshape(*lead, 64)

    def bands(t):
        t = t.reshape(*lead, 8, 2, 4, 8).sum((-4, -2))  # along-edge halves x 8 z cells
        return torch.stack([t[..., 0:3].sum(-1), t[..., 3:5].sum(-1), t[..., 5:8].sum(-1)], -1)

    mean_damage = bands(damage) / bands(count).clamp_min(1.0)
    return torch.cat([occupancy, mean_damage.reshape(*lead, 6)], -1)


def pose_features(p, q, v, w):
    """[instances][12]: position, rotation vector, velocity, angular velocity.  The network
    sees the position relative to the one at the wrench's measurement (`samples`,
    `CutRun.features`), not the absolute one."""
    return np.concatenate([p, rotvec_of(q), v, w], axis=1)


def assemble(pose, wrench, wrench_before, patch, log_modulus, age):
    """[rows][96] network inputs; age: robot steps since `wrench` and `patch` were measured."""
    rows = len(pose)
    return np.concatenate([pose, wrench, wrench_before, patch, np.broadcast_to(log_modulus, (rows,))[:, None],
                           np.broadcast_to(age, (rows,))[:, None]], axis=1).astype(np.float32)


# ---------------------------------------------------------------------------- the task
class ThreeCutTask:
    """Three cuts through the banana by an impedance-controlled knife, one per time slot.

    Per instance: cut positions (one in each third of the banana), descent speed, roll tilt
    about x and lateral offset.  In each slot of `slot_s` the edge target descends from
    `z_start` to `z_end` (below the board top, so the cut separates) at the instance's speed
    and holds there (cutting on, cut index = slot), then rises in `lift_s` and moves to the
    next cut position in `move_s` (cutting off)."""

    def __init__(self, count, seed=0, half_z=0.02, z_start=0.03, z_end=-0.004, descend_s=0.8, lift_s=0.15,
                 move_s=0.15, cuts=3):
        rng = np.random.default_rng(seed)
        thirds = [(-0.035, -0.02), (-0.008, 0.008), (0.02, 0.035)][:cuts]
        self.cut_x = np.stack([rng.uniform(a, b, count) for a, b in thirds], 1)
        self.speed = rng.uniform(0.06, 0.12, count)
        tilt = np.deg2rad(rng.uniform(-5.0, 5.0, count))
        self.orientation = np.stack([np.sin(tilt / 2), 0 * tilt, 0 * tilt, np.cos(tilt / 2)], 1)
        self.y = rng.uniform(-0.005, 0.005, count)
        self.half_z = half_z
        self.z_start, self.z_end = z_start, z_end
        self.descend_s, self.lift_s, self.move_s = descend_s, lift_s, move_s
        self.slot_s = descend_s + lift_s + move_s
        self.cuts = cuts
        self.duration = cuts * self.slot_s

    def at(self, t):
        """Tool-origin target positions (I, 3), orientations (I, 4), cutting, cut index."""
        k = min(int(t // self.slot_s), self.cuts - 1)
        tau = t - k * self.slot_s
        x = self.cut_x[:, k].copy()
        if tau < self.descend_s:
            z = np.maximum(self.z_start - self.speed * tau, self.z_end)
            cutting = True
        elif tau < self.descend_s + self.lift_s:
            z = np.full_like(x, self.z_end + (self.z_start - self.z_end) * (tau - self.descend_s) / self.lift_s)
            cutting = False
        else:
            z = np.full_like(x, self.z_start)
            if k + 1 < self.cuts:
                f = min((tau - self.descend_s - self.lift_s) / self.move_s, 1.0)
                x = x + (self.cut_x[:, k + 1] - x) * f
            cutting = False
        edge = np.stack([x, self.y, z], 1)
        origin = edge + rotation_matrix(self.orientation)[:, :, 2] * self.half_z
        return origin, self.orientation, cutting, k


def task_configs(count, seed=0):
    """Banana instances (6 mm grid, 62.5 us substeps) with Young's modulus log-uniform in
    [2e5, 6e5] Pa, Poisson's ratio uniform in [0.35, 0.45] and their own sampling seeds.  The
    side bands of neighbouring cuts (6 cells either side of the blade) overlap, and a knife
    that stalls leaves uncut particles under its edge, so a node can meet 2 x 3 x 2 codes: 16
    field slots per node (the reference's linked lists have no cap)."""
    rng = np.random.default_rng(seed + 7919)
    base = replace(banana_config(), external_tool=True, tool_substep="predict", grid_storage="compact",
                   cache_topology=True, fields_per_node=16)
    return [replace(base, young_modulus=float(10 ** rng.uniform(np.log10(2e5), np.log10(6e5))),
                    poisson_ratio=float(rng.uniform(0.35, 0.45)), sampling_seed=int(rng.integers(0, 1 << 30)))
            for _ in range(count)]


# ---------------------------------------------------------------------------- coupled runs
class CutRun:
    """A batch of three-cut tasks coupled to impedance arms.

    window: robot steps per exchange (1 = full coupling every robot step).  net: an `Mlp`
    for the "mlp" predictor."""

    def __init__(self, count=1, seed=0, window=1, configs=None, task=None, device="cuda:0", brick_z=None,
                 arm_kw=None):
        self.configs = configs if configs is not None else task_configs(count, seed)
        count = len(self.configs)
        self.task = task if task is not None else ThreeCutTask(count, seed, half_z=self.configs[0].blade_half_size[2])
        self.solver = BatchSolver(self.configs, device=device, brick_z=brick_z)
        self.window = int(window)
        self.port = BatchPort(self.solver, slots=self.window, patch=True)
        origin, orientation, _, _ = self.task.at(0.0)
        self.arm = ImpedanceArm(origin, orientation, **(arm_kw or {}))
        self.log_modulus = np.array([np.log10(c.young_modulus) for c in self.configs])
        self.prev = np.zeros((count, 6))
        self.prev_before = np.zeros((count, 6))
        self.steps = 0
        dev = self.solver.device.torch
        self.port.load_records(self.arm.records())
        self.patch = patch_features(self.solver.blade_patch()).cpu().numpy().astype(np.float64)
        self._mlp_stream = torch.cuda.Stream(dev)
        self.log = dict(used=[], measured=[], predicted=[], residual=[], fallback=[])
        self.wall_s = 0.0

    @property
    def count(self):
        return len(self.configs)

    @property
    def sim_s(self):
        return self.steps * INTERVAL

    def features(self, age):
        a = self.arm
        pose = pose_features(a.p - self._measured_at, a.q, a.v, a.w)
        return assemble(pose, self.prev, self.prev_before, self.patch, self.log_modulus, age)

    def _collect(self, starts, slots):
        extra = patch_features(self.port.patches[slots[-1]])
        wrench, patch = self.port.collect_wrenches(starts, INTERVAL, slots=slots, extra=extra)
        self.patch = patch
        return wrench

    def full_step(self, record=None):
        """One robot step with full coupling (the reference's exchange).  record: a dict of
        lists that receives the pose features of this step, its measured wrench and the patch
        after it."""
        begin = time.perf_counter()
        origin, orientation, cutting, cut = self.task.at(self.steps * INTERVAL)
        if record is not None:
            a = self.arm
            record["pose"].append(pose_features(a.p, a.q, a.v, a.w).astype(np.float32))
        start = self.arm.p.copy()
        self.port.submit_records(self.arm.records(), INTERVAL, cutting=cutting, cut_index=cut, slot=0)
        w = self._collect(start[None], [0])[0]
        self.arm.step(origin, orientation, w)
        if record is not None:
            record["wrench"].append(w.astype(np.float32))
            record["patch"].append(self.patch.astype(np.float32))
        self.prev_before = self.prev
        self.prev = w
        self.steps += 1
        self.log["used"].append(w)
        self.log["measured"].append(w)
        self.wall_s += time.perf_counter() - begin
        return w

    def _predict_window(self, net, targets):
        """The network's wrenches for the W steps of a window, from one inference: the inputs
        of step k take the arm state k steps into the window along the path the held wrench
        would give (the arm's motion over a window depends little on the wrench it feels).
        A predicted change of the force below SHRINK newtons is taken as zero (hold): on
        validation tasks the network's error in steady contact is of that size."""
        saved = self.arm.save()
        rows = []
        for k, (origin, orientation) in enumerate(targets):
            rows.append(self.features(k))
            self.arm.step(origin, orientation, self.prev)
        self.arm.load(saved)
        x = np.concatenate(rows)
        if getattr(self, "_x_host", None) is None or self._x_host.shape != x.shape:
            self._x_host = torch.empty(x.shape, dtype=torch.float32, pin_memory=True)
        self._x_host.copy_(torch.from_numpy(x))
        x = self._x_host
        with torch.cuda.stream(self._mlp_stream):
            y = net.infer(x.to(net.device, non_blocking=True)).cpu().numpy().astype(np.float64)
        y = y.reshape(len(targets), self.count, OUTPUTS)
        small = np.linalg.norm(y[..., :3], axis=-1) < SHRINK
        y[small] = 0.0
        return self.prev[None] + y

    def window_step(self, predictor="mlp", net=None, threshold=THRESHOLD, floor=1.0):
        """One exchange window of `window` robot steps (see the module docstring)."""
        begin = time.perf_counter()
        W = self.window
        if predictor == "mlp" and W > MAX_AGE:
            raise ValueError(f"The network predicts at most {MAX_AGE} steps ahead")
        saved = self.arm.save()
        self._measured_at = self.arm.p.copy()
        plan = [self.task.at((self.steps + k) * INTERVAL) for k in range(W)]
        guesses = None
        preds, starts, targets = [], [], []
        for k in range(W):
            origin, orientation, cutting, cut = plan[k]
            starts.append(self.arm.p.copy())
            self.port.submit_records(self.arm.records(), INTERVAL, cutting=cutting, cut_index=cut, slot=k)
            if k == 0 and predictor != "hold":
                # The window's first interval starts from the arm's state now: it runs on the
                # GPU while the host predicts the window.
                guesses = self._predict_window(net, [(o, q) for o, q, _, _ in plan])
            w_hat = self.prev.copy() if guesses is None else guesses[k]
            self.arm.step(origin, orientation, w_hat)
            preds.append(w_hat)
            targets.append((origin, orientation))
        measured = self._collect(np.stack(starts), list(range(W)))
        preds = np.stack(preds)
        err = np.linalg.norm((preds - measured)[:, :, :3].mean(0), axis=1)
        scale = np.maximum(np.linalg.norm(measured[:, :, :3].mean(0), axis=1), floor)
        residual = err / scale
        fallback = residual > threshold
        used = preds.copy()
        if fallback.any():
            used[:, fallback] = measured[:, fallback]
            self.arm.load(saved)
            for k, (origin, orientation) in enumerate(targets):
                self.arm.step(origin, orientation, used[k])
        self.prev_before = measured[-2] if W > 1 else self.prev
        self.prev = measured[-1]
        self.steps += W
        self.log["used"].extend(used)
        self.log["measured"].extend(measured)
        self.log["predicted"].extend(preds)
        self.log["residual"].append(residual)
        self.log["fallback"].append(fallback)
        self.wall_s += time.perf_counter() - begin
        return residual

    def run(self, predictor=None, net=None, steps=None, threshold=THRESHOLD, floor=1.0):
        """The whole task (or `steps` robot steps): full coupling when window == 1 or no
        predictor is given, else windows with the predictor."""
        total = int(round(self.task.duration / INTERVAL)) if steps is None else int(steps)
        if net is not None:
            # The predictions run on a stream of their own: the weights must be final first.
            net.pack()
            torch.cuda.synchronize()
        while self.steps < total:
            if self.window == 1 or predictor is None:
                self.full_step()
            else:
                self.window_step(predictor, net, threshold, floor)
        return self.summary()

    def summary(self):
        used, measured = np.array(self.log["used"]), np.array(self.log["measured"])
        f_used, f_meas = used[..., :3], measured[..., :3]
        norm = np.linalg.norm(f_meas, axis=-1).sum()
        out = dict(instances=self.count, robot_steps=self.steps, sim_s=self.sim_s, wall_s=self.wall_s,
                   sim_per_wall=self.count * self.sim_s / max(self.wall_s, 1e-12),
                   force_error=float(np.linalg.norm(f_used - f_meas, axis=-1).sum() / max(norm, 1e-12)),
                   mean_force_n=float(np.linalg.norm(f_meas, axis=-1).mean()))
        W = self.window
        if W > 1 and len(f_used) % W == 0:
            # Impulse the arm received over each exchange window against the judged one.
            du = (f_used - f_meas).reshape(-1, W, *f_used.shape[1:]).sum(1)
            dm = f_meas.reshape(-1, W, *f_meas.shape[1:]).sum(1)
            out["window_force_error"] = float(np.linalg.norm(du, axis=-1).sum()
                                              / max(np.linalg.norm(dm, axis=-1).sum(), 1e-12))
        if self.log["predicted"]:
            pred = np.array(self.log["predicted"])[..., :3]
            out["predictor_error"] = float(np.linalg.norm(pred - f_meas, axis=-1).sum() / max(norm, 1e-12))
            fallback = np.array(self.log["fallback"])
            out["fallback_rate"] = float(fallback.mean())
            out["mean_residual"] = float(np.array(self.log["residual"]).mean())
        return out


# ---------------------------------------------------------------------------- data, training
def generate(count, seed=0, steps=None, device="cuda:0", brick_z=None, log=None):
    """Trajectories of `count` batched three-cut tasks under full coupling: per robot step
    the pose features (steps, count, 12), the measured wrench (steps, count, 6) and the patch
    after the step (steps, count, 76), plus the run."""
    run = CutRun(count, seed, window=1, device=device, brick_z=brick_z)
    total = int(round(run.task.duration / INTERVAL)) if steps is None else int(steps)
    record = dict(pose=[], wrench=[], patch=[run.patch.astype(np.float32)])
    for j in range(total):
        run.full_step(record)
        if log is not None and (j + 1) % 500 == 0:
            log(j + 1, total)
    traj = {k: np.stack(v) for k, v in record.items()}
    traj["log_modulus"] = run.log_modulus.astype(np.float32)
    return traj, run


def samples(traj, ages=range(MAX_AGE)):
    """Network samples from trajectories: for step i and age k, the pose at i (its position
    relative to the one at the start of step i - k), the wrenches measured k + 1 and k + 2
    steps before (zero before the start) and the patch of that time; the target is the wrench
    at i minus the first of them.  Returns X (rows, 96), Y (rows, 6) and the base wrench
    (rows, 6)."""
    pose, wrench, patch = traj["pose"], traj["wrench"], traj["patch"]
    steps, count = wrench.shape[:2]
    padded = np.concatenate([np.zeros((MAX_AGE + 1,) + wrench.shape[1:], np.float32), wrench])
    X, Y, base = [], [], []
    for k in ages:
        i = np.arange(steps)
        prev = padded[MAX_AGE + 1 + i - 1 - k]                   # (steps, count, 6)
        before = padded[MAX_AGE + 1 + i - 2 - k]
        # patch[t] is the patch at the start of step t (after step t - 1).
        old_patch = patch[np.maximum(i - k, 0)]
        rel = pose.copy()
        rel[:, :, :3] -= pose[np.maximum(i - k, 0), :, :3]
        x = np.concatenate([rel, prev, before, old_patch, np.broadcast_to(traj["log_modulus"][None, :, None],
                                                                  (steps, count, 1)),
                            np.full((steps, count, 1), k, np.float32)], axis=2)
        X.append(x.reshape(-1, INPUTS))
        Y.append((wrench - prev).reshape(-1, OUTPUTS))
        base.append(prev.reshape(-1, OUTPUTS))
    return np.concatenate(X), np.concatenate(Y), np.concatenate(base)


def train(X, Y, epochs=20, batch=4096, lr=1.0e-3, lr_end=1.0e-5, weight_decay=1.0e-4, seed=0, device="cuda:0",
          validation=None, log=None):
    """A fresh network trained by the HIP kernels on (X, Y) from `samples` (AdamW, cosine
    learning rate); validation = (Xv, Yv) from other tasks keeps the best epoch.  Returns
    (net, training loss per epoch)."""
    net = Mlp(INPUTS, HIDDEN, OUTPUTS, capacity=batch, device=device, seed=seed)
    Xd = torch.as_tensor(X, device=device)
    Yd = torch.as_tensor(Y, device=device)
    net.standardize(Xd, Yd, x_floor=INPUT_FLOOR, y_floor=OUTPUT_FLOOR)
    if validation is not None:
        validation = tuple(torch.as_tensor(v, device=device) for v in validation)
    history = net.fit(Xd, Yd, epochs=epochs, batch=batch, lr=lr, lr_end=lr_end, weight_decay=weight_decay,
                      seed=seed, validation=validation, log=log)
    return net, history


def evaluate_predictors(net, X, Y):
    """Mean relative force error on held-out samples from `samples`: the network against
    holding the last measured wrench (what a widened exchange has without a model)."""
    held = X[:, PREV]
    Y = Y + held
    pred = held + net.infer(torch.as_tensor(X, device=net.device)).cpu().numpy()
    norm = np.linalg.norm(Y[:, :3], axis=1).sum()
    return dict(mlp=float(np.linalg.norm(pred[:, :3] - Y[:, :3], axis=1).sum() / norm),
                hold=float(np.linalg.norm(held[:, :3] - Y[:, :3], axis=1).sum() / norm))


# ---------------------------------------------------------------------------- plain coupling
def plain_run(config, task, steps=None, arm_kw=None, device="cuda:0", record=None, frame_stride=10):
    """The reference's coupling loop for one knife (`franka_run._exchange_wrench`,
    franka_run.py:394-422, with `Franka.step`): per robot step `arm.state()`,
    `ExternalToolPort.advance` over one interval on a plain `Solver`, then the arm step with the
    interval's wrench.  Returns the summary with sim/wall and the force trace.

    record: a new directory for the run's files in the reference's formats (report.json,
    trajectory.npz, columns.json, energy_columns.json; franka_run.py:600-690).  The arm
    stand-in has no joints: its log and frames carry the end-effector pose (position, xyzw) as
    qpos, its twist as qvel, and NaN in the joint-torque columns."""
    from pathlib import Path
    from .coupling import ExternalToolPort
    from .diagnostics import Metrics
    from .runio import TrajectoryRecorder, edge_height, run_report, write
    from .solver import Solver

    solver = Solver(config, device)
    port = ExternalToolPort(solver)
    origin, orientation, _, _ = task.at(0.0)
    arm = ImpedanceArm(origin[:1], orientation[:1], **(arm_kw or {}))
    total = int(round(task.duration / INTERVAL)) if steps is None else int(steps)
    forces = np.empty((total, 3))
    out = None
    if record is not None:
        out = Path(record)
        out.mkdir(parents=True, exist_ok=False)
        recorder = TrajectoryRecorder(nq=7, nv=6, interval=INTERVAL, frame_stride=frame_stride)
        metrics = Metrics(solver)
    recording_s = 0.0
    begin = time.perf_counter()
    status, error = "completed", None
    try:
        for j in range(total):
            origin, orientation, cutting, cut = task.at(j * INTERVAL)
            reaction = port.advance(arm.state(0), INTERVAL, cutting=cutting, cut_index=cut)
            wrench = np.concatenate([reaction.force_n, reaction.torque_nm])
            arm.step(origin[:1], orientation[:1], wrench[None])
            forces[j] = reaction.force_n
            if out is not None:
                t_rec = time.perf_counter()
                pose = np.concatenate([arm.p[0], arm.q[0]])
                edge = origin[0] - rotation_matrix(orientation)[0, :, 2] * task.half_z
                nan7 = np.full(7, np.nan)
                if recorder.step((j + 1) * INTERVAL, edge, arm.p[0], wrench, pose, np.concatenate([arm.v[0], arm.w[0]]),
                                 nan7, nan7, wrench):
                    lowest, _ = edge_height([arm.p[:1]], [arm.q[:1]], config.blade_half_size)
                    energy = np.concatenate([[(j + 1) * INTERVAL], metrics.read(), [float(solver.work.numpy()[0])]])
                    recorder.frame(pose, solver, energy, np.concatenate([origin[0], orientation[0]]), lowest[0, 0])
                recording_s += time.perf_counter() - t_rec
    except Exception as exc:
        status, error = "failed", repr(exc)
        raise
    finally:
        wall = time.perf_counter() - begin - recording_s
        if out is not None:
            recorder.save(out, solver.origin, solver.radius, ["time_s"] + Metrics.columns +
                          ["directional_contact_work_J"])
            write(out / "report.json", run_report(status, error, recorder.steps * INTERVAL, wall,
                                                  {"material": port.timing_s["submission"] +
                                                   port.timing_s["completion_and_readback"],
                                                   "recording": recording_s}, port.timing_s, True))
    return dict(instances=1, robot_steps=total, sim_s=total * INTERVAL, wall_s=wall,
                sim_per_wall=total * INTERVAL / wall, mean_force_n=float(np.linalg.norm(forces, axis=1).mean())), forces


# ---------------------------------------------------------------------------- command line
def _train_command(args):
    import json

    def log(*a):
        print(*a, flush=True)

    t0 = time.perf_counter()
    traj, run = generate(args.tasks, seed=args.seed, steps=args.steps,
                         log=lambda j, n: log(f"generate: step {j}/{n}, {time.perf_counter() - t0:.1f} s"))
    generation = dict(tasks=args.tasks, robot_steps=args.steps, intervals=args.tasks * args.steps,
                      wall_s=time.perf_counter() - t0, sim_per_wall=run.summary()["sim_per_wall"])
    del run
    X, Y, _ = samples(traj)
    del traj
    vtraj, _ = generate(args.validation_tasks, seed=args.seed + 1, steps=args.steps)
    validation = samples(vtraj)[:2]
    t1 = time.perf_counter()
    net, history = train(X, Y, epochs=args.epochs, batch=args.batch, lr=args.lr, weight_decay=args.weight_decay,
                         seed=args.seed, validation=validation,
                         log=lambda e, loss: log(f"train: epoch {e} loss {loss:.4f}"))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    params = net.params.numel()
    # Forward 2 flops per weight and sample, backward 4 (weight and input gradients).
    flops = 6.0 * params * (len(X) // args.batch) * args.batch * args.epochs
    net.save(args.out)
    report = dict(generation=generation, samples=len(X), epochs=args.epochs, batch=args.batch, lr=args.lr,
                  weight_decay=args.weight_decay, train_s=train_s, train_tflops=flops / train_s / 1e12,
                  loss=history, validation_loss=net.validation_history, network=args.out)
    log(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


def _evaluate_command(args):
    import json

    def log(*a):
        print(*a, flush=True)

    net = Mlp.load(args.net)
    report = {}
    # Held-out tasks: per-age force error of the network and of holding the measured wrench.
    traj, _ = generate(args.test_tasks, seed=args.seed, steps=args.steps)
    X, Y, _ = samples(traj)
    rows = len(X) // MAX_AGE
    report["heldout_by_age"] = [dict(age=k, **evaluate_predictors(net, X[k * rows:(k + 1) * rows],
                                                                  Y[k * rows:(k + 1) * rows])) for k in range(MAX_AGE)]
    log(json.dumps(report["heldout_by_age"]))
    configs = task_configs(1, args.seed)
    task = ThreeCutTask(1, args.seed, half_z=configs[0].blade_half_size[2])
    report["plain_single"], _ = plain_run(configs[0], task, steps=args.steps)
    log("plain", json.dumps(report["plain_single"]))
    for name, window, predictor, count in (("full_batch", 1, None, args.instances),
                                           ("hold_batch", args.window, "hold", args.instances),
                                           ("mlp_batch", args.window, "mlp", args.instances),
                                           ("mlp_single", args.window, "mlp", 1)):
        run = CutRun(count, seed=args.seed, window=window)
        report[name] = run.run(predictor, net if predictor == "mlp" else None, steps=args.steps)
        log(name, json.dumps(report[name]))
    report["speedup_vs_plain"] = report["mlp_batch"]["sim_per_wall"] / report["plain_single"]["sim_per_wall"]
    # Fallback rate against the force error felt by the arm, as the residual threshold moves.
    report["threshold_sweep"] = []
    for threshold in (0.03, 0.05, 0.08):
        for predictor in ("hold", "mlp"):
            run = CutRun(args.instances, seed=args.seed, window=args.window)
            r = run.run(predictor, net if predictor == "mlp" else None, steps=args.steps, threshold=threshold)
            report["threshold_sweep"].append(dict(threshold=threshold, predictor=predictor, **r))
            log("sweep", threshold, predictor, json.dumps(r))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


def main(argv=None):
    import argparse

    p = argparse.ArgumentParser(prog="python -m blademaster_hip.surrogate",
                                description="Train and evaluate the interval-wrench network on the GPU.")
    sub = p.add_subparsers(dest="command", required=True)
    t = sub.add_parser("train", help="generate three-cut trajectories and train the network")
    t.add_argument("--tasks", type=int, default=1024)
    t.add_argument("--validation-tasks", type=int, default=32)
    t.add_argument("--steps", type=int, default=3300)
    t.add_argument("--epochs", type=int, default=16)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=3e-4)
    t.add_argument("--weight-decay", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", required=True)
    t.add_argument("--report")
    e = sub.add_parser("evaluate", help="held-out errors and coupled runs against holding the wrench")
    e.add_argument("--net", required=True)
    e.add_argument("--test-tasks", type=int, default=16)
    e.add_argument("--instances", type=int, default=64)
    e.add_argument("--window", type=int, default=4)
    e.add_argument("--steps", type=int, default=3300)
    e.add_argument("--seed", type=int, default=3)
    e.add_argument("--report")
    args = p.parse_args(argv)
    (_train_command if args.command == "train" else _evaluate_command)(args)


if __name__ == "__main__":
    main()
