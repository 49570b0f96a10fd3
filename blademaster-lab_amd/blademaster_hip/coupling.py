"""External rigid-tool interface: the tool state is set at the start of each coupling
interval, contact impulses are accumulated over its substeps and returned as an
interval-average wrench (`blademaster_warp/coupling.py:1-189`, paper III-G).

The packed formats are the reference's: 49 floats in ([position 3, quaternion xyzw 4,
velocity 3, angular velocity 3, mobility 36], tool.py:19-39) and 10 + 7 S floats out
([impulse 3, angular impulse about the world origin 3, error flags 3, poses 7 S, residual],
tool.py:42-68, blade_constraint.py:208-210).  One interval is one HIP graph replay.
"""

from dataclasses import dataclass
import ctypes
import itertools
import time
import weakref
import numpy as np
import torch

# Ports are told apart in the solver's graph cache by a serial number that is never reused
# (an id() can be reused by a later port once this one is collected).
_port_serials = itertools.count()


def _release_port_graphs(solver_ref, owner):
    solver = solver_ref()
    if solver is not None:
        solver.release_graphs(owner)


def register_port(port, kind):
    """Give `port` its graph-cache tag prefix and drop its graphs from the solver's cache
    when the port is collected."""
    port._owner = (kind, next(_port_serials))
    weakref.finalize(port, _release_port_graphs, weakref.ref(port.solver), port._owner)
    return port._owner


def vector(value, length, name):
    result = np.array(value, dtype=np.float64, copy=True)
    if result.shape != (length,) or not np.isfinite(result).all():
        raise ValueError(f"{name} must contain {length} finite values")
    return result


@dataclass(frozen=True)
class ToolState:
    # The knife frame's origin, world metres: the centre of the blade box the contact kernel
    # tests against, not a point elsewhere on the tool.
    position: object
    orientation_xyzw: object
    # Velocity of that frame origin, and the angular velocity, world frame.
    linear_velocity: object
    angular_velocity: object
    # 6 x 6 world-frame mobility at the frame origin, rows and columns (linear, angular);
    # None for a prescribed (infinitely heavy) knife.
    mobility: object = None

    def checked(self):
        """Validated copy (coupling.py:31-50)."""
        position = vector(self.position, 3, "position")
        q = vector(self.orientation_xyzw, 4, "orientation_xyzw")
        if abs(np.linalg.norm(q) - 1.0) > 1.0e-5:
            raise ValueError("Tool orientation must be a unit xyzw quaternion")
        mobility = np.zeros((6, 6)) if self.mobility is None else np.asarray(self.mobility, dtype=float)
        if (
            mobility.shape != (6, 6)
            or not np.isfinite(mobility).all()
            or np.abs(mobility - mobility.T).max() > 1.0e-10
            or np.linalg.eigvalsh(mobility).min() < -1.0e-10
        ):
            raise ValueError("Mobility must be a finite symmetric positive-semidefinite 6x6 matrix")
        return ToolState(
            position,
            q / np.linalg.norm(q),
            vector(self.linear_velocity, 3, "linear_velocity"),
            vector(self.angular_velocity, 3, "angular_velocity"),
            mobility.copy(),
        )


@dataclass(frozen=True)
class Reaction:
    duration_s: float
    impulse_ns: np.ndarray
    angular_impulse_nms: np.ndarray
    origin_world_m: np.ndarray

    @property
    def force_n(self):
        return self.impulse_ns / self.duration_s

    @property
    def torque_nm(self):
        return self.angular_impulse_nms / self.duration_s

    def about(self, origin_world_m):
        """Shift the moment reference, preserving the same physical wrench (coupling.py:68-72)."""
        origin = vector(origin_world_m, 3, "moment origin")
        angular = self.angular_impulse_nms + np.cross(self.origin_world_m - origin, self.impulse_ns)
        return Reaction(self.duration_s, self.impulse_ns.copy(), angular, origin)


class ExternalToolPort:
    def __init__(self, solver, record_trajectory=False):
        if not solver.config.external_tool:
            raise ValueError("ExternalToolPort requires Config(external_tool=True)")
        self.solver = solver
        dev = solver.device.torch
        self.packed = torch.zeros(49, dtype=torch.float32, device=dev)
        # Pinned staging ring for the per-interval tool state: the upload is asynchronous, so
        # the host can enqueue the next interval while the device runs this one; a slot is
        # reused only after its copy has executed.
        self._staging = [torch.zeros(49, dtype=torch.float32, pin_memory=True) for _ in range(4)]
        self._staged = [None] * len(self._staging)
        self._stage_next = 0
        self.result = torch.zeros(10, dtype=torch.float32, device=dev)
        self.record_trajectory = record_trajectory
        self.trajectory = None
        self.trajectory_times = None
        self._results = {}
        self.timing_s = dict(submission=0.0, completion_and_readback=0.0)
        register_port(self, "external")

    def _ptr(self, tensor):
        return ctypes.c_void_p(tensor.data_ptr())

    def advance(self, state, duration_s, *, cutting=True, cut_index=0, graph=True):
        """Run one coupling interval and return its reaction (coupling.py:88-189)."""
        submission_begin = time.perf_counter()
        state, steps, samples = self.submit(state, duration_s, cutting=cutting, cut_index=cut_index, graph=graph)
        return self.collect(state, duration_s, samples, submission_begin)

    def submit(self, state, duration_s, *, cutting=True, cut_index=0, graph=True):
        """Enqueue one interval on the current stream without waiting for it.

        Returns (checked state, substeps, trajectory samples); `collect` reads the result.
        Consecutive submissions are ordered on the stream, so a caller may enqueue many
        intervals and read back only the last one.
        """
        state = state.checked()
        s = self.solver
        if not np.isfinite(duration_s) or duration_s <= 0:
            raise ValueError("Coupling duration must be positive and finite")
        steps = int(round(duration_s / s.config.dt))
        if steps < 1 or abs(steps * s.config.dt - duration_s) > 1.0e-12 + 1.0e-10 * abs(duration_s):
            raise ValueError("Coupling duration must be an integer number of material substeps")
        cut_index = s._set_cut(cut_index)
        held_pose = s.config.tool_substep == "hold"
        samples = 2 if held_pose else steps + 1
        dev = s.device.torch
        if self.record_trajectory:
            if steps not in self._results:
                self._results[steps] = torch.zeros(10 + 7 * samples, dtype=torch.float32, device=dev)
            self.result = self._results[steps]
        host = np.concatenate(
            [state.position, state.orientation_xyzw, state.linear_velocity, state.angular_velocity,
             state.mobility.ravel()]
        ).astype(np.float32)
        k = self._stage_next
        self._stage_next = (k + 1) % len(self._staging)
        if self._staged[k] is not None:
            self._staged[k].synchronize()
        self._staging[k].copy_(torch.from_numpy(host))
        self.packed.copy_(self._staging[k], non_blocking=True)
        done = torch.cuda.Event()
        done.record(torch.cuda.current_stream(dev))
        self._staged[k] = done
        result = self.result
        packed = self.packed

        def load():
            s.launch("bm_load_external_state", self._ptr(packed))

        def record(index):
            s.launch("bm_record_pose", int(index), self._ptr(result))

        def finish():
            s.launch("bm_validate_interval")
            s.launch("bm_store_interval", self._ptr(result))
            if s.blade_constraint is not None:
                s.blade_constraint.store_residual(result)

        before = [load]
        sample = None
        if self.record_trajectory:
            def sample(i):
                # A held knife does not move inside the interval: its start and end poses only.
                if not held_pose or i == steps - 1:
                    record(1 if held_pose else i + 1)

            before.append(lambda: record(0))
        if graph:
            unused_targets = np.zeros((steps, 3), dtype=np.float32)
            _, _, captured = s.prepare(
                unused_targets, cutting, cut_index, before=before, after=[finish], tag=self._owner,
                after_step=sample,
            )
            captured.replay()
            s.step_count += steps
        else:
            for fn in before:
                fn()
            for i in range(steps):
                s._step_kernels((0.0, 0.0, 0.0), cutting, cut_index)
                if sample is not None:
                    sample(i)
            finish()
        return state, steps, samples

    def collect(self, state, duration_s, samples, submission_begin=None):
        """Wait for the last submitted interval and return its Reaction."""
        readback_begin = time.perf_counter()
        if submission_begin is not None:
            self.timing_s["submission"] += readback_begin - submission_begin
        out = self.result.cpu().numpy().astype(np.float64)
        self.timing_s["completion_and_readback"] += time.perf_counter() - readback_begin
        impulse, angular, errors = out[:3], out[3:6], out[6:9]
        if self.record_trajectory:
            self.trajectory = out[9 : 9 + 7 * samples].reshape(samples, 7)
            self.trajectory_times = np.linspace(0.0, duration_s, samples)
        (px, py, pz), (ix, iy, iz) = state.position, impulse
        angular = angular - np.array([py * iz - pz * iy, pz * ix - px * iz, px * iy - py * ix])
        if errors.any() or not np.isfinite([impulse, angular]).all():
            raise RuntimeError(f"Invalid material coupling interval: GPU errors {errors.astype(int).tolist()}")
        if not np.isfinite(out[-1]) or out[-1] > 1.0e-6:
            raise RuntimeError(f"Blade grid contact did not converge: inward residual {out[-1]:.6g} m/s")
        return Reaction(float(duration_s), impulse, angular, state.position.copy())
