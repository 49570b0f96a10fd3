"""Impedance-controlled knife carrier: a stand-in for the Franka arm of `franka_run`.

The reference couples the material to a MuJoCo Franka (`franka.py:341-397`): each robot step
reads the tool state with the mobility `J M^-1 J^T`, advances the material over one coupling
interval, and applies the interval-average wrench while the Cartesian impedance controller
integrates the arm.  MuJoCo and the arm model are not available here, so `ImpedanceArm` keeps
that interface (`state()` -> ToolState with a mobility, `step(target, wrench)`) for a rigid end
effector of mass m and inertia J at the tool origin under a critically damped Cartesian
impedance law:

    F = K (p_target - p) - D v + f_ext,      tau = K_r e_R - D_r w + tau_ext,

with e_R the rotation vector from the current to the target orientation; semi-implicit Euler
at the robot step.  The mobility is blockdiag(I / m, R J^-1 R^T).  Every array is per
instance, so one object drives a batch of knives.
"""

import numpy as np

from .coupling import ToolState


def quat_mul(a, b):
    """Hamilton product of xyzw quaternions (..., 4)."""
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    v = aw * bv + bw * av + np.cross(av, bv)
    w = aw * bw - np.sum(av * bv, axis=-1, keepdims=True)
    return np.concatenate([v, w], axis=-1)


def quat_conj(q):
    return np.concatenate([-q[..., :3], q[..., 3:]], axis=-1)


def rotvec_of(q):
    """Rotation vector (axis times angle) of xyzw quaternions, shortest arc."""
    q = np.where(q[..., 3:] < 0, -q, q)
    s = np.linalg.norm(q[..., :3], axis=-1, keepdims=True)
    angle = 2 * np.arctan2(s, q[..., 3:])
    scale = np.where(s > 1e-12, angle / np.maximum(s, 1e-300), 2.0)
    return q[..., :3] * scale


def quat_of_rotvec(r):
    angle = np.linalg.norm(r, axis=-1, keepdims=True)
    half = 0.5 * angle
    s = np.where(angle > 1e-12, np.sin(half) / np.maximum(angle, 1e-300), 0.5)
    return np.concatenate([r * s, np.cos(half)], axis=-1)


def rotation_matrix(q):
    x, y, z, w = (q[..., i] for i in range(4))
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - 2 * (y * y + z * z)
    R[..., 0, 1] = 2 * (x * y - z * w)
    R[..., 0, 2] = 2 * (x * z + y * w)
    R[..., 1, 0] = 2 * (x * y + z * w)
    R[..., 1, 1] = 1 - 2 * (x * x + z * z)
    R[..., 1, 2] = 2 * (y * z - x * w)
    R[..., 2, 0] = 2 * (x * z - y * w)
    R[..., 2, 1] = 2 * (y * z + x * w)
    R[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return R


class ImpedanceArm:
    """Batch of impedance-controlled knife carriers.

    mass (kg), inertia (kg m^2, isotropic), stiffness (N/m), rot_stiffness (N m/rad); the
    damping is critical for each.  Each may be a scalar or one value per instance."""

    def __init__(self, position, orientation, mass=2.0, inertia=0.02, stiffness=4000.0, rot_stiffness=20.0,
                 dt=0.001):
        self.p = np.array(position, dtype=np.float64).reshape(-1, 3)
        count = len(self.p)
        self.q = np.array(orientation, dtype=np.float64).reshape(-1, 4) * np.ones((count, 1))
        self.q /= np.linalg.norm(self.q, axis=1, keepdims=True)
        self.v = np.zeros((count, 3))
        self.w = np.zeros((count, 3))
        per = lambda value: np.broadcast_to(np.asarray(value, dtype=np.float64), (count,)).copy()  # noqa: E731
        self.mass, self.inertia = per(mass), per(inertia)
        self.k, self.k_rot = per(stiffness), per(rot_stiffness)
        self.d = 2 * np.sqrt(self.k * self.mass)
        self.d_rot = 2 * np.sqrt(self.k_rot * self.inertia)
        self.dt = float(dt)

    def __len__(self):
        return len(self.p)

    # ------------------------------------------------------------------ state
    def mobility(self):
        """World [linear, angular] mobility at the tool origin, (instances, 6, 6)."""
        M = np.zeros((len(self), 6, 6))
        M[:, :3, :3] = np.eye(3) / self.mass[:, None, None]
        # Isotropic inertia: R J^-1 R^T = I / J.
        M[:, 3:, 3:] = np.eye(3) / self.inertia[:, None, None]
        return M

    def records(self):
        """49-float tool records (position, xyzw, velocity, angular velocity, mobility)."""
        return np.concatenate([self.p, self.q, self.v, self.w, self.mobility().reshape(len(self), 36)], axis=1)

    def state(self, i=0):
        """ToolState of instance i, as `Franka.state` returns it (franka.py:341-356)."""
        return ToolState(self.p[i].copy(), self.q[i].copy(), self.v[i].copy(), self.w[i].copy(),
                         self.mobility()[i])

    def save(self):
        return tuple(a.copy() for a in (self.p, self.q, self.v, self.w))

    def load(self, saved):
        self.p, self.q, self.v, self.w = (a.copy() for a in saved)

    # ------------------------------------------------------------------ dynamics
    def step(self, target_position, target_orientation, wrench):
        """One robot step with the external wrench (force 3, torque 3 about the tool origin)
        held over it (franka.py:358-397)."""
        wrench = np.asarray(wrench, dtype=np.float64).reshape(len(self), 6)
        target_position = np.asarray(target_position, dtype=np.float64).reshape(len(self), 3)
        target_orientation = np.asarray(target_orientation, dtype=np.float64).reshape(len(self), 4)
        force = self.k[:, None] * (target_position - self.p) - self.d[:, None] * self.v + wrench[:, :3]
        error = rotvec_of(quat_mul(target_orientation, quat_conj(self.q)))
        torque = self.k_rot[:, None] * error - self.d_rot[:, None] * self.w + wrench[:, 3:]
        self.v = self.v + self.dt * force / self.mass[:, None]
        self.p = self.p + self.dt * self.v
        self.w = self.w + self.dt * torque / self.inertia[:, None]
        self.q = quat_mul(quat_of_rotvec(self.w * self.dt), self.q)
        self.q /= np.linalg.norm(self.q, axis=1, keepdims=True)
