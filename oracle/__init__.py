"""TEST INFRASTRUCTURE ONLY — Python side of the CPU oracle (oracle/tlmpm_oracle.cpp).

Used by tests/, `__graft_entry__.smoke()` and bench.py's cpu_baseline leg as the checker of
the HIP path; the product package never imports it.  `OracleSim` holds a full simulation
state in the oracle library (float32 mirror or float64), with per-run constants derived
here independently of the product from the reference formulas (solver.py:164-200).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

REAL_ARRAYS = {
    # name: per-particle width (None = fixed size)
    "x": 3, "v": 3, "F": 9, "Fp": 9, "stress": 9, "damage": 1, "ref": 3, "weights": 8, "gradients": 24,
    "impulses": 3, "histories": None, "tool_pos": None, "tool_vel": None, "tool_rot": None, "tool_omega": None,
    "tool_mobility": None, "reaction": None, "torque": None, "board": None, "work": None, "speed": None,
    "tool_table_impulse": None, "interval_impulse": None, "interval_angular_impulse": None,
    "bc_side": 1, "bc_impulse": 1, "bc_normal": 3, "viscous": 9,
    # knives 1 .. K-1: one row of n entries each (sized in OracleSim._size)
    "extra_histories": None, "extra_damage": None,
    # grid arrays: one entry per (particle, corner) field slot at most
    "grid_mass": 8, "old_v": 24, "new_v": 24, "grid_impulse": 24, "momentum": 24, "coupled_remapped": 24,
}
FIXED = {
    "histories": lambda n: 6 * n, "tool_pos": 3, "tool_vel": 3, "tool_rot": 4, "tool_omega": 3, "tool_mobility": 36,
    "reaction": 3, "torque": 3, "board": 3, "work": 1, "speed": 1, "tool_table_impulse": 3, "interval_impulse": 3,
    "interval_angular_impulse": 3, "bc_max_violation": 1,
}


# Arrays with one entry per knife (tool_count of them).
PER_KNIFE = {"tool_pos": 3, "tool_vel": 3, "tool_rot": 4, "tool_omega": 3, "reaction": 3, "torque": 3,
             "tool_table_impulse": 3}


class OracleConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("node_count", ctypes.c_int64), ("key_capacity", ctypes.c_int64),
        ("pair_capacity", ctypes.c_int64), ("max_cut_index", ctypes.c_int64),
    ] + [(k, ctypes.c_double) for k in (
        "dt spacing mass volume radius blade_offset mu lam yield_stress flip sn gk mus mud gc board_sn threshold"
    ).split()] + [
        ("half", ctypes.c_double * 3), ("bevel", ctypes.c_double), ("blade_mass", ctypes.c_double),
        ("inertia_inv", ctypes.c_double * 3), ("board_half", ctypes.c_double * 2),
        ("drive_k", ctypes.c_double), ("drive_c", ctypes.c_double), ("drive_limit", ctypes.c_double),
        ("rot_k", ctypes.c_double), ("rot_c", ctypes.c_double),
        ("free_blade", ctypes.c_int32), ("pose_servo", ctypes.c_int32), ("external_tool", ctypes.c_int32),
        ("predict", ctypes.c_int32), ("accumulate_tool_impulses", ctypes.c_int32), ("cache_topology", ctypes.c_int32),
        ("cut_work_advance", ctypes.c_int32), ("cut_hold", ctypes.c_int32),
        ("viscosity", ctypes.c_double), ("mu_maxwell", ctypes.c_double), ("maxwell_relax", ctypes.c_double),
        ("maxwell", ctypes.c_int32), ("tool_count", ctypes.c_int32),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.c_int, ctypes.POINTER(OracleConfig)]
        lib.oracle_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int]
        lib.oracle_substep.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
        lib.oracle_stage.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int64]
        lib.oracle_blade_query.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                           ctypes.c_void_p]
        lib.oracle_svd3.argtypes = [ctypes.c_void_p] * 4
        lib.oracle_compatible.argtypes = [ctypes.c_int64, ctypes.c_int64]
        lib.oracle_friction.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.oracle_resist_cut.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [
            ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
        lib.oracle_couple.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int64, ctypes.c_int64]
        lib.oracle_g2p.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_double, ctypes.c_double, ctypes.c_int64]
        lib.oracle_set_threads.restype = ctypes.c_int
        lib.oracle_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib = lib
    return _lib


def _pow2(v):
    return 1 << (max(int(v), 16) - 1).bit_length()


def constants(cfg):
    """Per-run constants restated from solver.py:181-200 (independent of the product)."""
    h = cfg.grid_spacing
    volume = (h / 2) ** 3
    radius = h / 4
    hx, hy, hz = cfg.blade_half_size
    nu, E = cfg.poisson_ratio, cfg.young_modulus
    ratio = 1.0 if cfg.contact_reference_dt == 0 else (cfg.dt / cfg.contact_reference_dt) ** 2
    sn = cfg.normal_scale * ratio
    return dict(
        volume=volume,
        mass=cfg.density * volume,
        radius=radius,
        blade_offset=max(cfg.blade_contact_radius or radius, radius - hx),
        mu=E / (2 * (1 + nu)),
        lam=E * nu / ((1 + nu) * (1 - 2 * nu)),
        threshold=cfg.cutting_energy * volume ** (2 / 3) * (1 - cfg.sharpness) / cfg.sharpness,
        sn=sn,
        gk=cfg.tangent_stiffness * ratio,
        board_sn=cfg.board_normal_scale or sn,
        inertia_inv=3 / (cfg.blade_mass * np.array([hy * hy + hz * hz, hx * hx + hz * hz, hx * hx + hy * hy])),
        # Maxwell branch (solver.py:189-191).
        maxwell=cfg.maxwell_tau > 0 and cfg.maxwell_fraction > 0,
        mu_maxwell=cfg.maxwell_fraction * E / (2 * (1 + nu)),
        maxwell_relax=float(-np.expm1(-cfg.dt / cfg.maxwell_tau)) if cfg.maxwell_tau > 0 else 0.0,
    )


class OracleSim:
    """One oracle simulation of `cfg` at `precision` 32 or 64 bits.

    reference/nodes/weights/gradients are the stencil arrays of the run (reference order).
    """

    def __init__(self, cfg, reference, nodes, weights, gradients, node_count, precision=32, max_cut_index=0):
        self.cfg = cfg
        self.lib = load()
        self.n = n = len(reference)
        k = constants(cfg)
        self.k = k
        c = OracleConfig()
        c.n, c.node_count = n, node_count
        c.key_capacity, c.pair_capacity = _pow2(16 * n), _pow2(64 * n)
        c.max_cut_index = max_cut_index
        c.dt, c.spacing, c.mass, c.volume = cfg.dt, cfg.grid_spacing, k["mass"], k["volume"]
        c.radius, c.blade_offset, c.mu, c.lam = k["radius"], k["blade_offset"], k["mu"], k["lam"]
        c.yield_stress, c.flip = cfg.yield_stress, cfg.flip_fraction
        c.sn, c.gk, c.mus, c.mud, c.gc = k["sn"], k["gk"], cfg.friction_static, cfg.friction_dynamic, cfg.tangent_damping
        c.board_sn, c.threshold = k["board_sn"], k["threshold"]
        for i in range(3):
            c.half[i] = cfg.blade_half_size[i]
            c.inertia_inv[i] = k["inertia_inv"][i]
        c.bevel, c.blade_mass = cfg.blade_bevel_height, cfg.blade_mass
        c.board_half[0], c.board_half[1] = cfg.board_half_extent
        c.drive_k, c.drive_c, c.drive_limit = cfg.drive_stiffness, cfg.drive_damping, cfg.drive_force_limit
        c.rot_k, c.rot_c = cfg.rotational_stiffness, cfg.rotational_damping
        c.free_blade, c.pose_servo, c.external_tool = int(cfg.free_blade), int(cfg.pose_servo), int(cfg.external_tool)
        c.predict = int(cfg.tool_substep == "predict")
        c.accumulate_tool_impulses = int(cfg.accumulate_tool_impulses)
        c.cache_topology = int(cfg.cache_topology)
        c.cut_work_advance = int(cfg.cut_work == "advance")
        c.cut_hold = int(cfg.cut_hold)
        c.viscosity = cfg.dynamic_viscosity
        c.maxwell = int(k["maxwell"])
        c.mu_maxwell, c.maxwell_relax = k["mu_maxwell"], k["maxwell_relax"]
        c.tool_count = self.knives = int(cfg.tool_count)
        self.c = c
        self.precision = precision
        self.handle = self.lib.oracle_create(precision, ctypes.byref(c))
        self.set("ref", reference)
        self.set("x", reference)
        self.set("nodes", np.asarray(nodes, dtype=np.int64))
        self.set("weights", weights)
        self.set("gradients", gradients)
        hz = cfg.blade_half_size[2]
        self.set("tool_pos", [[0.0, 0.0, cfg.size[2] + 0.012 + hz]] * self.knives)

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.oracle_destroy(self.handle)
            self.handle = None

    def _size(self, name):
        n = self.n
        if name in ("codes",):
            return n
        if name in ("nodes", "slots"):
            return 8 * n
        if name == "errors":
            return 3
        if name == "scalars":
            return 5
        if name in PER_KNIFE:
            return PER_KNIFE[name] * self.knives
        if name == "extra_histories":
            return 3 * n * (self.knives - 1)
        if name == "extra_damage":
            return n * (self.knives - 1)
        if name in FIXED:
            f = FIXED[name]
            return f(n) if callable(f) else f
        if name in REAL_ARRAYS:
            return REAL_ARRAYS[name] * n
        raise KeyError(name)

    def _is_int(self, name):
        return name in ("codes", "nodes", "errors", "slots", "scalars", "keys", "grid_keys", "heads", "links")

    def set(self, name, value):
        count = self._size(name)
        dtype = np.int64 if self._is_int(name) else np.float64
        buf = np.ascontiguousarray(np.asarray(value, dtype=dtype).ravel())
        if buf.size != count:
            raise ValueError(f"{name}: {buf.size} values, expected {count}")
        code = self.lib.oracle_array(self.handle, name.encode(), buf.ctypes.data, count, 1)
        if code:
            raise RuntimeError(f"oracle_array({name}) failed: {code}")

    def get(self, name, count=None):
        count = self._size(name) if count is None else count
        dtype = np.int64 if self._is_int(name) else np.float64
        buf = np.zeros(count, dtype=dtype)
        code = self.lib.oracle_array(self.handle, name.encode(), buf.ctypes.data, count, 0)
        if code:
            raise RuntimeError(f"oracle_array({name}) failed: {code}")
        return buf

    def set_threads(self, threads):
        """Run the per-particle and per-field loops on `threads` OpenMP threads (the results
        equal the one-thread run bit for bit); returns the count in effect."""
        return int(self.lib.oracle_set_threads(self.handle, int(threads)))

    def scalars(self):
        a = self.get("scalars")
        return dict(active_count=int(a[0]), pair_count=int(a[1]), step_count=int(a[2]),
                    topology_dirty=bool(a[3]), single_field=bool(a[4]))

    def set_single_field(self, enabled):
        a = self.get("scalars")
        a[3] = 1
        a[4] = int(bool(enabled))
        self.set("scalars", a)

    def substep(self, target=(0.0, 0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0), cutting=True, cut_index=0):
        t = np.ascontiguousarray(target, dtype=np.float64)
        r = np.ascontiguousarray(q, dtype=np.float64)
        self.lib.oracle_substep(self.handle, t.ctypes.data, r.ctypes.data, int(bool(cutting)), 3**cut_index)

    def stage(self, name, arg0=0, arg1=0):
        code = self.lib.oracle_stage(self.handle, name.encode(), int(arg0), int(arg1))
        if code:
            raise RuntimeError(f"oracle stage {name} failed: {code}")

    def resist_cut(self, p, j, origin, tool_velocity, q, omega, rad, cutting):
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (j, origin, tool_velocity, q, omega)]
        out = np.zeros(3)
        self.lib.oracle_resist_cut(self.handle, int(p), *[a.ctypes.data for a in args], float(rad), int(cutting),
                                   out.ctypes.data)
        return out

    def grid_keys(self):
        active = self.scalars()["active_count"]
        return self.get("grid_keys", 8 * self.n)[:active]


def blade_query(local, half, bevel, precision=64):
    lib = load()
    local = np.ascontiguousarray(local, dtype=np.float64)
    half = np.ascontiguousarray(half, dtype=np.float64)
    out = np.zeros(7)
    lib.oracle_blade_query(precision, local.ctypes.data, half.ctypes.data, float(bevel), out.ctypes.data)
    return out[:3], out[3:6], out[6]


def svd3(A):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib.oracle_svd3(A.ctypes.data, U.ctypes.data, s.ctypes.data, V.ctypes.data)
    return U.reshape(3, 3), s, V.reshape(3, 3)


def compatible(a, b):
    return bool(load().oracle_compatible(int(a), int(b)))


def friction(normal, relative, jn, history, me, dt, mus, mud, gk, gc, precision=64):
    lib = load()
    inp = np.zeros(13)
    inp[0:3], inp[3:6], inp[6], inp[7:10], inp[10] = normal, relative, jn, history, me
    prm = np.array([dt, mus, mud, gk, gc], dtype=np.float64)
    out = np.zeros(6)
    lib.oracle_friction(precision, inp.ctypes.data, prm.ctypes.data, out.ctypes.data)
    return out[:3], out[3:]


def couple(keys, heads, links, mass, state, node_count):
    lib = load()
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    heads = np.ascontiguousarray(heads, dtype=np.int32)
    links = np.ascontiguousarray(links, dtype=np.int32)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    state = np.ascontiguousarray(state, dtype=np.float64).reshape(-1, 3)
    result = np.zeros_like(state)
    lib.oracle_couple(keys.ctypes.data, heads.ctypes.data, links.ctypes.data, mass.ctypes.data, state.ctypes.data,
                      result.ctypes.data, len(keys), int(node_count))
    return result


def g2p(x, v, slots, weights, old_v, new_v, dt, flip):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    slots = np.ascontiguousarray(slots, dtype=np.int32)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    old_v = np.ascontiguousarray(old_v, dtype=np.float64)
    new_v = np.ascontiguousarray(new_v, dtype=np.float64)
    lib.oracle_g2p(x.ctypes.data, v.ctypes.data, slots.ctypes.data, weights.ctypes.data, old_v.ctypes.data,
                   new_v.ctypes.data, float(dt), float(flip), len(x))
    return x, v


class MlsConfig(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int64 * 3), ("plastic", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("origin", ctypes.c_double * 3)] + [
        (k, ctypes.c_double) for k in "dx dt mass volume mu lam damaged_shear softening compressed stretched".split()]


class OracleMls:
    """MLS-MPM baseline oracle (mlsmpm.py:293-393) for config `cfg`, TopoCut parameters
    `params`, on the grid (origin, dims) of the solver under test."""

    def __init__(self, cfg, params, reference, origin, dims, precision=32, damage_from_start=True):
        lib = load()
        lib.oracle_mls_create.restype = ctypes.c_void_p
        lib.oracle_mls_create.argtypes = [ctypes.c_int, ctypes.POINTER(OracleConfig), ctypes.POINTER(MlsConfig)]
        lib.oracle_mls_destroy.argtypes = [ctypes.c_void_p]
        lib.oracle_mls_substep.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_mls_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_int]
        self.lib = lib
        self.n = n = len(reference)
        k = constants(cfg)
        c = OracleConfig()
        c.n, c.node_count, c.key_capacity, c.pair_capacity = n, 1, 16, 16
        c.dt, c.spacing, c.mass, c.volume = cfg.dt, cfg.grid_spacing, k["mass"], k["volume"]
        c.radius, c.blade_offset, c.mu, c.lam = k["radius"], k["blade_offset"], k["mu"], k["lam"]
        c.yield_stress = cfg.yield_stress
        c.sn, c.gk, c.mus, c.mud, c.gc = k["sn"], k["gk"], cfg.friction_static, cfg.friction_dynamic, cfg.tangent_damping
        c.board_sn, c.threshold = k["board_sn"], 1.0
        for i in range(3):
            c.half[i] = cfg.blade_half_size[i]
        c.bevel, c.blade_mass = cfg.blade_bevel_height, cfg.blade_mass
        c.board_half[0], c.board_half[1] = cfg.board_half_extent
        c.external_tool = 1
        m = MlsConfig()
        for i in range(3):
            m.dims[i] = int(dims[i])
            m.origin[i] = float(origin[i])
        m.plastic = int(cfg.yield_stress > 0)
        m.dx, m.dt, m.mass, m.volume = cfg.grid_spacing, cfg.dt, k["mass"], k["volume"]
        m.mu, m.lam, m.damaged_shear, m.softening = k["mu"], k["lam"], params.damaged_shear, params.softening
        m.compressed = (1 - params.compression) ** params.sensitivity
        m.stretched = (1 + params.stretch) ** params.sensitivity
        self.c, self.m = c, m
        self.handle = lib.oracle_mls_create(precision, ctypes.byref(c), ctypes.byref(m))
        self.set("x", reference)
        self.set("damage_reference", np.full(n, 1.0 if damage_from_start else 0.0))

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.oracle_mls_destroy(self.handle)
            self.handle = None

    def _count(self, name):
        n = self.n
        return {"x": 3 * n, "v": 3 * n, "C": 9 * n, "F": 9 * n, "damage": n, "damage_reference": n,
                "yield_stress": n, "impulse": 3, "angular_impulse": 3, "tool_pos": 3, "tool_vel": 3, "tool_rot": 4,
                "tool_omega": 3, "inversions": 1, "histories": 6 * n}[name]

    def set(self, name, value):
        buf = np.ascontiguousarray(np.asarray(value, dtype=np.float64).ravel())
        code = self.lib.oracle_mls_array(self.handle, name.encode(), buf.ctypes.data, self._count(name), 1)
        if code:
            raise RuntimeError(f"oracle_mls_array({name}) failed: {code}")

    def get(self, name):
        buf = np.zeros(self._count(name))
        code = self.lib.oracle_mls_array(self.handle, name.encode(), buf.ctypes.data, self._count(name), 0)
        if code:
            raise RuntimeError(f"oracle_mls_array({name}) failed: {code}")
        return buf

    def substep(self, steps=1):
        self.lib.oracle_mls_substep(self.handle, int(steps))
