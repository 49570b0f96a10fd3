"""ctypes binding of libbm_hip.so (declared in include/bm_hip.h).

The library must exist next to this file (built by `__graft_entry__.build()` or
`make -C blademaster-lab_amd/csrc`); there is no fallback implementation, so a missing or
stale library raises immediately.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libbm_hip.so")

c_f32p = ctypes.POINTER(ctypes.c_float)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_u8p = ctypes.POINTER(ctypes.c_uint8)


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("node_count", ctypes.c_int64),
        ("fields", ctypes.c_int64),
        ("dims", ctypes.c_int64 * 3),
        ("pair_capacity", ctypes.c_int64),
        ("list_buckets", ctypes.c_int64),
        ("list_lanes", ctypes.c_int64),
        ("tool_model", ctypes.c_int64),
        ("free_blade", ctypes.c_int64),
        ("lock_rotation", ctypes.c_int64),
        ("predict", ctypes.c_int64),
        ("plastic", ctypes.c_int64),
        ("brick", ctypes.c_int64 * 3),
        ("bricks", ctypes.c_int64 * 3),
        ("rigid_blocks", ctypes.c_int64),
        ("blade_hold", ctypes.c_int64),
        ("changed_capacity", ctypes.c_int64),
        ("instances", ctypes.c_int64),
        ("active_bricks", ctypes.c_int64),
        ("active_nodes", ctypes.c_int64),
        ("pair_mode", ctypes.c_int64),
        ("regular", ctypes.c_int64),
        ("maxwell", ctypes.c_int64),
        ("tools", ctypes.c_int64),
    ] + [
        (name, ctypes.c_double)
        for name in (
            "dt spacing mass volume radius blade_offset mu lam yield_stress flip "
            "sn gk mus mud gc board_sn threshold"
        ).split()
    ] + [
        ("half", ctypes.c_double * 3),
        ("bevel", ctypes.c_double),
        ("blade_mass", ctypes.c_double),
        ("inertia_inv", ctypes.c_double * 3),
        ("board_half", ctypes.c_double * 2),
        ("drive_k", ctypes.c_double),
        ("drive_c", ctypes.c_double),
        ("drive_limit", ctypes.c_double),
        ("rot_k", ctypes.c_double),
        ("rot_c", ctypes.c_double),
        ("pair_skin", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("mu_maxwell", ctypes.c_double),
        ("maxwell_relax", ctypes.c_double),
        ("mover_capacity", ctypes.c_int64),
    ]


STATE_FIELDS = [
    ("x", c_f32p), ("v", c_f32p), ("F", c_f32p), ("Fp", c_f32p), ("stress", c_f32p), ("damage", c_f32p),
    ("ref", c_f32p), ("local", c_f32p), ("impulses", c_f32p), ("histories", c_f32p),
    ("codes", c_i64p), ("transfer_codes", c_i64p), ("base_node", c_i32p), ("field_idx", c_u8p),
    ("field_code", c_i64p), ("node_nf", c_i32p),
    ("grid_mass", c_f32p), ("momentum", c_f32p), ("grid_impulse", c_f32p), ("old_v", c_f32p), ("new_v", c_f32p),
    ("remapped", c_f32p), ("grid_vel", c_f32p),
    ("tool_pos", c_f32p), ("tool_vel", c_f32p), ("tool_rot", c_f32p), ("tool_omega", c_f32p), ("mobility", c_f32p),
    ("reaction", c_f32p), ("torque", c_f32p), ("board", c_f32p), ("work", c_f32p), ("speed", c_f32p),
    ("table", c_f32p), ("interval_impulse", c_f32p), ("interval_angular", c_f32p),
    ("errors", c_i32p), ("topology_dirty", c_i32p), ("fp_identity", c_i32p), ("update_pending", c_i32p), ("update_done", c_i32p), ("instance", c_i32p),
    ("inst_mu", c_f32p), ("inst_lam", c_f32p), ("inst_sums", c_f32p),
    ("inst_buckets", c_i32p), ("inst_pair_builds", c_i32p), ("inst_pair_dirty", c_i32p),
    ("pair_keys", c_i64p), ("pair_history", c_f32p), ("pair_used", c_i32p), ("pair_count", c_i32p),
    ("pair_dirty", c_i32p), ("pair_builds", c_i32p), ("list_reach", c_f32p), ("built_x", c_f32p),
    ("built_codes", c_i64p), ("partners", c_i32p), ("partner_count", c_i32p), ("list_heads", c_i64p),
    ("list_links", c_i32p), ("changed", c_i32p), ("changed_count", c_i32p), ("pair_mark", c_i32p),
    ("pair_epoch", c_i32p), ("perm", c_i32p), ("brick_list", c_i32p), ("node_list", c_i32p), ("p2g_partial", c_f32p), ("remap_partial", c_f32p),
    ("rigid_partial", c_f32p), ("counters", ctypes.POINTER(ctypes.c_uint32)),
    ("bc_side", c_f32p), ("bc_normal", c_f32p), ("bc_lever", c_f32p), ("bc_rhs", c_f32p), ("bc_bound", c_f32p),
    ("bc_impulse", c_f32p), ("bc_extrapolated", c_f32p), ("bc_active", c_i32p), ("bc_nodes", c_i32p),
    ("bc_node_mark", c_i32p), ("bc_grid", c_f32p), ("bc_coupled", c_f32p), ("bc_counts", c_i32p),
    ("bc_sums", c_f32p), ("history_mask", ctypes.POINTER(ctypes.c_uint64)),
    ("viscous", ctypes.POINTER(ctypes.c_double)), ("viscous_impulse", c_f32p),
    ("extra_histories", c_f32p), ("extra_damage", c_f32p), ("extra_partial", c_f32p),
    ("multi_nodes", c_i32p), ("multi_stamp", c_i32p), ("multi_count", c_i32p), ("field_idx_prev", c_u8p),
    ("movers", c_i32p), ("displaced", c_i32p), ("refresh_mark", c_i32p), ("refresh_state", c_i32p),
]


class State(ctypes.Structure):
    _fields_ = STATE_FIELDS


class MlsParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("dims", ctypes.c_int64 * 3),
        ("plastic", ctypes.c_int64),
        ("origin", ctypes.c_double * 3),
    ] + [(name, ctypes.c_double) for name in
         "dx dt mass volume mu lam damaged_shear softening compressed stretched".split()] + [
        ("groups", ctypes.c_int64)]


MLS_STATE_FIELDS = [
    (name, c_f32p) for name in
    "x v C F damage damage_reference yield_stress impulse grid_m grid_p grid_v interval_impulse "
    "interval_angular".split()
] + [("inversions", c_i32p), ("perm", c_i32p)]


class MlsState(ctypes.Structure):
    _fields_ = MLS_STATE_FIELDS


LABELS_FIELDS = [
    ("fields", ctypes.c_int64), ("corner_mask", c_u8p), ("field_code", c_i64p), ("node_nf", c_i32p), ("field_idx", c_u8p),
    ("parent", c_i32p), ("root", c_i32p), ("block_sum", c_i32p), ("labels", c_i32p), ("count", c_i32p),
    ("errors", c_i32p),
]


class Labels(ctypes.Structure):
    _fields_ = LABELS_FIELDS


MLP_FIELDS = [
    ("inputs", ctypes.c_int32),
    ("hidden", ctypes.c_int32),
    ("outputs", ctypes.c_int32),
] + [(name, c_f32p) for name in ("params", "grads", "adam_m", "adam_v", "h1", "h2", "out", "d_out", "d_h2", "d_h1",
                                 "loss")] + [
    ("step", c_i32p),
    ("lr", c_f32p),
    ("packed", ctypes.POINTER(ctypes.c_uint16)),
    ("work", c_f32p),
    ("work_len", ctypes.c_int64),
] + [(name, c_f32p) for name in ("x_shift", "x_scale", "y_shift", "y_scale")]


class Mlp(ctypes.Structure):
    _fields_ = MLP_FIELDS


MP = ctypes.POINTER(MlsParams)
MS = ctypes.POINTER(MlsState)
P = ctypes.POINTER(Params)
S = ctypes.POINTER(State)
VP = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
F32 = ctypes.c_float

# name -> argument types (all return int)
SIGNATURES = {
    "bm_version": [],
    "bm_refresh_stress": [P, S, VP],
    "bm_clear_step": [P, S, VP],
    "bm_rigid_contacts": [P, S, I32, VP],
    "bm_pair_contacts": [P, S, I32, VP],
    "bm_register": [P, S, VP],
    "bm_p2g": [P, S, VP],
    "bm_couple_and_velocity": [P, S, VP],
    "bm_g2p_remap": [P, S, VP],
    "bm_couple_remapped": [P, S, VP],
    "bm_update_F": [P, S, VP],
    "bm_material_step": [P, S, VP],
    "bm_tool_step": [P, S, VP, VP, I32, VP],
    "bm_finish_external": [P, S, VP],
    "bm_accumulate_reaction": [P, S, VP],
    "bm_assign": [P, S, I64, I32, VP],
    "bm_mark_relabelled": [P, S, VP],
    "bm_substep": [P, S, I32, I32, I64, I32, I32, VP, VP, I32, VP],
    "bm_substep_deferred": [P, S, I32, I32, I64, I32, I32, VP, VP, I32, VP],
    "bm_load_external_state": [P, S, VP, VP],
    "bm_record_pose": [P, S, I32, VP, VP],
    "bm_store_interval": [P, S, VP, VP],
    "bm_validate_interval": [P, S, VP],
    "bm_count_embedded": [P, S, VP, VP],
    "bm_metrics": [P, S, VP, VP],
    "bm_plastic_project": [P, S, VP],
    "bm_probe_blade_query": [VP, ctypes.POINTER(ctypes.c_float), F32, VP, I32, VP],
    "bm_probe_friction": [VP, F32, F32, F32, F32, F32, VP, I32, VP],
    "bm_probe_svd3": [VP, VP, VP, VP, I32, VP],
    "bm_probe_resist_cut": [VP, VP, VP, VP, VP, I32, VP, VP, I32, VP],
    "bm_couple": [VP, VP, VP, VP, VP, VP, I64, I64, VP],
    "bm_g2p": [VP, VP, VP, VP, VP, VP, F32, F32, I64, I64, VP],
    "bm_probe_stress": [VP, F32, F32, VP, VP, I32, VP],
    "bm_blade_constraint": [P, S, VP],
    "bm_store_residual": [P, S, VP, I32, VP],
    "bm_mls_p2g": [MP, MS, VP],
    "bm_mls_grid_update": [MP, MS, VP],
    "bm_mls_g2p": [MP, MS, VP],
    "bm_mls_substep": [P, S, MP, MS, VP],
    "bm_piece_labels": [P, S, ctypes.POINTER(Labels), VP],
    "bm_piece_labels_static": [P, S, I64, VP, VP, VP, VP, VP],
    "bm_load_external_batch": [P, S, VP, VP],
    "bm_store_interval_batch": [P, S, VP, VP],
    "bm_mlp_sizes": [I32, I32, I32, c_i64p, c_i64p, c_i64p],
    "bm_mlp_forward": [ctypes.POINTER(Mlp), VP, I32, VP],
    "bm_mlp_backward": [ctypes.POINTER(Mlp), VP, VP, I32, VP],
    "bm_mlp_train_step": [ctypes.POINTER(Mlp), VP, VP, I32, VP],
    "bm_mlp_pack": [ctypes.POINTER(Mlp), VP],
    "bm_mlp_infer": [ctypes.POINTER(Mlp), VP, VP, I32, VP],
    "bm_mlp_gemm": [VP, VP, VP, I32, I32, I32, I32, I32, I32, I32, I32, VP],
    "bm_mlp_gemm_split": [VP, VP, VP, I32, I32, I32, I32, I32, I32, I32, VP, I64, VP],
    "bm_blade_patch": [P, S, VP, VP],
}

_lib = None


def load():
    """Load the library once and declare every entry point; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise RuntimeError(f"HIP library not built: {LIBRARY_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIBRARY_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.bm_piece_labels_static_bytes.argtypes = [I64, I64, I64]
    lib.bm_piece_labels_static_bytes.restype = ctypes.c_int64
    _lib = lib
    return lib


def check(code, name):
    if code < 0:
        raise ValueError(f"{name}: invalid argument ({code})")
    if code > 0:
        raise RuntimeError(f"{name}: HIP launch failed with error {code}")
