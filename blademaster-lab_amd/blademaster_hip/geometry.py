"""Reference-configuration sampling and the trilinear TLMPM stencil.

Follows `blademaster_warp/geometry.py:13-90`: two particles per grid cell and axis at
sub-cell centres (seed 0) or jittered uniformly inside their sub-cell (seed > 0), a shape
filter by an analytic level set, and the trilinear stencil (8 nodes, weights and
reference gradients) evaluated once in float64 on the host.  The stencil never changes
in TLMPM, so the device kernels keep only a compact form of it (the base node and the
three local cell fractions) and rebuild the 8 weights and 24 gradient terms in registers.
"""

import numpy as np

# Domain and wedge of the DexLab banana demo (`geometry.py:6-10`).
BANANA_DEMO_SIZE_M = (0.12, 0.06, 0.03)
DEMO_KNIFE_HALF_SIZE_M = (0.0005, 0.04, 0.02)
DEMO_KNIFE_BEVEL_HEIGHT_M = 0.005

# Corner order of the 8 stencil nodes: bit 2 = x offset, bit 1 = y, bit 0 = z
# (itertools.product((0, 1), repeat=3) order, `geometry.py:81`).
CORNERS = np.array([[(a >> 2) & 1, (a >> 1) & 1, a & 1] for a in range(8)], dtype=np.int64)


def levelset(x, size, shape):
    """Signed analytic boundary, negative inside (`geometry.py:13-33`)."""
    x = np.asarray(x, dtype=np.float64)
    size = np.asarray(size, dtype=np.float64)
    if shape == "block":
        q = np.abs(x - np.array([0.0, 0.0, size[2] / 2])) - size / 2
        outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
        return outside + np.minimum(q.max(axis=-1), 0.0)
    if shape == "cylinder":
        radial = np.hypot(x[..., 0], x[..., 1]) - min(size[0], size[1]) / 2
        axial = np.abs(x[..., 2] - size[2] / 2) - size[2] / 2
        return np.maximum(radial, axial)
    if shape == "horizontal_cylinder":
        # Axis along x through (y = 0, z = r), lying on the board; r = min(sy, sz) / 2.  Not in
        # the reference, whose cylinder stands on its end: the three sequential cuts of a log
        # need the axis along x.
        r = min(size[1], size[2]) / 2
        radial = np.hypot(x[..., 1], x[..., 2] - r) - r
        return np.maximum(np.abs(x[..., 0]) - size[0] / 2, radial)
    if shape == "banana":
        u = x[..., 0] / (size[0] / 2)
        centre_y = 0.035 * (u * u - 0.5)
        radius = 0.003 + 0.009 * np.maximum(0.0, 1.0 - u * u) ** 0.55
        tube = np.sqrt((x[..., 1] - centre_y) ** 2 + (x[..., 2] - 0.012) ** 2) - radius
        return np.maximum(np.abs(x[..., 0]) - size[0] / 2, tube)
    raise ValueError(shape)


def check_domain(size, shape):
    """The banana tube must fit the sampling box (`geometry.py:36-44`)."""
    if shape != "banana":
        return
    u = np.linspace(-1.0, 1.0, 10001)
    centre_y = 0.035 * (u * u - 0.5)
    radius = 0.003 + 0.009 * np.maximum(0.0, 1.0 - u * u) ** 0.55
    reach_y = max(np.max(centre_y + radius), -np.min(centre_y - radius))
    if reach_y > size[1] / 2 + 1.0e-10 or 0.024 > size[2] + 1.0e-10:
        raise ValueError(
            "Banana analytic geometry is clipped by the sampling domain; use y size >= 0.059 m and z >= 0.024 m"
        )


def grid_cells(size, spacing):
    """Cells per axis covering the whole domain (`geometry.py:47-49`)."""
    return np.ceil(np.asarray(size, dtype=np.float64) / spacing - 1.0e-9).astype(int)


def points(size, spacing, shape, seed=0):
    """Reference particle positions, the grid origin and the cell counts (`geometry.py:52-74`).

    Particles come out in lexicographic lattice order (x slowest, z fastest).
    """
    size = np.asarray(size, dtype=np.float64)
    check_domain(size, shape)
    cells = grid_cells(size, spacing)
    origin = np.array([-size[0] / 2, -size[1] / 2, 0.0])
    axes = [origin[i] + (np.arange(2 * cells[i]) + 0.5) * spacing / 2 for i in range(3)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    x = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    if seed:
        x = x + np.random.default_rng(seed).uniform(-spacing / 4, spacing / 4, size=x.shape)
    if shape in ("banana", "cylinder", "horizontal_cylinder"):
        x = x[levelset(x, size, shape) <= 0]
    elif shape == "block":
        # Partial cells at the far faces when the spacing does not divide the size.
        x = x[np.all(x - origin <= size, axis=1)]
    else:
        raise ValueError(shape)
    if seed:
        # Reflect jittered bottom particles above the board contact distance h/4.
        x[:, 2] = spacing / 4 + np.abs(x[:, 2] - spacing / 4)
    return x, origin, cells


def stencil(reference, origin, spacing, cells):
    """Trilinear nodes (n, 8), weights (n, 8) and reference gradients (n, 8, 3), float64
    (`geometry.py:77-90`)."""
    frac = (np.asarray(reference, dtype=np.float64) - origin) / spacing
    base = np.floor(frac).astype(np.int64)
    local = frac - base
    # factor[p, a, axis] = local if the corner is offset along axis, else 1 - local.
    factor = np.where(CORNERS[None] == 1, local[:, None, :], 1.0 - local[:, None, :])
    weights = factor[:, :, 0] * factor[:, :, 1] * factor[:, :, 2]
    gradients = np.empty(weights.shape + (3,))
    for axis in range(3):
        b, c = [k for k in range(3) if k != axis]
        sign = (2 * CORNERS[:, axis] - 1)[None]
        gradients[:, :, axis] = sign * (factor[:, :, b] * factor[:, :, c]) / spacing
    ijk = base[:, None, :] + CORNERS[None]
    dims = np.asarray(cells) + 1
    nodes = (ijk[..., 0] * dims[1] + ijk[..., 1]) * dims[2] + ijk[..., 2]
    return nodes, weights, gradients


def compact_stencil(reference, origin, spacing, cells):
    """Device form of the stencil: base node index (int32) and local cell fractions (f32).

    Nodes of corner a are base + (dx * (dims_y * dims_z) + dy * dims_z + dz); weights and
    gradients are the trilinear products of the fractions.  Built from the same float64
    fractions as `stencil`.
    """
    frac = (np.asarray(reference, dtype=np.float64) - origin) / spacing
    base = np.floor(frac).astype(np.int64)
    local = frac - base
    dims = np.asarray(cells) + 1
    base_node = (base[:, 0] * dims[1] + base[:, 1]) * dims[2] + base[:, 2]
    return base_node.astype(np.int32), local.astype(np.float32), base.astype(np.int32)


# Cells per brick along x and y of the particle bins, and the default brick height along z
# (bricks are at most as tall as the kernels' LDS windows allow, and at most 1024 particle
# slots).
BRICK_XY = 2
# Measured on the 1M-particle block (MI355X): 4 cells per brick in z gives 120.5 us per
# substep, 16 gives 126.7 us on the same device, 2 and 3 are slower again.
BRICK_Z = 4
LDS_BYTES = 65536


def lattice_slots(local):
    """Cell slot 4 bx + 2 by + bz of each particle from its local cell coordinates when every
    particle sits on the regular lattice (1/4 or 3/4 of the cell per axis, exactly in
    float32), else None (jittered sampling)."""
    local = np.asarray(local, dtype=np.float32)
    if not np.all((local == np.float32(0.25)) | (local == np.float32(0.75))):
        return None
    bits = (local > 0.5).astype(np.int64)
    return 4 * bits[:, 0] + 2 * bits[:, 1] + bits[:, 2]


def bin_bricks(cells, base_cell, fields, brick_xy=BRICK_XY, brick_z=BRICK_Z, lattice=None):
    """Particle bins of the brick kernels.

    Returns (brick cells per axis, bricks per axis, slots): bricks are numbered x-major, the
    cells of a brick z-fastest, and the c-th cell in this order holds its particles in
    slots[8c:8c + 8], -1 for unused slots.  With `lattice` (from `lattice_slots`) a particle
    takes the slot of its lattice position; otherwise the cell's particles fill its first
    slots in reference order.  The sampling puts at most 8 particles in a cell (2 per axis,
    jitter stays inside the sub-cell).
    The brick kernels keep per (field slot, window node) 7 (P2G) or 3 (G2P) floats in LDS,
    plus 7 or 3 floats per (cell, corner); the larger must fit LDS_BYTES.
    """
    cells = np.asarray(cells, dtype=np.int64)

    def fits(bx, by, bz):
        window = fields * (bx + 1) * (by + 1) * (bz + 1)
        corners = 8 * bx * by * bz
        return max(7 * window + 7 * corners, 3 * window + 3 * corners) * 4 <= LDS_BYTES

    bx = by = brick_xy
    bz = int(max(1, min(cells[2], brick_z, 1024 // (8 * bx * by))))
    while bz > 1 and not fits(bx, by, bz):
        bz -= 1
    while not fits(bx, by, bz) and (bx > 1 or by > 1):
        bx, by = max(1, bx - 1), max(1, by - 1)
    if not fits(bx, by, bz):
        raise ValueError("fields_per_node too large for the LDS window of one brick")
    brick = np.array([bx, by, bz], dtype=np.int64)
    bricks = -(-cells // brick)
    base_cell = np.asarray(base_cell, dtype=np.int64)
    ids = ((base_cell[:, 0] // bx) * bricks[1] + base_cell[:, 1] // by) * bricks[2] + base_cell[:, 2] // bz
    local = ((base_cell[:, 0] % bx) * by + base_cell[:, 1] % by) * bz + base_cell[:, 2] % bz
    per_brick = int(bx * by * bz)
    key = ids * per_brick + local
    perm = np.argsort(key, kind="stable")
    nb = int(np.prod(bricks))
    sorted_key = key[perm]
    slots = np.full(8 * nb * per_brick, -1, dtype=np.int32)
    if lattice is not None:
        index = 8 * key + np.asarray(lattice, dtype=np.int64)
        if len(np.unique(index)) != len(index):
            raise ValueError("two particles on one lattice position")
        slots[index] = np.arange(len(key), dtype=np.int32)
        return brick, bricks, slots
    start = np.searchsorted(sorted_key, sorted_key)
    rank = np.arange(len(perm)) - start
    if len(rank) and rank.max() >= 8:
        raise ValueError("more than 8 particles in a grid cell")
    slots[8 * sorted_key + rank] = perm
    return brick, bricks, slots
