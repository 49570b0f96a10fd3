"""Piece labels: connected components of particles through compatible grid fields.

Restates `blademaster_warp/topology.py:6-57`.  Graph vertices are the particles plus one
vertex per distinct (grid node, cut code) field; a particle joins the fields of its
stencil nodes with positive weight, and two fields of the same node join when their codes
are compatible.  Labels are ranked by the lowest particle index of each piece, so the
output does not depend on the traversal order.
"""

import numpy as np


def compatible(a, b):
    """Ternary cut codes agree on every digit where both are non-zero (`topology.py:6-14`)."""
    a, b = int(a), int(b)
    while a or b:
        da, db = a % 3, b % 3
        if da and db and da != db:
            return False
        a //= 3
        b //= 3
    return True


def side_masks(codes):
    """Bit masks of the digits equal to 1 and to 2 of non-negative int64 ternary codes."""
    codes = np.asarray(codes, dtype=np.int64).copy()
    one = np.zeros(codes.shape, dtype=np.uint64)
    two = np.zeros(codes.shape, dtype=np.uint64)
    bit = np.uint64(1)
    while np.any(codes):
        digit = codes % 3
        one |= np.where(digit == 1, bit, np.uint64(0))
        two |= np.where(digit == 2, bit, np.uint64(0))
        codes //= 3
        bit = bit << np.uint64(1)
    return one, two


def components(nodes, weights, codes):
    """Labels (n,) ordered by lowest particle index, and the piece sizes."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    nodes = np.asarray(nodes)
    weights = np.asarray(weights)
    codes = np.asarray(codes, dtype=np.int64)
    n = len(codes)
    if n == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
    used = weights > 0
    particle = np.nonzero(used)[0]
    field_node = nodes[used].astype(np.int64)
    field_code = codes[particle]
    # Distinct (node, code) fields.
    pairs = np.stack([field_node, field_code], axis=1)
    unique, inverse = np.unique(pairs, axis=0, return_inverse=True)
    inverse = inverse.ravel()
    rows = [particle]
    cols = [n + inverse]
    # Fields sharing a node: link the compatible ones.  Fields are sorted by node, so the
    # fields of one node form a run.
    one, two = side_masks(unique[:, 1])
    starts = np.flatnonzero(np.r_[True, unique[1:, 0] != unique[:-1, 0]])
    ends = np.r_[starts[1:], len(unique)]
    multi = np.flatnonzero(ends - starts > 1)
    for i in multi:
        lo, hi = starts[i], ends[i]
        idx = np.arange(lo, hi)
        a, b = np.triu_indices(hi - lo, 1)
        clash = (one[idx[a]] & two[idx[b]]) | (two[idx[a]] & one[idx[b]])
        keep = clash == 0
        rows.append(n + idx[a][keep])
        cols.append(n + idx[b][keep])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    size = n + len(unique)
    graph = coo_matrix((np.ones(len(r), dtype=np.int8), (r, c)), shape=(size, size))
    _, label = connected_components(graph, directed=False)
    label = label[:n]
    first = np.full(label.max() + 1, n, dtype=np.int64)
    np.minimum.at(first, label, np.arange(n))
    present = first < n
    order = np.argsort(first[present], kind="stable")
    rank = np.full(len(first), -1, dtype=np.int64)
    rank[np.flatnonzero(present)[order]] = np.arange(order.size)
    labels = rank[label]
    return labels, np.bincount(labels)
