"""Once-per-mesh host preparation of the cortical submesh (numpy).

Counterparts of the reference's mesh set-up, run once per hemisphere before
the device pipeline:

* ``validate_mesh_inputs``   core_analysis.py:779-816
* ``build_cortex_submesh``   core_analysis.py:859-911
* ``boundary_indices``       core_analysis.py:651-665
* ``bfs_order_all``          core_analysis.py:1336-1347, 1384-1417 (sorted adjacency, BFS from 0,
                             then every remaining component)
* ``face_geometry``          core_analysis.py:913-978 (face areas, max edge, vertex areas)
* ``mean_edge_length``       distance_engines.py:17-21
* ``default_bin_width``      lean_geometry.py:99-114
* ``normalize_scales`` / ``scale_token`` / ``boundary_area_loss_fraction``
                             core_analysis.py:727-777
"""

from __future__ import annotations

import math
from collections import deque

import numpy as np


def validate_mesh_inputs(vertices, faces, cortex_mask):
    V = np.asarray(vertices, dtype=np.float64)
    F_raw = np.asarray(faces)
    M = np.asarray(cortex_mask)
    if V.ndim != 2 or V.shape[1] != 3:
        raise ValueError(f"Invalid vertex array shape {V.shape}; expected (N, 3).")
    if F_raw.ndim != 2 or F_raw.shape[1] != 3:
        raise ValueError(f"Invalid face array shape {F_raw.shape}; expected (M, 3).")
    if F_raw.shape[0] == 0:
        raise ValueError("Face array is empty.")
    if M.ndim != 1:
        raise ValueError(f"Invalid cortex mask shape {M.shape}; expected (N,).")
    if M.shape[0] != V.shape[0]:
        raise ValueError(f"Mask length mismatch: mask has {M.shape[0]} entries, vertices has {V.shape[0]}.")
    if not np.issubdtype(F_raw.dtype, np.integer):
        if np.any(np.abs(F_raw - np.rint(F_raw)) > 0):
            raise ValueError("Face indices must be integers.")
        F = np.rint(F_raw).astype(np.int32)
    else:
        F = F_raw.astype(np.int32, copy=False)
    if np.any(F < 0):
        raise ValueError("Face indices contain negative values.")
    if np.any(F >= V.shape[0]):
        raise ValueError(f"Face indices out of range: max face index {int(np.max(F))}, "
                         f"but only {V.shape[0]} vertices provided.")
    M = M.astype(bool, copy=False)
    if not np.any(M):
        raise ValueError("Cortex mask is empty (no True vertices).")
    return V, F, M


def build_cortex_submesh(V, F, mask):
    """(V_sub, F_sub, sub_to_orig, orig_to_sub): all-cortex faces with area > 1e-12, reindexed."""
    keep = mask[F].all(axis=1)
    Fk = F[keep]
    if Fk.shape[0] == 0:
        return (np.empty((0, 3)), np.empty((0, 3), np.int32), np.empty(0, np.int32),
                np.full(V.shape[0], -1, np.int32))
    p0, p1, p2 = V[Fk[:, 0]], V[Fk[:, 1]], V[Fk[:, 2]]
    areas = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)
    Fk = Fk[areas > 1e-12]
    if Fk.shape[0] == 0:
        return (np.empty((0, 3)), np.empty((0, 3), np.int32), np.empty(0, np.int32),
                np.full(V.shape[0], -1, np.int32))
    used = np.unique(Fk)
    sub_to_orig = used.astype(np.int32)
    orig_to_sub = np.full(V.shape[0], -1, dtype=np.int32)
    orig_to_sub[sub_to_orig] = np.arange(sub_to_orig.size, dtype=np.int32)
    return (np.ascontiguousarray(V[sub_to_orig]), np.ascontiguousarray(orig_to_sub[Fk]), sub_to_orig,
            orig_to_sub)


def boundary_indices(F):
    """Sorted vertices on edges that belong to one face (core_analysis.py:651-665); edges keyed as
    one int64 (lo * n + hi) so the counting is a flat sort."""
    F = np.asarray(F, dtype=np.int64)
    if F.size == 0:
        return np.empty(0, dtype=np.int32)
    n = int(F.max()) + 1
    a = np.concatenate([F[:, 0], F[:, 1], F[:, 2]])
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0]])
    key = np.minimum(a, b) * n + np.maximum(a, b)
    uk, cnt = np.unique(key, return_counts=True)
    bk = uk[cnt == 1]
    if bk.size:
        return np.unique(np.concatenate([bk // n, bk % n])).astype(np.int32)
    return np.empty(0, dtype=np.int32)


def adjacency_csr(n, F):
    """Sorted unique neighbour lists as CSR (indptr, indices)."""
    F = np.asarray(F, dtype=np.int64)
    a = np.concatenate([F[:, 0], F[:, 0], F[:, 1], F[:, 1], F[:, 2], F[:, 2]])
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0], F[:, 2], F[:, 0], F[:, 1]])
    key = np.unique(a * n + b)
    rows = key // n
    cols = (key % n).astype(np.int32)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, rows + 1, 1)
    np.cumsum(indptr, out=indptr)
    return indptr, cols


def bfs_order_all(n, F, start=0):
    """Breadth-first order over every component (core_analysis.py:1384-1417): from ``start``,
    neighbours in ascending order, then from each unvisited vertex in index order. scipy's
    compiled breadth-first order visits a vertex's neighbours in CSR order, which is ascending
    here, so the order is the same as a queue over sorted neighbour lists."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import breadth_first_order

    if n == 0:
        return np.empty(0, dtype=np.int64)
    indptr, cols = adjacency_csr(n, F)
    G = sp.csr_matrix((np.ones(cols.size, dtype=np.int8), cols, indptr), shape=(n, n))
    G.has_sorted_indices = True
    visited = np.zeros(n, dtype=bool)
    parts = []
    s0 = start if 0 <= start < n else 0
    for s in [s0] + list(range(n)):
        if visited[s]:
            continue
        part = breadth_first_order(G, s, directed=True, return_predecessors=False)
        visited[part] = True
        parts.append(part)
        if visited.all():
            break
    return np.concatenate(parts).astype(np.int64)


def face_geometry(V, F):
    """(face_areas, face_L, vertex_areas) with the scale-aware zero-area threshold."""
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e01, e12, e20 = p1 - p0, p2 - p1, p0 - p2
    nn = np.linalg.norm(np.cross(e01, p2 - p0), axis=1)
    L = np.maximum.reduce([np.linalg.norm(e01, axis=1), np.linalg.norm(e12, axis=1),
                           np.linalg.norm(e20, axis=1)])
    tiny = np.finfo(np.float64).eps * (L * L) * 10.0
    areas = 0.5 * nn
    areas[nn < tiny] = 0.0
    va = np.zeros(V.shape[0])
    for i in range(3):
        np.add.at(va, F[:, i], areas / 3.0)
    return areas, L, va


def mean_edge_length(V, F):
    F = np.asarray(F, dtype=np.int64)
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    E.sort(axis=1)
    E = np.unique(E, axis=0)
    return float(np.mean(np.linalg.norm(V[E[:, 1]] - V[E[:, 0]], axis=1)))


def mean_edge_from_adjacency(V, indptr, cols):
    """mean_edge_length from a sorted adjacency CSR: the edges (a, b), a < b, in row order are the
    sorted unique edges, so the mean is taken over the same values in the same order."""
    indptr = np.asarray(indptr, dtype=np.int64)
    rows = np.repeat(np.arange(indptr.size - 1), np.diff(indptr))
    cols = np.asarray(cols, dtype=np.int64)
    up = cols > rows
    return float(np.mean(np.linalg.norm(V[cols[up]] - V[rows[up]], axis=1)))


def bin_width_from_lengths(L, quantile=0.999, pad=1.05):
    """default_bin_width from the faces' longest edges (face_geometry's L)."""
    return float(np.quantile(L, quantile) * pad)


def default_bin_width(V, F, quantile=0.999, pad=1.05):
    L = np.maximum.reduce([
        np.linalg.norm(V[F[:, 1]] - V[F[:, 0]], axis=1),
        np.linalg.norm(V[F[:, 2]] - V[F[:, 1]], axis=1),
        np.linalg.norm(V[F[:, 0]] - V[F[:, 2]], axis=1),
    ])
    return float(np.quantile(L, quantile) * pad)


def normalize_scales(scale, default=(0.0006, 0.001, 0.002, 0.004, 0.008, 0.016)):
    if scale is None:
        raw = list(default)
    elif np.isscalar(scale):
        raw = [scale]
    else:
        raw = list(scale)
    if not raw:
        raise ValueError("At least one scale must be provided.")
    out, seen = [], set()
    for item in raw:
        v = float(item)
        if not np.isfinite(v):
            raise ValueError(f"Invalid scale {item!r}: must be finite.")
        if v <= 0.0 or v > 1.0:
            raise ValueError(f"Invalid scale {item!r}: expected 0 < scale <= 1.")
        if v not in seen:
            seen.add(v)
            out.append(v)
    out.sort()
    return tuple(out)


def scale_token(scale):
    return format(float(scale), "g")


def boundary_area_loss_fraction(radius, boundary_distance):
    r = float(radius)
    d = float(boundary_distance)
    if not np.isfinite(r) or r <= 0.0:
        return np.nan
    if not np.isfinite(d):
        return 0.0
    if d >= r:
        return 0.0
    x = min(1.0, max(0.0, d / r))
    return float((math.acos(x) - x * math.sqrt(max(0.0, 1.0 - x * x))) / math.pi)
