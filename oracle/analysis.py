"""Oracle: mesh preparation and the per-source wiring loop (test infrastructure only).

Restates:
* ``build_cortex_submesh``     core_analysis.py:859-911
* ``boundary_indices``         core_analysis.py:651-665
* ``bfs_order_all``            core_analysis.py:1336-1347, 1384-1417
* ``face_geometry``            core_analysis.py:913-957 (areas, face_L), 959-978 (vertex areas)
* ``clean_vertex_samples``     core_analysis.py:1424-1439
* ``wiring_costs``             core_analysis.py:1441-1632 (per source: distance -> field_health
  -> distance to boundary -> MSD -> per_source_geometry -> samples; CSR over the full vertex space)
"""

from collections import deque

import numpy as np

from . import geometry as og
from . import heat as oh


def build_cortex_submesh(V, F, mask):
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int32)
    mask = np.asarray(mask, dtype=bool)
    keep = mask[F].all(axis=1)
    Fk = F[keep]
    if Fk.shape[0] == 0:
        raise ValueError("No valid cortical faces remain after applying cortex mask.")
    p0, p1, p2 = V[Fk[:, 0]], V[Fk[:, 1]], V[Fk[:, 2]]
    areas = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1)
    Fk = Fk[areas > 1e-12]
    used = np.unique(Fk)
    sub_to_orig = used.astype(np.int32)
    orig_to_sub = np.full(V.shape[0], -1, dtype=np.int32)
    orig_to_sub[sub_to_orig] = np.arange(sub_to_orig.size, dtype=np.int32)
    return (np.ascontiguousarray(V[sub_to_orig]), np.ascontiguousarray(orig_to_sub[Fk]),
            sub_to_orig, orig_to_sub)


def boundary_indices(F):
    F = np.asarray(F, dtype=np.int32)
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    e.sort(axis=1)
    ue, cnt = np.unique(e, axis=0, return_counts=True)
    be = ue[cnt == 1]
    if be.size:
        return np.unique(be.reshape(-1)).astype(np.int32)
    return np.empty(0, dtype=np.int32)


def bfs_order_all(n, F, start=0):
    adj = [set() for _ in range(n)]
    for a, b, c in np.asarray(F, dtype=np.int64).tolist():
        adj[a].update((b, c))
        adj[b].update((a, c))
        adj[c].update((a, b))
    adj = [sorted(s) for s in adj]
    visited = np.zeros(n, dtype=bool)
    order = []

    def bfs_from(s):
        q = deque([s])
        visited[s] = True
        while q:
            v = q.popleft()
            order.append(v)
            for nb in adj[v]:
                if not visited[nb]:
                    visited[nb] = True
                    q.append(nb)

    bfs_from(start if 0 <= start < n else 0)
    for v in range(n):
        if not visited[v]:
            bfs_from(v)
    return order


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


def clean_vertex_samples(samples):
    if not samples:
        return []
    arr = np.asarray(samples, dtype=np.float64)
    arr = arr[np.argsort(arr[:, 0], kind="stable")]
    if arr.shape[0] == 1:
        return [(float(arr[0, 0]), float(arr[0, 1]))]
    r = arr[:, 0]
    tol = max(1e-12, 1e-9 * float(r[-1] - r[0]))
    keep = np.empty(arr.shape[0], dtype=bool)
    keep[0] = True
    keep[1:] = np.diff(r) > tol
    arr = arr[keep]
    return [(float(a0), float(a1)) for a0, a1 in arr]


class OracleMesh:
    """Cortical submesh plus everything the per-source loop precomputes (core_analysis.py:545-724)."""

    def __init__(self, V, F, mask, eps=1e-6):
        self.V_full = np.asarray(V, dtype=np.float64)
        self.n_full = self.V_full.shape[0]
        self.mask_full = np.asarray(mask, dtype=bool)
        self.V, self.F, self.sub_to_orig, self.orig_to_sub = build_cortex_submesh(V, F, mask)
        self.n = self.V.shape[0]
        self.boundary = boundary_indices(self.F)
        self.face_areas, self.face_L, self.vertex_areas = face_geometry(self.V, self.F)
        self.lean_area, self.g11, self.g12, self.g22 = og.precompute_face_metric(self.V, self.F)
        self.width = og.default_bin_width(self.V, self.F)
        self.health_h = float(np.median(self.face_L))
        self.eps = float(eps)
        self._bfs = None

    @property
    def bfs_order(self):
        if self._bfs is None:
            self._bfs = bfs_order_all(self.n, self.F)
        return self._bfs


def wiring_costs(mesh, distance_batch, scales, n_samples_between_scales=3, boundary_cap_fraction=None,
                 sources_sub=None, block=64):
    """The reference per-source loop (core_analysis.py:1441-1632) over ``sources_sub`` (default: BFS order).

    ``distance_batch(list_of_sub_idx) -> (k, n)`` supplies the fields (row per source).
    Returns a dict of float32 full-vertex-space arrays plus field_flag and the sample CSR.
    """
    scales = tuple(sorted(float(s) for s in scales))
    nF = mesh.n_full
    res = {
        "msd_unweighted": np.full(nF, np.nan, np.float32),
        "msd_weighted": np.full(nF, np.nan, np.float32),
        "dist_to_boundary": np.full(nF, np.nan, np.float32),
        "radius": np.full((len(scales), nF), np.nan, np.float32),
        "perimeter": np.full((len(scales), nF), np.nan, np.float32),
        "field_flag": np.zeros(nF, np.uint8),
        "neg_frac": np.full(nF, np.nan),
        "viol_frac": np.full(nF, np.nan),
    }
    total_area = float(np.sum(mesh.vertex_areas))
    targets = np.array([total_area * s for s in scales], dtype=np.float64)
    m = int(n_samples_between_scales)
    fracs = np.arange(1, m + 1, dtype=np.float64) / (m + 1.0)
    order = list(mesh.bfs_order) if sources_sub is None else [int(s) for s in sources_sub]
    samples = [None] * nF
    n_eval = 0
    for b0 in range(0, len(order), block):
        blk = order[b0:b0 + block]
        D = np.atleast_2d(distance_batch(blk))
        good_rows, good_src = [], []
        for row, sub_idx in enumerate(blk):
            d = np.ascontiguousarray(D[row], dtype=np.float64)
            oi = int(mesh.sub_to_orig[sub_idx])
            neg, viol, bad = og.field_health(d, mesh.V, sub_idx, mesh.health_h)
            res["neg_frac"][oi], res["viol_frac"][oi] = neg, viol
            if bad:
                res["field_flag"][oi] = 1
                continue
            if mesh.boundary.size:
                bd = d[mesh.boundary]
                bd = bd[np.isfinite(bd)]
                res["dist_to_boundary"][oi] = np.float32(np.min(bd)) if bd.size else np.nan
            else:
                res["dist_to_boundary"][oi] = np.float32(np.inf)
            valid = (d > mesh.eps) & np.isfinite(d)
            if np.any(valid):
                dv = d[valid]
                w = mesh.vertex_areas[valid]
                wsum = float(np.sum(w))
                res["msd_unweighted"][oi] = np.float32(np.mean(dv))
                res["msd_weighted"][oi] = np.float32(float((dv * w).sum() / wsum)
                                                     if np.isfinite(wsum) and wsum > 0.0 else np.nan)
            good_rows.append(row)
            good_src.append(sub_idx)
        if not good_rows:
            continue
        geo = og.per_source_geometry(D[good_rows], mesh.F, mesh.lean_area, mesh.g11, mesh.g12, mesh.g22,
                                     mesh.width, targets, fracs, 1e-10, 60)
        n_eval += int(geo["n_eval"].sum())
        for i, sub_idx in enumerate(good_src):
            oi = int(mesh.sub_to_orig[sub_idx])
            ro, po = geo["radii"][i], geo["perim"][i]
            res["radius"][:, oi] = ro.astype(np.float32)
            res["perimeter"][:, oi] = po.astype(np.float32)
            pairs = [(float(r), float(a)) for r, a in zip(ro, targets) if np.isfinite(r)]
            er, ea = geo["extra_r"][i], geo["extra_a"][i]
            if er.size:
                ok = np.isfinite(er) & np.isfinite(ea)
                if boundary_cap_fraction is not None:
                    d_b = float(res["dist_to_boundary"][oi])
                    for j in np.flatnonzero(ok):
                        if og.boundary_area_loss_fraction(float(er[j]), d_b) > boundary_cap_fraction:
                            ok[j] = False
                pairs.extend(zip(er[ok].tolist(), ea[ok].tolist()))
            samples[oi] = clean_vertex_samples(pairs)
    indptr = np.zeros(nF + 1, dtype=np.int64)
    for i, row in enumerate(samples):
        indptr[i + 1] = indptr[i] + (0 if row is None else len(row))
    sr = np.empty(int(indptr[-1]), np.float32)
    sa = np.empty(int(indptr[-1]), np.float32)
    for i, row in enumerate(samples):
        if row:
            sr[indptr[i]:indptr[i + 1]] = np.asarray([x[0] for x in row], np.float32)
            sa[indptr[i]:indptr[i + 1]] = np.asarray([x[1] for x in row], np.float32)
    res.update(sample_indptr=indptr, sample_radii=sr, sample_areas=sa, targets=targets, scales=scales,
               n_eval=n_eval)
    return res


def heat_oracle_for(mesh, **kw):
    return oh.HeatOracle(mesh.V, mesh.F, **kw)
