"""Oracle: robust (intrinsic Delaunay) heat-method operators (test infrastructure only).

The reference's production engine calls potpourri3d's
``MeshHeatMethodDistanceSolver(V, F, t_coef, use_robust=True)`` (distance_engines.py:78-79,
119-125). potpourri3d is a third-party package absent here (requirements.txt:5-7 pins no
version); its robust mode follows the published algorithm restated below:

* intrinsic mollification (Sharp & Crane, "A Laplacian for Nonmanifold Triangle Meshes",
  SGP 2020, section 4.5): eps = factor * mean edge length, delta = max(0, max over triangles and
  corners of eps - (l_a + l_b - l_c)) added to every edge length (factor 1e-6);
* intrinsic Delaunay flipping (Fisher, Springborn, Bobenko, Schroder, "An algorithm for the
  construction of intrinsic Delaunay triangulations", 2006): an interior edge whose opposite
  angles sum to more than pi is replaced by the other diagonal of its two triangles, the new
  length taken from the planar layout of the pair, until none is left;
* the cotan Laplacian L_ij = -(cot a_ij + cot b_ij)/2, lumped mass sum(A_f)/3, the heat method
  on that triangulation (t = t_coef * mean edge length of the input surface), the divergence of
  the normalised gradient computed in each triangle's own planar layout.

This restatement shares no code with the library's (intrinsic.cpp uses halfedge arrays and
the Gram form; this uses a directed-edge map, cotangents and planar layouts).
Parity is pinned by invariants (every interior edge Delaunay, area preserved, no flips and the
extrinsic operators back on a Delaunay input, a planar known answer) in tests/test_oracle.py;
against potpourri3d itself it is unpinned (not installed).
"""

import math

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu

from .heat import mean_edge_length


def _cot(a, b, c):
    """cot of the angle opposite side c (law of cosines, area from Heron)."""
    s = sorted((a, b, c), reverse=True)
    q = (s[0] + (s[1] + s[2])) * (s[2] - (s[0] - s[1])) * (s[2] + (s[0] - s[1])) * (s[0] + (s[1] - s[2]))
    A = 0.25 * math.sqrt(max(q, 0.0))
    return (a * a + b * b - c * c) / (4.0 * A)


def intrinsic_delaunay(V, F, mollify=1e-6, max_sweeps=1000):
    """(faces, lengths opposite each corner, flips, delta) of the intrinsic Delaunay triangulation."""
    V = np.asarray(V, dtype=np.float64)
    faces = [list(map(int, f)) for f in np.asarray(F)]
    # directed edge (a, b) -> face holding it; edge lengths by unordered pair
    length = {}
    for f in faces:
        for c in range(3):
            a, b = f[c], f[(c + 1) % 3]
            key = (min(a, b), max(a, b))
            if key not in length:
                length[key] = float(np.linalg.norm(V[a] - V[b]))
    eps = mollify * (sum(length.values()) / len(length))
    delta = 0.0
    for f in faces:
        l = [length[(min(f[c], f[(c + 1) % 3]), max(f[c], f[(c + 1) % 3]))] for c in range(3)]
        for c in range(3):
            delta = max(delta, eps - (l[c] + l[(c + 1) % 3] - l[(c + 2) % 3]))
    # per-face edge lengths (edges can become parallel after flips, so lengths live on faces)
    flen = []
    for f in faces:
        flen.append([length[(min(f[c], f[(c + 1) % 3]), max(f[c], f[(c + 1) % 3]))] + delta for c in range(3)])
    owner = {}
    for fi, f in enumerate(faces):
        for c in range(3):
            owner[(f[c], f[(c + 1) % 3])] = (fi, c)
    flips = 0
    for _ in range(max_sweeps):
        changed = False
        for fi in range(len(faces)):
            for c in range(3):
                f = faces[fi]
                i, j, k = f[c], f[(c + 1) % 3], f[(c + 2) % 3]
                other = owner.get((j, i))
                if other is None or i > j:
                    continue
                gi, gc = other
                g = faces[gi]
                l = g[(gc + 2) % 3]
                if l == k:
                    continue
                cij, cjk, cki = flen[fi][c], flen[fi][(c + 1) % 3], flen[fi][(c + 2) % 3]
                cil, clj = flen[gi][(gc + 1) % 3], flen[gi][(gc + 2) % 3]
                if _cot(cjk, cki, cij) + _cot(cil, clj, cij) >= -1e-12:
                    continue
                # planar layout: i at 0, j on +x, k above, l below; new edge k-l
                xk = (cij * cij + cki * cki - cjk * cjk) / (2 * cij)
                yk = math.sqrt(max(cki * cki - xk * xk, 0.0))
                xl = (cij * cij + cil * cil - clj * clj) / (2 * cij)
                yl = -math.sqrt(max(cil * cil - xl * xl, 0.0))
                lkl = math.hypot(xk - xl, yk - yl)
                for key in ((i, j), (j, k), (k, i), (j, i), (i, l), (l, j)):
                    owner.pop(key, None)
                faces[fi] = [k, i, l]
                flen[fi] = [cki, cil, lkl]
                faces[gi] = [l, j, k]
                flen[gi] = [clj, cjk, lkl]
                for fj in (fi, gi):
                    ff = faces[fj]
                    for cc in range(3):
                        owner[(ff[cc], ff[(cc + 1) % 3])] = (fj, cc)
                flips += 1
                changed = True
        if not changed:
            break
    else:
        raise RuntimeError("intrinsic Delaunay flips did not converge")
    Fo = np.array(faces, dtype=np.int64)
    # length opposite corner c = edge (c+1, c+2)
    Lo = np.array([[fl[1], fl[2], fl[0]] for fl in flen])
    return Fo, Lo, flips, delta


def is_delaunay(F, Lopp, tol=1e-10):
    """Every interior edge has cot alpha + cot beta >= -tol."""
    F = np.asarray(F)
    cots = {}
    for fi, f in enumerate(F):
        for c in range(3):
            a, b = int(f[(c + 1) % 3]), int(f[(c + 2) % 3])  # edge opposite corner c
            lc = Lopp[fi][c]
            la, lb = Lopp[fi][(c + 1) % 3], Lopp[fi][(c + 2) % 3]
            cots.setdefault((min(a, b), max(a, b), round(lc, 12)), []).append(_cot(la, lb, lc))
    return all(sum(v) >= -tol for v in cots.values() if len(v) == 2)


def intrinsic_operators(F, Lopp, n):
    """(L, M, area): cotan Laplacian and lumped mass of the triangulation with the given lengths."""
    F = np.asarray(F, dtype=np.int64)
    rows, cols, vals = [], [], []
    mass = np.zeros(n)
    area = np.empty(F.shape[0])
    for fi, f in enumerate(F):
        l0, l1, l2 = Lopp[fi]
        s = sorted((l0, l1, l2), reverse=True)
        q = (s[0] + (s[1] + s[2])) * (s[2] - (s[0] - s[1])) * (s[2] + (s[0] - s[1])) * (s[0] + (s[1] - s[2]))
        A = 0.25 * math.sqrt(max(q, 0.0))
        area[fi] = A
        for c in range(3):
            a, b = int(f[(c + 1) % 3]), int(f[(c + 2) % 3])
            w = 0.5 * _cot(Lopp[fi][(c + 1) % 3], Lopp[fi][(c + 2) % 3], Lopp[fi][c])  # angle at corner c
            rows += [a, b, a, b]
            cols += [b, a, a, b]
            vals += [-w, -w, w, w]
            mass[int(f[c])] += A / 3.0
    L = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    L.sum_duplicates()
    return L, sp.diags(mass, format="csc"), area


def layout_divergence(u, F, Lopp):
    """sum_f A_f (X_f . grad B_i), X_f = -grad u/|grad u|, each face laid out in the plane."""
    u = np.asarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    F = np.asarray(F, dtype=np.int64)
    out = np.zeros_like(u)
    for fi, f in enumerate(F):
        l0, l1, l2 = Lopp[fi]
        # v0 at the origin, v1 on +x (|v0 v1| = l2), v2 from |v0 v2| = l1, |v1 v2| = l0
        x2 = (l2 * l2 + l1 * l1 - l0 * l0) / (2 * l2)
        y2 = math.sqrt(max(l1 * l1 - x2 * x2, 0.0))
        P = np.array([[0.0, 0.0], [l2, 0.0], [x2, y2]])
        A = 0.5 * l2 * y2
        # grad B_c = rot90(edge opposite c) / (2A), counter-clockwise layout
        g = np.empty((3, 2))
        for c in range(3):
            e = P[(c + 2) % 3] - P[(c + 1) % 3]
            g[c] = np.array([-e[1], e[0]]) / (2 * A)
        grad = g.T @ u[f]  # 2 x k
        nn = np.sqrt(np.sum(grad * grad, axis=0))
        with np.errstate(divide="ignore", invalid="ignore"):
            X = np.where(nn > 1e-300, -grad / nn, 0.0)
        out[f] += A * (g @ X)
    return out


class RobustHeatOracle:
    """FP64 heat-method distances on the intrinsic Delaunay triangulation (k sources per pass)."""

    def __init__(self, vertices, faces, t_coef=1.0, poisson_shift=1e-10, t=None, mollify=1e-6):
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        self.F, self.Lopp, self.flips, self.delta = intrinsic_delaunay(V, faces, mollify)
        n = V.shape[0]
        self.L, self.M, self.area = intrinsic_operators(self.F, self.Lopp, n)
        self.mean_edge = mean_edge_length(V, faces)
        self.t = float(t) if t is not None else float(t_coef) * self.mean_edge ** 2
        sym = dict(diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
        self._heat = splu(sp.csc_matrix(self.M + self.t * self.L), permc_spec="MMD_AT_PLUS_A", **sym)
        self._poisson = splu(sp.csc_matrix(self.L + poisson_shift * self.M), permc_spec="MMD_AT_PLUS_A", **sym)
        self.n = n

    def compute_distance_batch(self, sources):
        src = np.asarray(sources, dtype=np.int64).reshape(-1)
        rhs = np.zeros((self.n, src.size))
        rhs[src, np.arange(src.size)] = 1.0
        u = np.asarray(self._heat.solve(np.asfortranarray(rhs)))
        div = layout_divergence(u, self.F, self.Lopp)
        phi = np.asarray(self._poisson.solve(np.asfortranarray(div)))
        phi = phi - phi[src, np.arange(src.size)][None, :]
        return np.ascontiguousarray(phi.T)
