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
  the normalised gradient computed in each triangle's own planar layout;
* for meshes with non-manifold edges or inconsistently oriented neighbours, the tufted cover
  (Sharp & Crane 2020, section 4.2): each face taken once per side, the sides glued around every
  edge in the cyclic order of the faces about it, the flips run on that cover, and every cover
  face weighted 1/2 (each input face is counted twice); a face and its mirror copy that end up
  identical are merged into one face of weight 1.

This restatement shares no code with the library's (intrinsic.cpp uses halfedge arrays and
the Gram form; this uses a directed-edge map, or for the cover a (face, slot) twin map,
cotangents and planar layouts).
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


def _is_edge_manifold(F):
    seen = set()
    for f in np.asarray(F).tolist():
        for c in range(3):
            d = (f[c], f[(c + 1) % 3])
            if d in seen:
                return False
            seen.add(d)
    return True


def tufted_cover(V, F):
    """Cover faces (face f side 0 = f's vertex order, side 1 = reversed as (c, b, a); cover face
    index 2f + side) and the twin map {(cover face, slot): (cover face, slot)}, slot c being the
    edge from vertex c to vertex c + 1, glued around every input edge in the faces' angular order."""
    V = np.asarray(V, dtype=np.float64)
    F = [list(map(int, f)) for f in np.asarray(F)]
    faces = []
    for a, b, c in F:
        faces.append([a, b, c])
        faces.append([c, b, a])
    about = {}
    for fi, f in enumerate(F):
        for c in range(3):
            a, b = f[c], f[(c + 1) % 3]
            about.setdefault((min(a, b), max(a, b)), []).append(fi)

    def slot(cf, a, b):
        t = faces[cf]
        for c in range(3):
            if t[c] == a and t[(c + 1) % 3] == b:
                return c
        return None

    twin = {}
    nonmanifold = 0
    for (i, j), fl in about.items():
        nonmanifold += len(fl) >= 3
        e = V[j] - V[i]
        e = e / np.linalg.norm(e)
        ws, ups = [], []
        for fi in fl:
            k = [v for v in F[fi] if v != i and v != j][0]
            w = V[k] - V[i]
            w = w - np.dot(w, e) * e
            ws.append(w)
            p0, p1, p2 = V[F[fi][0]], V[F[fi][1]], V[F[fi][2]]
            ups.append(float(np.dot(np.cross(p1 - p0, p2 - p0), np.cross(e, w))) > 0.0)
        u = ws[0] / np.linalg.norm(ws[0])
        v = np.cross(e, u)
        theta = [math.atan2(float(np.dot(w, v)), float(np.dot(w, u))) for w in ws]
        ring = sorted(range(len(fl)), key=lambda q: (theta[q], fl[q]))
        for r in range(len(ring)):
            qa, qb = ring[r], ring[(r + 1) % len(ring)]
            ca = 2 * fl[qa] + (0 if ups[qa] else 1)  # side of face a facing increasing angle
            cb = 2 * fl[qb] + (1 if ups[qb] else 0)  # side of face b facing decreasing angle
            sa = slot(ca, i, j)
            x, y = (i, j) if sa is not None else (j, i)
            if sa is None:
                sa = slot(ca, j, i)
            sb = slot(cb, y, x)
            if sb is None or (ca, sa) in twin or (cb, sb) in twin:
                raise RuntimeError("tufted cover: inconsistent gluing")
            twin[(ca, sa)] = (cb, sb)
            twin[(cb, sb)] = (ca, sa)
    return faces, twin, nonmanifold


def tufted_delaunay(V, F, mollify=1e-6, max_sweeps=1000):
    """Intrinsic Delaunay triangulation of the tufted cover: (faces, lengths opposite each corner,
    weight per face, flips, delta, non-manifold edge count)."""
    V = np.asarray(V, dtype=np.float64)
    faces, twin, nonmanifold = tufted_cover(V, F)
    Fl = [list(map(int, f)) for f in np.asarray(F)]
    length = {}
    for f in Fl:
        for c in range(3):
            a, b = f[c], f[(c + 1) % 3]
            length.setdefault((min(a, b), max(a, b)), float(np.linalg.norm(V[a] - V[b])))
    eps = mollify * (sum(length.values()) / len(length))
    delta = 0.0
    for f in Fl:
        l = [length[(min(f[c], f[(c + 1) % 3]), max(f[c], f[(c + 1) % 3]))] for c in range(3)]
        for c in range(3):
            delta = max(delta, eps - (l[c] + l[(c + 1) % 3] - l[(c + 2) % 3]))
    # slot lengths per cover face
    flen = [[length[(min(t[c], t[(c + 1) % 3]), max(t[c], t[(c + 1) % 3]))] + delta for c in range(3)] for t in faces]
    flips = 0
    for _ in range(max_sweeps):
        changed = False
        for fi in range(len(faces)):
            for c in range(3):
                if (fi, c) not in twin:
                    continue
                gi, d = twin[(fi, c)]
                if (gi, d) < (fi, c):
                    continue
                f, g = faces[fi], faces[gi]
                i, j, k = f[c], f[(c + 1) % 3], f[(c + 2) % 3]
                l = g[(d + 2) % 3]
                if k == l:
                    continue
                cij, cjk, cki = flen[fi][c], flen[fi][(c + 1) % 3], flen[fi][(c + 2) % 3]
                cil, clj = flen[gi][(d + 1) % 3], flen[gi][(d + 2) % 3]
                if _cot(cjk, cki, cij) + _cot(cil, clj, cij) >= -1e-12:
                    continue
                xk = (cij * cij + cki * cki - cjk * cjk) / (2 * cij)
                yk = math.sqrt(max(cki * cki - xk * xk, 0.0))
                xl = (cij * cij + cil * cil - clj * clj) / (2 * cij)
                yl = -math.sqrt(max(cil * cil - xl * xl, 0.0))
                lkl = math.hypot(xk - xl, yk - yl)
                out_ki = twin.get((fi, (c + 2) % 3))
                out_jk = twin.get((fi, (c + 1) % 3))
                out_il = twin.get((gi, (d + 1) % 3))
                out_lj = twin.get((gi, (d + 2) % 3))
                for key in ((fi, 0), (fi, 1), (fi, 2), (gi, 0), (gi, 1), (gi, 2)):
                    twin.pop(key, None)
                faces[fi] = [k, i, l]  # slots k->i, i->l, l->k
                flen[fi] = [cki, cil, lkl]
                faces[gi] = [l, j, k]  # slots l->j, j->k, k->l
                flen[gi] = [clj, cjk, lkl]
                pairs = [((fi, 0), out_ki), ((fi, 1), out_il), ((gi, 0), out_lj), ((gi, 1), out_jk),
                         ((fi, 2), (gi, 2))]
                for a, b in pairs:
                    if b is not None:
                        twin[a] = b
                        twin[b] = a
                flips += 1
                changed = True
        if not changed:
            break
    else:
        raise RuntimeError("intrinsic Delaunay flips did not converge")
    # merge a face with an identical mirror copy
    groups = {}
    for fi, t in enumerate(faces):
        groups.setdefault(tuple(sorted(t)), []).append(fi)
    Fo, Lo, W = [], [], []
    for key in sorted(groups):
        left = list(groups[key])
        while left:
            fa = left.pop(0)
            la = sorted(flen[fa])
            mate = None
            for fb in left:
                lb = sorted(flen[fb])
                if all(abs(x - y) <= 1e-12 * x for x, y in zip(la, lb)):
                    mate = fb
                    break
            if mate is not None:
                left.remove(mate)
            Fo.append(faces[fa])
            fl = flen[fa]
            Lo.append([fl[1], fl[2], fl[0]])  # opposite corner c = slot c + 1
            W.append(1.0 if mate is not None else 0.5)
    return np.array(Fo, dtype=np.int64), np.array(Lo), np.array(W), flips, delta, nonmanifold


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


def intrinsic_operators(F, Lopp, n, weight=None):
    """(L, M, area): cotan Laplacian and lumped mass of the triangulation with the given lengths;
    each face's contribution times its weight (tufted cover), area returned unweighted."""
    F = np.asarray(F, dtype=np.int64)
    wf = np.ones(F.shape[0]) if weight is None else np.asarray(weight, dtype=np.float64)
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
            w = 0.5 * wf[fi] * _cot(Lopp[fi][(c + 1) % 3], Lopp[fi][(c + 2) % 3], Lopp[fi][c])  # corner c
            rows += [a, b, a, b]
            cols += [b, a, a, b]
            vals += [-w, -w, w, w]
            mass[int(f[c])] += wf[fi] * A / 3.0
    L = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    L.sum_duplicates()
    return L, sp.diags(mass, format="csc"), area


def layout_divergence(u, F, Lopp, weight=None):
    """sum_f w_f A_f (X_f . grad B_i), X_f = -grad u/|grad u|, each face laid out in the plane."""
    u = np.asarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    F = np.asarray(F, dtype=np.int64)
    wf = np.ones(F.shape[0]) if weight is None else np.asarray(weight, dtype=np.float64)
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
        out[f] += wf[fi] * A * (g @ X)
    return out


class RobustHeatOracle:
    """FP64 heat-method distances on the intrinsic Delaunay triangulation (k sources per pass)."""

    def __init__(self, vertices, faces, t_coef=1.0, poisson_shift=1e-10, t=None, mollify=1e-6, cover=None):
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        self.cover = (not _is_edge_manifold(faces)) if cover is None else bool(cover)
        if self.cover:
            self.F, self.Lopp, self.weight, self.flips, self.delta, _ = tufted_delaunay(V, faces, mollify)
        else:
            self.F, self.Lopp, self.flips, self.delta = intrinsic_delaunay(V, faces, mollify)
            self.weight = np.ones(self.F.shape[0])
        n = V.shape[0]
        self.L, self.M, self.area = intrinsic_operators(self.F, self.Lopp, n, self.weight)
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
        div = layout_divergence(u, self.F, self.Lopp, self.weight)
        phi = np.asarray(self._poisson.solve(np.asfortranarray(div)))
        phi = phi - phi[src, np.arange(src.size)][None, :]
        return np.ascontiguousarray(phi.T)
