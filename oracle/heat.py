"""Oracle: FP64 heat-method distances (test infrastructure only).

Restates fastcw_lean.py:375-534 (``ConsistentBatchHeatEngine``):

* ``face_gradient_basis``  fastcw_lean.py:375-391
* ``consistent_operators`` fastcw_lean.py:394-413  (L = G^T diag(A) G, M = diag(sum A/3))
* ``fused_normalized_divergence`` fastcw_lean.py:416-442 (C loop in lean_geometry.c,
  plus an independent sparse-matrix form ``divergence_sparse`` used to pin it)
* ``HeatOracle``           fastcw_lean.py:467-534 (SuperLU factors of M + tL and
  L + shift*M, multi-RHS solves, shift phi -= phi[src])
* ``mean_edge_length``     distance_engines.py:17-21 / lean_geometry.py:360-367
"""

import time

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu

from . import _clib


def mean_edge_length(V, F):
    """Mean length over unique undirected edges (distance_engines.py:17-21)."""
    F = np.asarray(F, dtype=np.int64)
    E = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    E.sort(axis=1)
    E = np.unique(E, axis=0)
    V = np.asarray(V, dtype=np.float64)
    return float(np.mean(np.linalg.norm(V[E[:, 1]] - V[E[:, 0]], axis=1)))


def face_gradient_basis(V, F):
    """Per-face area and gradients of the three hat functions (fastcw_lean.py:375-391)."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F, dtype=np.int64)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    e01 = p1 - p0
    e02 = p2 - p0
    nrm = np.cross(e01, e02)
    nl = np.linalg.norm(nrm, axis=1)
    if np.any(nl <= 0.0):
        raise ValueError("Zero-area faces must be removed before building the heat engine.")
    area = 0.5 * nl
    nhat = nrm / nl[:, None]
    inv2a = 1.0 / (2.0 * area)
    gb1 = np.cross(e02, nhat) * inv2a[:, None]
    gb2 = np.cross(nhat, e01) * inv2a[:, None]
    gb0 = -gb1 - gb2
    return area, np.ascontiguousarray(gb0), np.ascontiguousarray(gb1), np.ascontiguousarray(gb2)


def consistent_operators(V, F):
    """(L, M, area, gb0, gb1, gb2) with L = sum_f A_f gradB_i . gradB_j and M lumped (fastcw_lean.py:394-413)."""
    F = np.asarray(F, dtype=np.int64)
    nV = int(np.asarray(V).shape[0])
    area, gb0, gb1, gb2 = face_gradient_basis(V, F)
    gb = (gb0, gb1, gb2)
    rows, cols, vals = [], [], []
    for i in range(3):
        for j in range(3):
            rows.append(F[:, i])
            cols.append(F[:, j])
            vals.append(area * np.einsum("ij,ij->i", gb[i], gb[j]))
    L = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nV, nV))
    L.sum_duplicates()
    mass = np.zeros(nV)
    for i in range(3):
        np.add.at(mass, F[:, i], area / 3.0)
    return L, sp.diags(mass, format="csc"), area, gb0, gb1, gb2


def fused_normalized_divergence(u, F, area, gb0, gb1, gb2):
    """out (n,k) = sum_f A_f (X_f . gradB_i), X_f = -grad u/|grad u| (fastcw_lean.py:416-442), C loop."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    n, k = u.shape
    F32 = np.ascontiguousarray(F, dtype=np.int32)
    out = np.empty((n, k), dtype=np.float64)
    _clib.lib().oc_fused_normalized_divergence(
        _clib.dptr(u), n, k, _clib.iptr(F32), F32.shape[0], _clib.dptr(np.ascontiguousarray(area)),
        _clib.dptr(np.ascontiguousarray(gb0)), _clib.dptr(np.ascontiguousarray(gb1)),
        _clib.dptr(np.ascontiguousarray(gb2)), _clib.dptr(out))
    return out


def gradient_operator(V, F):
    """Sparse G (3m x n): rows 3f..3f+2 hold grad of the P1 interpolant on face f (component-major per face)."""
    F = np.asarray(F, dtype=np.int64)
    area, gb0, gb1, gb2 = face_gradient_basis(V, F)
    m = F.shape[0]
    n = int(np.asarray(V).shape[0])
    rows, cols, vals = [], [], []
    for corner, gb in enumerate((gb0, gb1, gb2)):
        for c in range(3):
            rows.append(3 * np.arange(m) + c)
            cols.append(F[:, corner])
            vals.append(gb[:, c])
    G = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * m, n))
    return G, area


def divergence_sparse(u, V, F):
    """Independent form of the fused divergence: G^T diag(A) X with X = -Gu/|Gu| (zero where |Gu| <= 1e-300)."""
    G, area = gradient_operator(V, F)
    u = np.asarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    g = (G @ u).reshape(-1, 3, u.shape[1])
    nn = np.sqrt(np.sum(g * g, axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        X = np.where((nn > 1e-300)[:, None, :], -g / nn[:, None, :], 0.0)
    AX = (X * area[:, None, None]).reshape(-1, u.shape[1])
    return np.asarray(G.T @ AX)


class HeatOracle:
    """FP64 heat-method distances, k sources per factor pass (fastcw_lean.py:467-534).

    t = t_coef * mean_edge^2 unless ``t`` is given; Poisson matrix L + shift*M.
    SuperLU with MMD on A^T+A is the reference's own CHOLMOD stand-in
    (fastcw_lean.py:445-454). Both matrices are symmetric positive definite, so
    the factorisation runs in SuperLU's symmetric mode (diagonal pivots,
    diag_pivot_thresh=0): the same solution up to rounding as the reference's
    default partial pivoting, which on the near-singular Poisson matrix spends
    minutes per factorisation at 40k+ vertices.
    """

    def __init__(self, vertices, faces, t_coef=1.0, poisson_shift=1e-10, t=None, permc_spec="MMD_AT_PLUS_A"):
        V = np.ascontiguousarray(vertices, dtype=np.float64)
        F = np.ascontiguousarray(faces, dtype=np.int32)
        t0 = time.perf_counter()
        L, M, area, gb0, gb1, gb2 = consistent_operators(V, F)
        self.mean_edge = mean_edge_length(V, F)
        self.t = float(t) if t is not None else float(t_coef) * self.mean_edge ** 2
        self.L, self.M = L, M
        self.poisson_shift = float(poisson_shift)
        sym = dict(diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
        self._heat = splu(sp.csc_matrix(M + self.t * L), permc_spec=permc_spec, **sym)
        self._poisson = splu(sp.csc_matrix(L + self.poisson_shift * M), permc_spec=permc_spec, **sym)
        self.F, self.area, self.gb0, self.gb1, self.gb2 = F, area, gb0, gb1, gb2
        self.n_vertices = V.shape[0]
        self.setup_seconds = time.perf_counter() - t0

    def heat_solve(self, sources):
        src = np.asarray(sources, dtype=np.int64).reshape(-1)
        rhs = np.zeros((self.n_vertices, src.size))
        rhs[src, np.arange(src.size)] = 1.0
        return np.ascontiguousarray(self._heat.solve(np.asfortranarray(rhs)))

    def compute_distance_batch(self, sources, timings=None):
        """(k, n) C-contiguous; row j = distances from sources[j], shifted to 0 at the source."""
        src = np.asarray(sources, dtype=np.int64).reshape(-1)
        k = src.size
        t0 = time.perf_counter()
        u = self.heat_solve(src)
        t1 = time.perf_counter()
        div = fused_normalized_divergence(u, self.F, self.area, self.gb0, self.gb1, self.gb2)
        t2 = time.perf_counter()
        phi = np.asarray(self._poisson.solve(np.asfortranarray(div)))
        phi = phi - phi[src, np.arange(k)][None, :]
        out = np.ascontiguousarray(phi.T)
        t3 = time.perf_counter()
        if timings is not None:
            for key, dt in (("heat_solve", t1 - t0), ("fused_div", t2 - t1), ("poisson_solve", t3 - t2)):
                timings[key] = timings.get(key, 0.0) + dt
        return out

    def compute_distance(self, source_idx):
        return self.compute_distance_batch([int(source_idx)])[0]
